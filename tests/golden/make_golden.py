#!/usr/bin/env python3
"""Write the golden vectors in tests/golden/ from the reference routines.

Runs oracle/_ref/ref_driver -- the reference's own sipsol.f90, cgstab_sip.f90,
GRAD_GAUSS.f90, bpres.f90, fluxmass2.f90, PISO_assemble_pressure_eq.f90,
assemble_pressure_eq.f90, fluxmc.f90 and continuityErrors.h compiled with
amdflang by `make -C oracle ref`
(build container only: the reference sources are not on the GPU machines) --
on seeded inputs, and stores inputs' seeds and the reference outputs as
small .npz fixtures.  Inputs are regenerated by the tests from the same seeds
(oracle builders + numpy PCG64), so only outputs are stored.

usage: python tests/golden/make_golden.py
"""
import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import oracle as O  # noqa: E402

B_FIELDS = ("bb", "bs", "bw", "bp", "bn", "be", "bt")
SOLVER_OUT = ("fi", "res") + B_FIELDS


def solver_inputs(kind, dims, seed):
    """Seeded inputs: the system, fi, and random halo data in res / coefb."""
    if kind == "general":
        s, fi = O.build_general(dims, seed)
    else:
        s = O.build_lognormal(dims, seed, 1.0)
        fi = np.zeros(s.n)
    rng = np.random.default_rng(seed)
    s.res[:] = rng.uniform(-1, 1, s.n)
    for f in B_FIELDS:
        getattr(s, f)[:] = rng.uniform(-0.2, 0.2, s.n)
    return s, fi


def geometry(dims, seed):
    rng = np.random.default_rng(seed)
    n = dims[0] * dims[1] * dims[2]
    g = {f: rng.uniform(0.2, 0.8, n) for f in ("fx", "fy", "fz")}
    for f in ("ar1x", "ar1y", "ar1z", "ar2x", "ar2y", "ar2z", "ar3x", "ar3y", "ar3z"):
        g[f] = rng.uniform(-1, 1, n)
    g["vol"] = rng.uniform(0.5, 2.0, n)
    u = rng.normal(size=n)
    return g, u


FIELD_NAMES = ("x", "y", "z", "xc", "yc", "zc", "vol", "fx", "fy", "fz", "ar1x", "ar1y",
               "ar1z", "ar2x", "ar2y", "ar2z", "ar3x", "ar3y", "ar3z", "den", "apu", "apv",
               "apw", "u", "v", "w", "p", "gradu", "gradv", "gradw", "gradp")


def fields(dims, seed):
    """Seeded fields for the p' assembly stencils: a jittered, stretched
    structured mesh (nodes x,y,z; centres xc,yc,zc) and random flow data."""
    rng = np.random.default_rng(seed)
    ni, nj, nk = dims
    n = ni * nj * nk
    k, i, j = np.meshgrid(np.arange(nk), np.arange(ni), np.arange(nj), indexing="ij")
    jit = lambda: rng.uniform(-0.05, 0.05, (nk, ni, nj))
    f = {}
    f["x"] = (1.02 ** i + jit()).ravel()
    f["y"] = (0.9 * j + jit()).ravel()
    f["z"] = (1.1 * k + 0.1 * i + jit()).ravel()
    f["xc"] = (1.02 ** i - 0.5 + jit()).ravel()
    f["yc"] = (0.9 * j - 0.45 + jit()).ravel()
    f["zc"] = (1.1 * k - 0.55 + 0.1 * i + jit()).ravel()
    f["vol"] = rng.uniform(0.5, 1.5, n)
    for q in ("fx", "fy", "fz"):
        f[q] = rng.uniform(0.3, 0.7, n)
    for d, comp in (("1", "x"), ("2", "y"), ("3", "z")):
        for c in "xyz":
            f["ar" + d + c] = rng.uniform(0.8, 1.2, n) if c == comp else rng.uniform(-0.2, 0.2, n)
    f["den"] = rng.uniform(0.9, 1.1, n)
    for q in ("apu", "apv", "apw"):
        f[q] = rng.uniform(0.1, 1.0, n)
    for q in ("u", "v", "w", "p"):
        f[q] = rng.normal(size=n)
    for q in ("gradu", "gradv", "gradw", "gradp"):
        f[q] = rng.normal(size=3 * n)
    f["su"], f["sv"], f["sw"] = rng.normal(size=n), rng.normal(size=n), rng.normal(size=n)
    f["out0"] = [rng.normal(size=n) for _ in range(4)]   # acfe, acfw, fcf, bp before the call
    f["f123"] = [rng.normal(size=n) for _ in range(3)]
    f["vis"] = rng.uniform(0.01, 0.1, n)
    f["bt0"], f["bb0"] = rng.normal(size=n), rng.normal(size=n)   # bt, bb before the call
    return f


# fluxuvw scheme flags (the reference's module logicals), as the oracle's
# ORC_UVW_* bits: cds 1, cds4 2, luds 4, quick 8, flux_limiter 16, smart 32,
# avl 64, muscl 128, umist 256, gamma 512
UVW_SCHEMES = {"cds": 1, "cds4": 2, "luds": 4, "quick": 8, "cds_quick": 9,
               "smart": 16 | 32, "avl": 16 | 64, "muscl": 16 | 128, "umist": 16 | 256,
               "gamma": 16 | 512, "uds": 0}


def fluxuvw_inputs(f):
    """Array list of the reference driver's fluxuvw mode (fields, su, sv, sw,
    acfe, acfw, fcf, bp, then vis, bt, bb)."""
    return [f[q] for q in FIELD_NAMES] + [f["su"], f["sv"], f["sw"]] + f["out0"] + \
        [f["vis"], f["bt0"], f["bb0"]]


# (name, routine, kind, dims, seed, nsw, alfa, full arrays?)
SOLVER_CASES = [
    ("sipsol_general_12x10x9", "sipsol", "general", (12, 10, 9), 101, 25, 0.92, True),
    ("sipsol_general_20x37x23", "sipsol", "general", (20, 37, 23), 102, 12, 0.92, False),
    ("sipsol_lognormal_alfa0_16", "sipsol", "lognormal", (16, 16, 16), 103, 30, 0.0, True),
    ("cgstab_general_12x10x9", "cgstab_sip", "general", (12, 10, 9), 104, 400, 0.92, True),
    ("cgstab_lognormal_20", "cgstab_sip", "lognormal", (20, 20, 20), 105, 400, 0.92, True),
    ("cgstab_lognormal_36", "cgstab_sip", "lognormal", (36, 36, 36), 106, 400, 0.92, False),
    ("bicgilu_general_12x10x9", "cgstab", "general", (12, 10, 9), 107, 15, 0.0, True),
    ("bicgilu_lognormal_20", "cgstab", "lognormal", (20, 20, 20), 108, 400, 0.0, True),
    ("pcgjac_general_12x10x9", "pcgjac", "general", (12, 10, 9), 109, 40, 0.0, True),
    ("pcgjac_lognormal_20", "pcgjac", "lognormal", (20, 20, 20), 110, 600, 0.0, True),
]
ASSEMBLY_CASES = [
    ("fluxmass", (12, 10, 9), 301), ("fluxmass", (20, 17, 11), 302),
    ("piso_assemble", (12, 10, 9), 303), ("piso_assemble_ipc0", (12, 10, 9), 304),
    ("assemble_p", (12, 10, 9), 305), ("fluxmc", (12, 10, 9), 306),
]
GHOST_VARS = ("u", "v", "w", "te", "ed", "t", "vis")
GEO12 = ("fx", "fy", "fz", "ar1x", "ar1y", "ar1z", "ar2x", "ar2y", "ar2z", "ar3x", "ar3y", "ar3z")


def ghost_inputs(dims, seed):
    """Seeded inputs of the ghost-cell update: the seven variables, the
    geometry, boundary types 1..4 per boundary cell (lbw, lbe, lbs, lbn, lbb,
    lbt) and the inlet values uin, vin, win, tein, edin, tin."""
    rng = np.random.default_rng(seed)
    ni, nj, nk = dims
    n = ni * nj * nk
    var = {q: rng.normal(size=n) for q in GHOST_VARS}
    var["te"] = rng.uniform(0.1, 1.0, n)
    var["ed"] = rng.uniform(0.5, 2.0, n)
    var["vis"] = rng.uniform(0.01, 0.1, n)
    geo = {q: rng.uniform(0.3, 0.7, n) for q in ("fx", "fy", "fz")}
    for q in GEO12[3:]:
        geo[q] = rng.uniform(-1, 1, n)
    lens = (nj * nk, nj * nk, ni * nk, ni * nk, ni * nj, ni * nj)
    lb = [rng.integers(1, 5, m).astype(np.int32) for m in lens]
    inl = rng.normal(size=6)
    inl[3:5] = rng.uniform(0.5, 1.0, 2)
    return var, geo, lb, inl


def ghost_ref_arrays(dims, var, geo, lb, inl):
    n = dims[0] * dims[1] * dims[2]
    pads = []
    for a in lb:
        x = np.zeros(n)
        x[:a.size] = a
        pads.append(x)
    il = np.zeros(n)
    il[:6] = inl
    return [var[q] for q in GHOST_VARS] + [geo[q] for q in GEO12] + pads + [il]


GHOST_CASES = [((12, 10, 9), 601), ((20, 17, 11), 602)]
# the periodic variant (update_values_at_ghostcells.f90) through the driver
# linked with it
GHOST_PERIODIC_CASES = [((12, 10, 9), 611), ((20, 17, 11), 612)]
FLUXUVW_CASES = [((12, 10, 9), 501, "cds", 1.0), ((12, 10, 9), 502, "luds", 1.0),
                 ((12, 10, 9), 503, "quick", 1.0), ((12, 10, 9), 504, "cds4", 0.7),
                 ((12, 10, 9), 505, "smart", 1.0), ((12, 10, 9), 506, "gamma", 1.0),
                 ((12, 10, 9), 507, "cds_quick", 0.5), ((12, 10, 9), 508, "umist", 1.0),
                 ((12, 10, 9), 509, "avl", 1.0), ((12, 10, 9), 510, "muscl", 1.0),
                 ((12, 10, 9), 511, "uds", 1.0), ((20, 17, 11), 512, "cds", 1.0)]
UVW_OUT = ("acfe", "acfw", "su", "sv", "sw", "bp", "bt", "bb")
STENCIL_CASES = [
    ("grad_gauss_12x10x9", "grad_gauss", (12, 10, 9), 201),
    ("grad_gauss_20x37x23", "grad_gauss", (20, 37, 23), 202),
    ("bpres_12x10x9", "bpres", (12, 10, 9), 203),
]


def digest(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype="<f8").tobytes()).hexdigest()


# outbc: (dims, seed, face flags lout + 2 low + 4 loe + ... + 64 lot)
OUTBC_VARS = ("u", "v", "w", "f1", "f2", "f3", "te", "ed", "vis", "t", "vart", "con", "utt",
              "vtt", "wtt")
OUTBC_CASES = [((12, 10, 9), 701, 1 + 4), ((12, 10, 9), 702, 127), ((20, 17, 11), 703, 1 + 2 + 8 + 64),
               ((12, 10, 9), 704, 126)]


def outbc_inputs(dims, seed):
    """u..wtt, den, ar1x..ar1z, boundary types (outlets common) and flowin"""
    rng = np.random.default_rng(seed)
    ni, nj, nk = dims
    n = ni * nj * nk
    var = {q: rng.normal(size=n) for q in OUTBC_VARS}
    var["u"] = 1.0 + 0.2 * rng.normal(size=n)
    geo = {"den": rng.uniform(0.9, 1.1, n)}
    for q in ("ar1x", "ar1y", "ar1z"):
        geo[q] = rng.uniform(0.5, 1.5, n) if q == "ar1x" else rng.uniform(-0.1, 0.1, n)
    lens = (nj * nk, nj * nk, ni * nk, ni * nk, ni * nj, ni * nj)
    lb = [np.where(rng.uniform(size=m) < 0.7, 2, rng.integers(1, 5, m)).astype(np.int32)
          for m in lens]
    return var, geo, lb, float(rng.uniform(2.0, 5.0))


def outbc_ref_arrays(dims, var, geo, lb):
    n = dims[0] * dims[1] * dims[2]
    pads = []
    for a in lb:
        x = np.zeros(n)
        x[:a.size] = a
        pads.append(x)
    return [var[q] for q in ("u", "v", "w", "f1", "f2", "f3")] + \
        [geo[q] for q in ("den", "ar1x", "ar1y", "ar1z")] + \
        [var[q] for q in OUTBC_VARS[6:]] + [np.stack(pads, axis=1).T.copy().ravel()]


CBC_CASES = [((12, 12, 9), 801), ((17, 17, 11), 802)]
CBC_VARS = ("u", "v", "w", "f1", "f2")

# calc_vis_les: (dims, seed); ni <= nk + 1 keeps the top wall plane lk(nimm)
# and its ghost plane inside the array
LES_CASES = [((12, 10, 12), 901), ((10, 13, 11), 902), ((16, 14, 16), 903)]
LES_IN = ("u", "v", "w", "vol", "den", "fz", "xc", "yc", "zc", "ar3x", "ar3y", "ar3z", "vis",
          "strain", "walldist")


def les_inputs(dims, seed):
    rng = np.random.default_rng(seed)
    ni, nj, nk = dims
    n = ni * nj * nk
    k, i, j = np.meshgrid(np.arange(nk), np.arange(ni), np.arange(nj), indexing="ij")
    f = {}
    sc = rng.uniform(0.01, 2.0, n)
    f["u"], f["v"], f["w"] = sc * rng.normal(size=n), sc * rng.normal(size=n), 0.3 * rng.normal(size=n)
    f["vol"] = rng.uniform(0.5, 1.5, n)
    f["den"] = rng.uniform(0.95, 1.05, n)
    f["fz"] = rng.uniform(0.3, 0.7, n)
    f["xc"] = (i + 0.1 * rng.normal(size=i.shape)).ravel()
    f["yc"] = (j + 0.1 * rng.normal(size=j.shape)).ravel()
    f["zc"] = (k + 0.1 * rng.normal(size=k.shape)).ravel()
    f["ar3x"], f["ar3y"] = rng.uniform(-0.1, 0.1, n), rng.uniform(-0.1, 0.1, n)
    f["ar3z"] = rng.uniform(0.8, 1.2, n)
    f["vis"] = rng.uniform(0.02, 0.05, n)
    f["strain"] = rng.uniform(0.0, 3.0, n)
    f["walldist"] = rng.uniform(0.05, 2.0, n)
    par = dict(viscos=0.02, densit=1.0, cappa=0.41, elog=8.432, urf=0.7,
               nx=ni - 2, ny=nj - 2, nz=nk - 2)
    return f, par


def les_ref_arrays(f, par):
    return [f[q] for q in LES_IN] + [np.array([par["viscos"], par["densit"], par["cappa"],
                                               par["elog"]]),
                                     np.array([par["nx"], par["ny"], par["nz"]], dtype=np.int32)]


def solver_cases(only=None):
    """Solver fixtures from the reference build; `only`: names to rewrite
    (None: all)."""
    for name, routine, kind, dims, seed, nsw, alfa, full in SOLVER_CASES:
        if only is not None and name not in only:
            continue
        s, fi = solver_inputs(kind, dims, seed)
        arrays = [fi, s.ae, s.aw, s.an, s.as_, s.at, s.ab, s.ap, s.su, s.res] + \
                 [getattr(s, f) for f in B_FIELDS]
        outs, resor, log = O.run_ref(routine, dims, arrays, nsw=nsw, alfa=alfa, sor=1e-8)
        info = O.parse_log(log)
        rec = dict(routine=routine, kind=kind, dims=np.array(dims), seed=seed, nsw=nsw,
                   alfa=alfa, resor=resor, iters=info["iters"], resl_printed=info["resl"])
        for f, a in zip(SOLVER_OUT, outs):
            if full:
                rec[f] = a
            rec[f + "_sha256"] = digest(a)
        np.savez_compressed(os.path.join(HERE, name + ".npz"), **rec)
        print(f"{name}: iters {info['iters']} resor {resor:.6e}")


# SHA-256 of the oracle's log-normal systems at sizes too large for a GPU
# test to rebuild on the CPU: the generator's device output must match them
LOGNORMAL_DIGESTS = [((260, 260, 260), 2024, 1.0), ((132, 132, 132), 7, 1.0)]


def lognormal_digests():
    import json
    out = []
    for dims, seed, sigma in LOGNORMAL_DIGESTS:
        sy = O.build_lognormal(dims, seed, sigma)
        out.append(dict(dims=list(dims), seed=seed, sigma=sigma,
                        sha256={f: digest(getattr(sy, f))
                                for f in ("ae", "aw", "an", "as_", "at", "ab", "ap", "su")}))
    with open(os.path.join(HERE, "lognormal_digests.json"), "w") as fh:
        json.dump(out, fh, indent=1)
    print("lognormal_digests.json: written")


def ghost_periodic_cases():
    for dims, seed in GHOST_PERIODIC_CASES:
        var, geo, lb, inl = ghost_inputs(dims, seed)
        outs, _, _ = O.run_ref("ghost", dims, ghost_ref_arrays(dims, var, geo, lb, inl),
                               driver=O.REF_DRIVER_PERIODIC)
        rec = dict(dims=np.array(dims), seed=seed, **{q: a for q, a in zip(GHOST_VARS, outs)})
        name = f"ghost_periodic_{dims[0]}x{dims[1]}x{dims[2]}"
        np.savez_compressed(os.path.join(HERE, name + ".npz"), **rec)
        print(f"{name}: written")


def main():
    if len(sys.argv) > 1 and sys.argv[1] == "--ghost-periodic":
        ghost_periodic_cases()
        return
    if len(sys.argv) > 1 and sys.argv[1] == "--lognormal":
        lognormal_digests()         # oracle only: needs no reference build
        return
    if not O.ref_available():
        sys.exit("oracle/_ref/ref_driver is not built (make -C oracle ref)")
    if len(sys.argv) > 1 and sys.argv[1] == "--solvers":
        # rewrite only the named solver fixtures (all solver fixtures when none is named)
        solver_cases(set(sys.argv[2:]) or None)
        return
    solver_cases()
    for routine, dims, seed in ASSEMBLY_CASES:
        f = fields(dims, seed)
        ipc = 0 if routine.endswith("_ipc0") else 1
        mode = routine.replace("_ipc0", "")
        rec = dict(dims=np.array(dims), seed=seed, ipc=ipc)
        for d in range(3):
            arrays = [f[q] for q in FIELD_NAMES] + [f["su"], f["sv"], f["sw"]] + f["out0"]
            outs, _, _ = O.run_ref(mode, dims, arrays, nsw=d, alfa=float(ipc))
            for q, a in zip(("acfe", "acfw", "fcf", "bp", "su"), outs):
                rec[f"{q}_{d}"] = a
        name = f"{routine}_{dims[0]}x{dims[1]}x{dims[2]}"
        np.savez_compressed(os.path.join(HERE, name + ".npz"), **rec)
        print(f"{name}: written")
    for dims, seed, flags in OUTBC_CASES:
        var, geo, lb, flowin = outbc_inputs(dims, seed)
        outs, _, _ = O.run_ref("outbc", dims, outbc_ref_arrays(dims, var, geo, lb), nsw=flags,
                               alfa=flowin)
        rec = dict(dims=np.array(dims), seed=seed, flags=flags,
                   **{q: a for q, a in zip(OUTBC_VARS, outs)})
        name = f"outbc_{dims[0]}x{dims[1]}x{dims[2]}_{seed}"
        np.savez_compressed(os.path.join(HERE, name + ".npz"), **rec)
    for dims, seed in CBC_CASES:
        rng = np.random.default_rng(seed)
        arrs = [rng.normal(size=dims[0] * dims[1] * dims[2]) for _ in CBC_VARS]
        outs, _, _ = O.run_ref("cbc", dims, arrs)
        rec = dict(dims=np.array(dims), seed=seed, **{q: a for q, a in zip(CBC_VARS, outs)})
        np.savez_compressed(os.path.join(HERE, f"cbc_{dims[0]}x{dims[1]}x{dims[2]}.npz"), **rec)
    for dims, seed in LES_CASES:
        f, par = les_inputs(dims, seed)
        outs, _, _ = O.run_ref("les", dims, les_ref_arrays(f, par), alfa=par["urf"])
        rec = dict(dims=np.array(dims), seed=seed, vis=outs[0])
        np.savez_compressed(os.path.join(HERE, f"les_{dims[0]}x{dims[1]}x{dims[2]}.npz"), **rec)
    for dims, seed in GHOST_CASES:
        var, geo, lb, inl = ghost_inputs(dims, seed)
        outs, _, _ = O.run_ref("ghost", dims, ghost_ref_arrays(dims, var, geo, lb, inl))
        rec = dict(dims=np.array(dims), seed=seed, **{q: a for q, a in zip(GHOST_VARS, outs)})
        name = f"ghost_{dims[0]}x{dims[1]}x{dims[2]}"
        np.savez_compressed(os.path.join(HERE, name + ".npz"), **rec)
        print(f"{name}: written")
    ghost_periodic_cases()
    for dims, seed, scheme, gds in FLUXUVW_CASES:
        f = fields(dims, seed)
        rec = dict(dims=np.array(dims), seed=seed, flags=UVW_SCHEMES[scheme], gds=gds)
        for d in range(3):
            outs, _, _ = O.run_ref("fluxuvw", dims, fluxuvw_inputs(f),
                                   nsw=d + 8 * UVW_SCHEMES[scheme], alfa=gds)
            for q, a in zip(UVW_OUT, outs):
                rec[f"{q}_{d}"] = a
        name = f"fluxuvw_{scheme}_{dims[0]}x{dims[1]}x{dims[2]}"
        np.savez_compressed(os.path.join(HERE, name + ".npz"), **rec)
        print(f"{name}: written")
    for dims, seed in (((12, 10, 9), 401), ((20, 17, 11), 402)):
        f = fields(dims, seed)
        outs, _, _ = O.run_ref("continuity", dims, f["f123"])
        np.savez_compressed(os.path.join(HERE, f"continuity_{dims[0]}x{dims[1]}x{dims[2]}.npz"),
                            dims=np.array(dims), seed=seed, sums=outs[0])
        print(f"continuity {dims}: {outs[0]}")
    for name, routine, dims, seed in STENCIL_CASES:
        g, u = geometry(dims, seed)
        if routine == "grad_gauss":
            arrays = [u] + [g[f] for f in O.GEOM_FIELDS]
            outs, _, _ = O.run_ref("grad_gauss", dims, arrays)
            rec = dict(dims=np.array(dims), seed=seed, dudx=outs[0], dudy=outs[1], dudz=outs[2])
        else:
            arrays = [u, g["fx"], g["fy"], g["fz"]]
            outs, _, _ = O.run_ref("bpres", dims, arrays)
            rec = dict(dims=np.array(dims), seed=seed, p=outs[0])
        np.savez_compressed(os.path.join(HERE, name + ".npz"), **rec)
        print(f"{name}: written")


if __name__ == "__main__":
    main()
