"""GPU p' assembly stencils against the oracle (bit-identical) and the
continuity sums (deterministic FP64 reduction, reassociation only)."""
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
import make_golden as MG  # noqa: E402
from test_oracle import oracle_assembly, oracle_courant, oracle_fluxuvw, oracle_ghost  # noqa: E402

DIMS = [(12, 10, 9), (20, 17, 11), (40, 23, 19)]


def fields_for(gpu, dims, seed):
    f = MG.fields(dims, seed)
    return f, {q: f[q] for q in MG.FIELD_NAMES}


@pytest.mark.parametrize("dims", DIMS)
@pytest.mark.parametrize("routine", ["fluxmass", "piso_ipc1", "piso_ipc0", "assemble_p", "fluxmc"])
def test_assembly_bitwise(gpu, O, dims, routine):
    f, flds = fields_for(gpu, dims, 77)
    with gpu.Context(*dims) as ctx:
        for d in range(3):
            ipc = 0 if routine == "piso_ipc0" else 1
            name = {"piso_ipc1": "piso", "piso_ipc0": "piso"}.get(routine, routine)
            ref = oracle_assembly(O, name, dims, f, d, ipc)
            acfe, acfw, fcf, bp = [a.copy() for a in f["out0"]]
            su = f["su"].copy()
            if routine == "fluxmass":
                ctx.fluxmass(flds, d, acfe, acfw, fcf)
            elif routine.startswith("piso"):
                ctx.piso_assemble_pressure_eq(flds, d, acfe, acfw, fcf, ipc, bp, su)
            elif routine == "assemble_p":
                ctx.assemble_pressure_eq(flds, d, acfe, acfw, fcf, f["su"], f["sv"], f["sw"])
            else:
                ctx.fluxmc(flds, d, acfe)
            got = dict(acfe=acfe, acfw=acfw, fcf=fcf, bp=bp, su=su)
            keys = ("acfe",) if routine == "fluxmc" else ("acfe", "acfw", "fcf", "bp", "su")
            for q in keys:
                np.testing.assert_array_equal(got[q], ref[q], err_msg=f"{routine} {q} dir {d}")


@pytest.mark.parametrize("dims", DIMS)
def test_continuity(gpu, O, dims):
    import ctypes as C
    f = MG.fields(dims, 91)
    out = (C.c_double * 2)()
    O.lib().orc_continuity(*dims, *[a.ctypes.data_as(C.POINTER(C.c_double)) for a in f["f123"]], out)
    with gpu.Context(*dims) as ctx:
        got = ctx.continuity(*f["f123"])
    assert got[1] == pytest.approx(out[1], rel=1e-13)
    assert got[0] == pytest.approx(out[0], rel=1e-10, abs=1e-12 * out[1])


@pytest.mark.parametrize("dims", DIMS)
@pytest.mark.parametrize("scheme", sorted(MG.UVW_SCHEMES))
def test_fluxuvw_bitwise(gpu, O, dims, scheme):
    """Every convection scheme of fluxuvw, three directions: bit-identical to
    the oracle (which is bit-identical to the reference build), including the
    stale high-order flux on faces QUICK skips."""
    f, flds = fields_for(gpu, dims, 81)
    flags, gds = MG.UVW_SCHEMES[scheme], 0.8
    with gpu.Context(*dims) as ctx:
        for d in range(3):
            ref = oracle_fluxuvw(O, dims, f, d, flags, gds)
            acfe, acfw, fcf, bp = [a.copy() for a in f["out0"]]
            su, sv, sw = f["su"].copy(), f["sv"].copy(), f["sw"].copy()
            bt, bb = f["bt0"].copy(), f["bb0"].copy()
            ctx.fluxuvw(flds, d, f["vis"], acfe, acfw, fcf, su, sv, sw, bp, bt, bb, flags, gds)
            got = dict(acfe=acfe, acfw=acfw, su=su, sv=sv, sw=sw, bp=bp, bt=bt, bb=bb)
            for q in MG.UVW_OUT:
                np.testing.assert_array_equal(got[q], ref[q], err_msg=f"{scheme} {q} dir {d}")


@pytest.mark.parametrize("dims", DIMS)
def test_fluxuvw_fp32_close_to_fp64(gpu, dims):
    """fluxuvw's CDS branch with its face arithmetic in FP32 (CAP_UVW_FP32)
    against the FP64 kernel on the same fields.  The bound is set by the
    fields' own conditioning: the FP64 kernel run on the inputs rounded to
    FP32 moves each output by some amount (large where a random face's
    xpn . S nearly vanishes); the FP32 kernel must stay within 16 times that
    plus 1e-6 of the array's largest value.  Other schemes with
    CAP_UVW_FP32 are refused."""
    f, flds = fields_for(gpu, dims, 81)
    cds, fp32, gds = MG.UVW_SCHEMES["cds"], 1024, 0.8
    r32 = lambda a: a.astype(np.float32).astype(np.float64)  # noqa: E731
    flds_r = {q: r32(a) for q, a in flds.items()}

    def run(fl, vis, fcf0, flags):
        acfe, acfw, fcf, bp = [a.copy() for a in f["out0"]]
        fcf[:] = fcf0
        su, sv, sw = f["su"].copy(), f["sv"].copy(), f["sw"].copy()
        bt, bb = f["bt0"].copy(), f["bb0"].copy()
        ctx.fluxuvw(fl, d, vis, acfe, acfw, fcf, su, sv, sw, bp, bt, bb, flags, gds)
        return dict(acfe=acfe, acfw=acfw, su=su - f["su"], sv=sv - f["sv"], sw=sw - f["sw"],
                    bp=bp, bt=bt, bb=bb)
    with gpu.Context(*dims) as ctx:
        for d in range(3):
            fcf0 = f["out0"][2]
            ref = run(flds, f["vis"], fcf0, cds)
            rnd = run(flds_r, r32(f["vis"]), r32(fcf0), cds)
            got = run(flds, f["vis"], fcf0, cds | fp32)
            for q in MG.UVW_OUT:
                scale = np.max(np.abs(ref[q]))
                if q in ("bp", "bt", "bb"):     # scratch, cleared after use
                    assert scale == 0 and not np.any(got[q]), q
                    continue
                cond = np.max(np.abs(rnd[q] - ref[q]))
                err = np.max(np.abs(got[q] - ref[q]))
                print(f"dir {d} {q}: FP32 error {err / scale:.2e}, input-rounding effect "
                      f"{cond / scale:.2e} (of the largest value)")
                assert err <= 16 * cond + 1e-6 * scale, f"{q} dir {d}"
        acfe, acfw, fcf, bp = [a.copy() for a in f["out0"]]
        with pytest.raises(gpu.CapError):
            ctx.fluxuvw(flds, 0, f["vis"], acfe, acfw, fcf, f["su"].copy(), f["sv"].copy(),
                        f["sw"].copy(), bp, f["bt0"].copy(), f["bb0"].copy(),
                        MG.UVW_SCHEMES["quick"] | fp32, gds)


@pytest.mark.parametrize("dims", DIMS + [(36, 36, 36)])
@pytest.mark.parametrize("types", ["mixed", "inlet", "outlet", "symmetry", "wall"])
def test_ghostcells_bitwise(gpu, O, dims, types):
    """The on-device ghost-cell fill: every boundary type on all six faces,
    bit-identical to the oracle (itself pinned to the reference build)."""
    var, geo, lb, inl = MG.ghost_inputs(dims, 97)
    if types != "mixed":
        code = {"inlet": 1, "outlet": 2, "symmetry": 3, "wall": 4}[types]
        for a in lb:
            a[:] = code
    ref = oracle_ghost(O, dims, var, geo, lb, inl)
    got = [var[q].copy() for q in MG.GHOST_VARS]
    with gpu.Context(*dims) as ctx:
        ctx.update_ghostcells(got, [geo[q] for q in MG.GEO12], lb, inl)
    for q, a in zip(MG.GHOST_VARS, got):
        np.testing.assert_array_equal(a, ref[q], err_msg=q)


@pytest.mark.parametrize("dims", DIMS + [(36, 36, 36)])
@pytest.mark.parametrize("fix", [0, 1])
def test_courant(gpu, O, dims, fix):
    """su and CoNum bit-identical; meanCoNum bit-identical to the oracle with
    compensated sums and within 1e-13 of the reference's running sums."""
    f = MG.fields(dims, 93)
    f1, f2, f3 = f["f123"]
    vol = f["vol"]
    su0, co0, mean0, ts0, tm0 = oracle_courant(O, dims, f1, f2, f3, vol, 0.01, 2.0, fix, 0.8)
    with O.plane_sums(2):
        _, coa, meana, tsa, tma = oracle_courant(O, dims, f1, f2, f3, vol, 0.01, 2.0, fix, 0.8)
    su = np.full_like(f1, np.nan)
    with gpu.Context(*dims) as ctx:
        co, mean, ts, tm = ctx.courant(f1, f2, f3, vol, su, 0.01, 2.0, bool(fix), 0.8)
    np.testing.assert_array_equal(su, su0)
    assert co == coa and mean == meana and ts == tsa and tm == tma
    assert co == co0 and mean == pytest.approx(mean0, rel=1e-13)


@pytest.mark.parametrize("dims", DIMS + [(36, 36, 36), (68, 52, 36)])
@pytest.mark.parametrize("device", [False, True])
def test_ghostcells_periodic_bitwise(gpu, O, dims, device):
    """The channel's periodic ghost fill (update_values_at_ghostcells.f90):
    bit-identical to the oracle (itself pinned to the reference build)."""
    import torch
    var, geo, lb, inl = MG.ghost_inputs(dims, 98)
    ref = {q: var[q].copy() for q in MG.GHOST_VARS}
    O.ghostcells_periodic(dims, [ref[q] for q in MG.GHOST_VARS], geo["fz"])
    if device:
        got = [torch.from_numpy(var[q].copy()).cuda() for q in MG.GHOST_VARS]
        fz = torch.from_numpy(geo["fz"].copy()).cuda()
    else:
        got = [var[q].copy() for q in MG.GHOST_VARS]
        fz = geo["fz"]
    with gpu.Context(*dims, device_pointers=device) as ctx:
        ctx.update_ghostcells_periodic(got, fz)
    for q, a in zip(MG.GHOST_VARS, got):
        np.testing.assert_array_equal(a.cpu().numpy() if device else a, ref[q], err_msg=q)
