"""Two of the benchmark's configurations end to end on the GPU against the
oracle:

* C1, the 32^3 lid-driven cavity (Re 100, seeded U[-0.1, 0.1] velocities,
  the lid as the top face's inlet ghost values): one SIMPLE iteration --
  ghost fill, calcuvw, calcp with sipsol (alpha 0.92) to 1e-8 -- every
  array of the module state and every solver report bit-identical;
* C2, the 128^3 grid stretched 1.05 toward every wall: GRAD_GAUSS and fluxuvw
  (CDS, gds 1) on the seeded field, the p' system that calcuvw + calcp
  assemble, and its solution by pcgsip (alpha 0), iccg and cgstab_sip
  (alpha 0.92) to 1e-8, each against SHA-256 digests of the oracle's arrays
  (tests/golden/c2_128.json, made by tests/golden/make_configs.py): bit-
  identical to the oracle with the GPU's compensated sums, whose p' is within
  the recorded max_rel_vs_plain of the reference's plain sums: 1e-10, or four
  times the solve's own spread between two summation orders when larger.
"""
import hashlib
import json
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
import configcase as CC  # noqa: E402
import flowcase as FC  # noqa: E402

pytestmark = pytest.mark.gpu


def _digest(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype="<f8").tobytes()).hexdigest()


def _dev(f):
    import torch
    return {q: torch.from_numpy(np.ascontiguousarray(a).copy()).cuda() for q, a in f.items()}


def _host(d):
    return {q: t.cpu().numpy() for q, t in d.items()}


def test_c1_cavity_simple_iteration(gpu, O):
    import torch
    dims, f, lb, inl = CC.state(O, "cavity", 32, 1, fill=False)
    print(f"C1 cavity {dims}: ghost fill, calcuvw, calcp (sipsol alfa 0.92)", flush=True)
    fo = {q: a.copy() for q, a in f.items()}
    spv_o = np.zeros_like(f["u"])
    O.ghostcells(dims, [fo[q] for q in CC.GHOST7], [fo[q] for q in CC.GEO12], lb, inl)
    with O.plane_sums(2):
        ru_o = O.calcuvw(dims, fo, spv_o, CC.drv_opts(O.DrvOpts, dims), CC.uvw_opts(O.UvwOpts, dims))
        rp_o = O.calcp(dims, fo, CC.drv_opts(O.DrvOpts, dims))
    fd = _dev(f)
    spv_g = torch.zeros(f["u"].size, dtype=torch.float64, device="cuda")
    lbd = [torch.from_numpy(a.copy()).cuda() for a in lb]
    with gpu.Context(*dims, device_pointers=True) as ctx:
        ctx.update_ghostcells([fd[q] for q in CC.GHOST7], [fd[q] for q in CC.GEO12], lbd, inl)
        ru_g = ctx.calcuvw(fd, spv_g, CC.drv_opts(gpu.DrvOpts, dims), CC.uvw_opts(gpu.UvwOpts, dims))
        rp_g = ctx.calcp(fd, CC.drv_opts(gpu.DrvOpts, dims))
    fg = _host(fd)
    print("  momentum sweeps", [r["iters"] for r in ru_g], "p' sweeps", rp_g[0][0]["iters"],
          "continuity", rp_g[1][0], flush=True)
    for a, b in zip(ru_o + rp_o[0], ru_g + rp_g[0]):
        assert (a["iters"], a["resor"], a["resl"]) == (b["iters"], b["resor"], b["resl"])
    assert rp_g[0][0]["converged"]
    np.testing.assert_allclose(rp_g[1][0], rp_o[1][0], rtol=1e-12, atol=1e-14)
    bad = [q for q in FC.OUTPUTS if not np.array_equal(fo[q], fg[q])]
    assert not bad, f"arrays differ from the oracle: {bad}"
    np.testing.assert_array_equal(spv_g.cpu().numpy(), spv_o)
    # the lid moved the fluid: the top interior row carries the largest u
    ni, nj, nk = dims
    u = fg["u"].reshape(nk, ni, nj)
    assert np.abs(u[-3]).max() > 3 * np.abs(u[nk // 2]).max()


@pytest.fixture(scope="module")
def c2(O):
    path = os.path.join(HERE, "golden", "c2_128.json")
    with open(path) as fh:
        rec = json.load(fh)
    dims, f, lb, inl = CC.state(O, "stretched", rec["n"], rec["seed"])
    for q, h in rec["state"].items():
        assert _digest(f[q]) == h, f"C2 inputs differ from the fixture's: {q}"
    return rec, dims, f


def test_c2_gradients_and_fluxuvw(gpu, O, c2):
    import torch
    rec, dims, f = c2
    print(f"C2 {dims}: GRAD_GAUSS and fluxuvw (CDS, gds 1)", flush=True)
    fd = _dev(f)
    n = f["u"].size
    with gpu.Context(*dims, device_pointers=True) as ctx:
        geom = {q: fd[q] for q in gpu.GEOM_FIELDS}
        for q in ("u", "v", "w"):
            g3 = [torch.zeros(n, dtype=torch.float64, device="cuda") for _ in range(3)]
            ctx.grad_gauss(fd[q], geom, *g3)
            fd["grad" + q] = torch.stack(g3, dim=1).reshape(-1).contiguous()
            assert _digest(fd["grad" + q].cpu().numpy()) == rec["stencils"]["grad" + q], q
        for d in range(3):
            o = {q: torch.zeros(n, dtype=torch.float64, device="cuda")
                 for q in ("acfe", "acfw", "su", "sv", "sw", "bp", "bt", "bb")}
            ctx.fluxuvw(fd, d, fd["vis"], o["acfe"], o["acfw"], fd[("f1", "f2", "f3")[d]], o["su"],
                        o["sv"], o["sw"], o["bp"], o["bt"], o["bb"], 1, 1.0)
            for q, t in o.items():
                assert _digest(t.cpu().numpy()) == rec["stencils"][f"fluxuvw_{q}_{d}"], (q, d)


@pytest.mark.parametrize("solver", ["pcgsip", "iccg", "cgstab_sip"])
def test_c2_p_solves(gpu, O, c2, solver):
    import torch
    rec, dims, f = c2
    want = rec["solvers"][solver]
    print(f"C2 {dims}: calcuvw + calcp assembly, then {solver} to 1e-8", flush=True)
    # 1e-10 of the reference's summation order, or what the solve itself moves
    # between two summation orders when that is more (BiCGStab amplifies
    # rounding: 2.9e-9 against a per-plane/running-sum spread of the same size)
    bound = max(1e-10, 4.0 * want["order_sensitivity"])
    assert want["max_rel_vs_plain"] <= bound and want["iters"] == want["iters_plain"]
    fd = _dev(f)
    spv = torch.zeros(f["u"].size, dtype=torch.float64, device="cuda")
    with gpu.Context(*dims, device_pointers=True) as ctx:
        ctx.calcuvw(fd, spv, CC.drv_opts(gpu.DrvOpts, dims), CC.uvw_opts(gpu.UvwOpts, dims))
        ctx.calcp(fd, CC.drv_opts(gpu.DrvOpts, dims, solver=4, alfa=0.0, nsw=1))
        s = CC.p_system(O, dims, {q: fd[q].cpu().numpy() for q in CC.SYSTEM})
        for q in CC.SYSTEM:
            assert _digest(getattr(s, "as_" if q == "as" else q)) == rec["system"][q], q
        coef = gpu.Coef(**{q: torch.from_numpy(getattr(s, q).copy()).cuda()
                           for q in ("ae", "aw", "an", "as_", "at", "ab", "ap", "su")})
        fi = torch.zeros(s.n, dtype=torch.float64, device="cuda")
        kw = {k: v for k, v in want.items() if k == "alfa"}
        rep = getattr(ctx, solver)(fi, coef, sor=1e-8, nsw=4000, **kw)
    print(f"  {rep['iters']} iterations, {rep['ms']:.1f} ms, oracle {want['iters']}", flush=True)
    assert rep["converged"] and rep["iters"] == want["iters"]
    assert (rep["resor"], rep["resl"]) == (want["resor"], want["resl"])
    assert _digest(fi.cpu().numpy()) == want["fi"]
