"""The pressure-correction drivers on the GPU (calcp, PISO, PIMPLE, outbc,
correctBoundaryConditions, PISO_getHbyA) against the oracle's restatement,
run with the GPU's compensated scalar products (plane_sums(2)).

Every array of the module state must be bit-identical, and so must the
solver reports (iterations, initial and final L1 residuals).  The continuity
sums are reductions in a different order (relative 1e-12)."""
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
import flowcase as FC  # noqa: E402

DIMS = (18, 14, 12)


def _set_lb(o, lb, dev):
    keep = []
    for name, a in lb.items():
        if dev:
            import torch
            t = torch.from_numpy(a).cuda()
            keep.append(t)
            setattr(o, name, t.data_ptr())
        else:
            keep.append(a)
            setattr(o, name, a.ctypes.data)
    return keep


def run_pair(gpu, O, dims, seed, kind, outlet=False, **kw):
    """(oracle state, GPU state, oracle result, GPU result) of one driver call"""
    import torch
    f = FC.flow_case(dims, seed)
    kw.setdefault("ijkpr", FC.ref_cell(dims))
    if outlet:
        kw.update(lout=1, loe=1, low=1, flowin=float(np.sum(f["ar1x"][:dims[1] * dims[0]])),
                  const_mflux=0)
    oo, go = FC.opts(O.DrvOpts, **kw), FC.opts(gpu.DrvOpts, **kw)
    keep = []
    if outlet:
        lb = FC.boundary_types(dims)
        keep += _set_lb(oo, lb, False) + _set_lb(go, lb, True)
    fo = {q: a.copy() for q, a in f.items()}
    fd = {q: torch.from_numpy(a.copy()).cuda() for q, a in f.items()}
    with gpu.Context(*dims, device_pointers=True) as ctx:
        with O.plane_sums(2):
            if kind == "calcp":
                ro = O.calcp(dims, fo, oo)
            elif kind in ("piso", "pimple"):
                ro = O.piso(dims, fo, oo, pimple=kind == "pimple")
            elif kind == "outbc":
                ro = O.outbc(dims, fo, oo)
            elif kind == "correct_bc":
                ro = O.correct_bc(dims, fo)
            else:
                ro = O.gethbya(dims, fo, oo)
        if kind == "calcp":
            rg = ctx.calcp(fd, go)
        elif kind in ("piso", "pimple"):
            rg = ctx.piso(fd, go, pimple=kind == "pimple")
        elif kind == "outbc":
            rg = ctx.outbc(fd, go)
        elif kind == "correct_bc":
            rg = ctx.correct_bc(fd)
        else:
            rg = ctx.gethbya(fd, go)
        torch.cuda.synchronize()
    del keep
    return fo, {q: t.cpu().numpy() for q, t in fd.items()}, ro, rg


def check_state(fo, fg, what):
    bad = []
    for q in FC.OUTPUTS:
        if not np.array_equal(fo[q], fg[q]):
            d = np.max(np.abs(fo[q] - fg[q])) / max(np.max(np.abs(fo[q])), 1e-300)
            bad.append(f"{q} (max rel {d:.2e})")
    assert not bad, f"{what}: arrays differ from the oracle: {', '.join(bad)}"


def check_reports(ro, rg, what):
    (rep_o, cont_o), (rep_g, cont_g) = ro, rg
    for a, b in zip(rep_o, rep_g):
        assert (a["iters"], a["resor"], a["resl"]) == (b["iters"], b["resor"], b["resl"]), what
        assert a["converged"] == b["converged"], what
    for a, b in zip(cont_o, cont_g):
        np.testing.assert_allclose(b, a, rtol=1e-12, atol=1e-12 * abs(a[0]), err_msg=what)


@pytest.mark.parametrize("solver,alfa,ncorr,npcor", [(4, 0.0, 2, 2), (1, 0.92, 1, 1),
                                                     (3, 0.0, 1, 2), (5, 0.92, 2, 1)])
def test_calcp_bitwise(gpu, O, solver, alfa, ncorr, npcor):
    what = f"calcp solver {solver} ncorr {ncorr} npcor {npcor}"
    print(what, flush=True)
    fo, fg, ro, rg = run_pair(gpu, O, DIMS, 11, "calcp", solver=solver, alfa=alfa, ncorr=ncorr,
                              npcor=npcor)
    print("  iterations", [r["iters"] for r in rg[0]], "continuity", rg[1][-1], flush=True)
    assert all(r["converged"] for r in rg[0])
    check_reports(ro, rg, what)
    check_state(fo, fg, what)


def test_calcp_with_outlet(gpu, O):
    fo, fg, ro, rg = run_pair(gpu, O, DIMS, 12, "calcp", outlet=True, solver=5, alfa=0.92)
    check_reports(ro, rg, "calcp outlet")
    check_state(fo, fg, "calcp outlet")


@pytest.mark.parametrize("kind,opts", [("piso", dict(bdf=1)),
                                       ("piso", dict(cn=1, buoy=1, boussinesq=1)),
                                       ("piso", dict(buoy=1, boussinesq=0, npcor=1)),
                                       ("pimple", dict())])
def test_piso_pimple_bitwise(gpu, O, kind, opts):
    kw = dict(ncorr=2, npcor=2, nsw=300)
    kw.update(opts)
    what = f"{kind} {opts}"
    print(what, flush=True)
    fo, fg, ro, rg = run_pair(gpu, O, DIMS, 21, kind, **kw)
    print("  iterations", [r["iters"] for r in rg[0]], flush=True)
    check_reports(ro, rg, what)
    check_state(fo, fg, what)


@pytest.mark.parametrize("dims", [(12, 10, 9), (18, 14, 12), (33, 21, 17)])
def test_boundary_routines_bitwise(gpu, O, dims):
    fo, fg, _, _ = run_pair(gpu, O, dims, 31, "outbc", outlet=True)
    check_state(fo, fg, f"outbc {dims}")
    fo, fg, _, _ = run_pair(gpu, O, (dims[0], dims[0], dims[2]), 32, "correct_bc")
    check_state(fo, fg, f"correct_bc {dims}")
    for opts in (dict(bdf=1), dict(cn=1, buoy=1, boussinesq=1), dict(buoy=1)):
        fo, fg, _, _ = run_pair(gpu, O, dims, 33, "gethbya", **opts)
        check_state(fo, fg, f"gethbya {opts} {dims}")
