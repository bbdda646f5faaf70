"""The Fortran drop-ins: a host program built on this repository's module
file (tests/fortran/host_modules.f90, which restates the module names and
kinds a host program of the solvers provides) calls `sipsol/pcgsip/iccg/cgstab_sip(fi, ip)`, resolved at link time to
cappuccino-ghostcells-mpi_amd/fortran/cap_dropin.f90 (ISO_C_BINDING ->
libcapamd.so).  Outputs must match the oracle like the C-ABI path does, and
the solver report line must carry the same iteration count."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "tests", "fortran", "_build", "dropin_driver")
B_FIELDS = ("bb", "bs", "bw", "bp", "bn", "be", "bt")


@pytest.mark.skipif(not os.path.exists(DRIVER), reason="drop-in host program not built")
@pytest.mark.parametrize("routine,kind,alfa,nsw", [
    ("sipsol", "general", 0.92, 20), ("pcgsip", "lognormal", 0.0, 2000),
    ("iccg", "lognormal", 0.92, 2000), ("cgstab_sip", "general", 0.92, 400),
    ("cgstab", "lognormal", 0.0, 400), ("pcgjac", "lognormal", 0.0, 4000)])
def test_fortran_dropin(gpu, O, routine, kind, alfa, nsw):
    dims = (20, 37, 23)
    if kind == "general":
        s, fi = O.build_general(dims, 61)
    else:
        s = O.build_lognormal(dims, 61, 1.0)
        fi = np.zeros(s.n)
    rng = np.random.default_rng(61)
    s.res[:] = rng.uniform(-1, 1, s.n)
    for f in B_FIELDS:
        getattr(s, f)[:] = rng.uniform(-0.2, 0.2, s.n)
    arrays = [fi, s.ae, s.aw, s.an, s.as_, s.at, s.ab, s.ap, s.su, s.res] + \
             [getattr(s, f) for f in B_FIELDS]
    outs, resor, log = O.run_ref(routine, dims, [a.copy() for a in arrays], nsw=nsw,
                                 alfa=alfa, driver=DRIVER)
    fn = {"sipsol": O.sipsol, "pcgsip": O.pcgsip, "cgstab_sip": O.cgstab_sip,
          "iccg": O.iccg, "cgstab": O.cgstab, "pcgjac": O.pcg_jacobi}[routine]
    no_alfa = routine in ("iccg", "cgstab", "pcgjac")
    kw = dict(sor=1e-8, nsw=nsw) if no_alfa else dict(alfa=alfa, sor=1e-8, nsw=nsw)
    s0, fi0 = s.copy(), fi.copy()
    rep = fn(s, fi, **kw)
    info = O.parse_log(log)
    print(routine, log.strip(), "| oracle iters", rep["iters"])
    assert resor == pytest.approx(rep["resor"], rel=1e-12)
    # the iteration count of the reference's summation order agrees within
    # what the solve itself moves between summation orders (cgstab: several)
    rep_acc, _, fi_acc = O.accurate_sums(fn, s0, fi0, **kw)
    spread = abs(rep_acc["iters"] - rep["iters"])
    assert abs(info["iters"] - rep["iters"]) <= max(1, spread)
    if routine == "sipsol":
        np.testing.assert_array_equal(outs[0], fi)
    else:
        # the Krylov solvers: bit-identical to the oracle with the GPU's
        # compensated sums; within 1e-10 of the reference's plain sums, or
        # within the solve's own spread between summation orders when larger
        assert info["iters"] == rep_acc["iters"]
        np.testing.assert_array_equal(outs[0], fi_acc)
        scale = np.max(np.abs(fi))
        sens = O.order_sensitivity(fn, s0, fi0, **kw)
        tol = max(1e-10 if info["iters"] == rep["iters"] else 1e-6, 4.0 * sens)
        assert np.max(np.abs(outs[0] - fi)) <= tol * scale
    if routine not in ("iccg", "cgstab", "pcgjac"):
        for f, a in zip(B_FIELDS, outs[2:]):
            np.testing.assert_array_equal(a, getattr(s, f), err_msg=f)


GOLD = os.path.join(ROOT, "tests", "golden")


@pytest.mark.skipif(not os.path.exists(DRIVER), reason="drop-in host program not built")
@pytest.mark.parametrize("scheme", ["cds", "quick", "smart"])
def test_fortran_fluxuvw_dropin(gpu, O, scheme):
    """`call fluxuvw(...)` resolved to the drop-in: bit-identical to the
    reference build's outputs in the golden fixtures, three directions."""
    import sys
    sys.path.insert(0, GOLD)
    import make_golden as MG
    g = np.load(os.path.join(GOLD, f"fluxuvw_{scheme}_12x10x9.npz"))
    dims = tuple(int(x) for x in g["dims"])
    f = MG.fields(dims, int(g["seed"]))
    for d in range(3):
        outs, _, _ = O.run_ref("fluxuvw", dims, MG.fluxuvw_inputs(f),
                               nsw=d + 8 * int(g["flags"]), alfa=float(g["gds"]), driver=DRIVER)
        for q, a in zip(MG.UVW_OUT, outs):
            np.testing.assert_array_equal(a, g[f"{q}_{d}"], err_msg=f"{q} dir {d}")


@pytest.mark.skipif(not os.path.exists(DRIVER), reason="drop-in host program not built")
def test_fortran_fluxmc_dropin(gpu, O):
    """`call fluxmc(inp, ...)` cell by cell in calcp's loop order, resolved
    to the drop-in (the device computes each direction once, at the loop's
    first cell): bit-identical to the reference build's per-cell values."""
    import sys
    sys.path.insert(0, GOLD)
    import make_golden as MG
    g = np.load(os.path.join(GOLD, "fluxmc_12x10x9.npz"))
    dims = tuple(int(x) for x in g["dims"])
    f = MG.fields(dims, int(g["seed"]))
    arrays = [f[q] for q in MG.FIELD_NAMES] + [f["su"], f["sv"], f["sw"]] + f["out0"]
    for d in range(3):
        outs, _, _ = O.run_ref("fluxmc", dims, arrays, nsw=d, alfa=1.0, driver=DRIVER)
        np.testing.assert_array_equal(outs[0], g[f"acfe_{d}"], err_msg=f"dir {d}")


@pytest.mark.skipif(not os.path.exists(DRIVER), reason="drop-in host program not built")
@pytest.mark.parametrize("name", ["ghost_12x10x9", "ghost_20x17x11"])
def test_fortran_ghostcells_dropin(gpu, O, name):
    """`call update_values_at_ghostcells` resolved to the drop-in: bit-
    identical to the reference build (golden fixtures, every boundary type)."""
    import sys
    sys.path.insert(0, GOLD)
    import make_golden as MG
    g = np.load(os.path.join(GOLD, name + ".npz"))
    dims = tuple(int(x) for x in g["dims"])
    var, geo, lb, inl = MG.ghost_inputs(dims, int(g["seed"]))
    outs, _, _ = O.run_ref("ghost", dims, MG.ghost_ref_arrays(dims, var, geo, lb, inl),
                           driver=DRIVER)
    for q, a in zip(MG.GHOST_VARS, outs):
        np.testing.assert_array_equal(a, g[q], err_msg=q)
