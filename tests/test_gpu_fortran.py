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
    ("iccg", "lognormal", 0.92, 2000), ("cgstab_sip", "general", 0.92, 400)])
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
          "iccg": O.iccg}[routine]
    kw = dict(sor=1e-8, nsw=nsw) if routine == "iccg" else dict(alfa=alfa, sor=1e-8, nsw=nsw)
    s0, fi0 = s.copy(), fi.copy()
    rep = fn(s, fi, **kw)
    info = O.parse_log(log)
    print(routine, log.strip(), "| oracle iters", rep["iters"])
    assert resor == pytest.approx(rep["resor"], rel=1e-12)
    assert abs(info["iters"] - rep["iters"]) <= 1
    if routine == "sipsol":
        np.testing.assert_array_equal(outs[0], fi)
    else:
        # the Krylov solvers: bit-identical to the oracle with the GPU's
        # compensated sums, within 1e-10 of the reference's plain sums
        rep_acc, _, fi_acc = O.accurate_sums(fn, s0, fi0, **kw)
        assert info["iters"] == rep_acc["iters"]
        np.testing.assert_array_equal(outs[0], fi_acc)
        scale = np.max(np.abs(fi))
        tol = 1e-10 if info["iters"] == rep["iters"] else 1e-6
        assert np.max(np.abs(outs[0] - fi)) <= tol * scale
    if routine != "iccg":
        for f, a in zip(B_FIELDS, outs[2:]):
            np.testing.assert_array_equal(a, getattr(s, f), err_msg=f)
