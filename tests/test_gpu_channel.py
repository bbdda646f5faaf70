"""The channel (periodic x and z, walls at y mapped to the reference's k,
tanh-stretched by 2.0, Re_bulk 5600) at the learned-guess configuration's
size, 128 x 96 x 64, advanced by the library's routines on the device
(cappuccino-ghostcells-mpi_amd/channel.py).  Size-independent properties:

* the geometry producer's outward face vectors: +i, +j, +k;
* the Courant-number step holds Co = 0.5 exactly;
* every p' solve converges, the last corrector to 1e-8 of its start;
* the periodic ghost planes equal the opposite interior planes, and the wall
  ghost values make the face velocity zero;
* the driving pressure gradient keeps the bulk velocity at 1 within 1e-3;
* the continuity error after the last corrector is a small fraction of the
  total face flux.
"""
import importlib.util
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _channel_mod():
    spec = importlib.util.spec_from_file_location(
        "cap_channel", os.path.join(ROOT, "cappuccino-ghostcells-mpi_amd", "channel.py"))
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    return m


@pytest.mark.parametrize("solver", [4, 6])
def test_channel_steps(gpu, solver):
    import torch
    CH = _channel_mod()
    print(f"channel 128x96x64 solver {solver}", flush=True)
    ch = CH.Channel(gpu, torch, 128, 96, 64, solver=solver)
    try:
        f = ch.f
        ni, nj, nk = ch.dims
        sh = (nk, ni, nj)
        assert float(f["ar1z"].min()) >= 0 and float(f["ar2x"].min()) >= 0
        assert float(f["ar3y"].min()) >= 0 and float(f["ar3y"].max()) > 0
        for s in range(4):
            reps = ch.step()
            L = ch.last
            print(f"  step {s}: Co {L['courant']:.6f} dt {L['dt']:.4e} iterations "
                  f"{[r['iters'] for r in reps]} continuity {L['cont'][-1]} bulk "
                  f"{ch.bulk_velocity():.6f}", flush=True)
            assert L["courant"] == pytest.approx(0.5, rel=1e-12)
            assert all(r["converged"] for r in reps)
            assert reps[-1]["resl"] < 1e-8 * reps[-1]["resor"]
        u = f["u"].view(sh).cpu().numpy()
        w = f["w"].view(sh).cpu().numpy()
        fz = f["fz"].view(sh).cpu().numpy()
        # periodic ghosts: i = 2 <- nimm, i = nim <- 3; j = 2 <- njmm, njm <- 3
        np.testing.assert_array_equal(u[2:-2, 1, 2:-2], u[2:-2, ni - 3, 2:-2])
        np.testing.assert_array_equal(u[2:-2, ni - 2, 2:-2], u[2:-2, 2, 2:-2])
        np.testing.assert_array_equal(u[2:-2, 2:-2, 1], u[2:-2, 2:-2, nj - 3])
        np.testing.assert_array_equal(u[2:-2, 2:-2, nj - 2], u[2:-2, 2:-2, 2])
        # no-slip walls: the linearly interpolated face value vanishes
        lam = fz[1, 2:-2, 2:-2]
        face = lam * u[2, 2:-2, 2:-2] + (1 - lam) * u[1, 2:-2, 2:-2]
        assert np.max(np.abs(face)) < 1e-14
        lam = fz[nk - 3, 2:-2, 2:-2]
        face = lam * w[nk - 2, 2:-2, 2:-2] + (1 - lam) * w[nk - 3, 2:-2, 2:-2]
        assert np.max(np.abs(face)) < 1e-14
        assert ch.bulk_velocity() == pytest.approx(1.0, abs=1e-3)
        tot = sum(float(f[q].abs().sum()) for q in ("f1", "f2", "f3"))
        assert ch.last["cont"][-1][0] < 2e-3 * tot
    finally:
        ch.close()
