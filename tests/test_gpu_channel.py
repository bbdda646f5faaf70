"""The channel (periodic x and z, walls at y mapped to the reference's k,
tanh-stretched by 2.0, Re_bulk 5600) at the learned-guess configuration's
size, 128 x 96 x 64, and at the full-step configuration's, 384 x 256 x 192
(two steps with the mixed-precision p' solve), advanced by the library's
routines on the device (cappuccino-ghostcells-mpi_amd/channel.py).
Size-independent properties:

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


@pytest.mark.parametrize("size,solver,steps", [((128, 96, 64), 4, 4), ((128, 96, 64), 6, 4),
                                               ((384, 256, 192), 6, 2)])
def test_channel_steps(gpu, size, solver, steps):
    import torch
    CH = _channel_mod()
    print(f"channel {size} solver {solver}", flush=True)
    ch = CH.Channel(gpu, torch, *size, solver=solver)
    try:
        f = ch.f
        ni, nj, nk = ch.dims
        sh = (nk, ni, nj)
        assert float(f["ar1z"].min()) >= 0 and float(f["ar2x"].min()) >= 0
        assert float(f["ar3y"].min()) >= 0 and float(f["ar3y"].max()) > 0
        for s in range(steps):
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


def test_warm_and_learned_starts(gpu):
    """The last corrector's p solve (PISO_multiple_correction.f90:141) of
    the 128 x 96 x 64 channel from three starts, on the same state: the
    reference's (the previous corrector's p), the previous step's p (warm)
    and the learned guess, the U-Net trained on pairs of the channel's own
    earlier steps.  Each start is accepted only when it lowers the L1
    residual; the solve then stops at the absolute residual the reference's
    start stops at.  The warm start saves iterations; the network's guess
    lowers the start's residual below the warm start's."""
    import torch
    CH = _channel_mod()
    ch = CH.Channel(gpu, torch, 128, 96, 64, solver=4)
    try:
        for _ in range(10):
            ch.step(timed=False)
        X, T = [], []
        for _ in range(300):
            prev = ch.f["p"].clone()
            ch.step(timed=False)
            x, t = ch.pressure_pair(prev)
            X.append(x)
            T.append(t)
        ni, nj, nk = ch.dims
        net = gpu.UNet(ch.ctx, nj - 4, ni - 4, nk - 4, seed=1)
        g = torch.Generator().manual_seed(0)
        for ep in range(8):
            lr = 5e-4 if ep < 6 else 1e-4
            loss = np.mean([net.train_step(X[q], T[q], lr)
                            for q in torch.randperm(len(X), generator=g).tolist()])
        print(f"training loss, last epoch {loss:.4f}", flush=True)
        del X, T
        its = {"reference": [], "warm": [], "learned": []}
        for s in range(3):
            snap = ch.snapshot()
            l1s = {}
            for kind in its:
                ch.restore(snap)
                reps = ch.step(timed=False, net=net if kind == "learned" else None,
                               warm=kind == "warm")
                r = reps[-1]
                assert r["converged"]
                its[kind].append(r["iters"])
                if kind == "reference":
                    l1_ref = r["resor"]
                    resl_ref = r["resl"]
                else:
                    assert ch.last["accepted"]
                    ref, start = ch.last["l1"]
                    assert ref == pytest.approx(l1_ref, rel=1e-9)
                    l1s[kind] = start
                    # the same absolute stopping point as the reference's start
                    assert r["resl"] < 1e-8 * l1_ref * 1.01
                    assert r["resor"] == pytest.approx(start, rel=1e-9)
            print(f"step {s}: iterations {[its[k][-1] for k in its]}, L1 start / reference "
                  f"warm {l1s['warm'] / l1_ref:.4f} learned {l1s['learned'] / l1_ref:.4f} "
                  f"(reference stop {resl_ref:.3e})", flush=True)
            assert l1s["warm"] < 0.2 * l1_ref
            assert l1s["learned"] < l1s["warm"]
            ch.restore(snap)
            ch.step(timed=False)
        assert np.mean(its["warm"]) < np.mean(its["reference"])
        assert np.mean(its["learned"]) <= np.mean(its["warm"]) + 1
        net.close()
    finally:
        ch.close()


def test_fp32_momentum_fluxes(gpu):
    """The channel with fluxuvw's face arithmetic in FP32 (CAP_UVW_FP32)
    against FP64 from the same start: three steps later u, v, w agree to
    within 1e-5 of the bulk velocity, and the p' solves converge alike."""
    import torch
    CH = _channel_mod()
    out = []
    for fp32 in (False, True):
        ch = CH.Channel(gpu, torch, 128, 96, 64, solver=4, fp32_flux=fp32)
        try:
            its = [[r["iters"] for r in ch.step()] for _ in range(3)]
            out.append(({q: ch.f[q].clone() for q in ("u", "v", "w")}, its, ch.phase_ms()))
        finally:
            ch.close()
    (a, ia, pa), (b, ib, pb) = out
    diff = max(float((a[q] - b[q]).abs().max()) for q in ("u", "v", "w"))
    print(f"max |velocity difference| {diff:.3e}; iterations FP64 {ia} FP32 {ib}; calcuvw ms "
          f"{pa['calcuvw']:.2f} / {pb['calcuvw']:.2f}", flush=True)
    assert diff < 1e-5
    for x, y in zip(ia, ib):
        assert abs(x[-1] - y[-1]) <= 3
