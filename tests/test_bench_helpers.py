"""bench.py's own helpers (no GPU): the torch L1 residual that the learned
guess + mixed-precision comparison scales its stop test with must equal the
oracle's residual sum |su - M x| over the pcgsip solve range."""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


def test_l1_residual_matches_oracle(O):
    import torch
    dims = (14, 11, 9)
    s = O.build_lognormal(dims, 3, 1.0)
    rng = np.random.default_rng(5)
    x = rng.normal(size=s.n)
    mask = np.zeros(s.n, bool)
    mask.reshape(dims[2], dims[0], dims[1])[2:-2, 2:-2, 2:-2] = True
    want = np.abs((s.su - O.spmv(s, 3, dims[2] - 2, x))[mask]).sum()
    arrs = {f: torch.from_numpy(getattr(s, f).copy())
            for f in ("ae", "aw", "an", "as_", "at", "ab", "ap", "su")}
    got = bench._l1_residual(arrs, torch.from_numpy(x), dims)
    np.testing.assert_allclose(got, want, rtol=1e-12)
    # at x = 0 it is sum |su|, the solver's initial residual from zero
    z = bench._l1_residual(arrs, torch.zeros(s.n, dtype=torch.float64), dims)
    np.testing.assert_allclose(z, np.abs(s.su[mask]).sum(), rtol=1e-12)


def test_traffic_only_at_the_profiled_grid():
    """roofline.traffic comes from counters read at one grid size: another
    --n must report null rather than that size's bytes."""
    assert bench._pmc_traffic("sip_fwd_res", 256) is not None
    assert bench._pmc_traffic("sip_fwd_res", 384) is None
    assert bench._pmc_traffic("no_such_kernel", 256) is None


def test_mixed_error_report(O):
    """the mixed solve's error report: zero for the same solution; a constant
    shift of p' moves p' but no face-flux correction; a perturbation of one
    cell moves exactly its six faces, by a_f times the perturbation"""
    import torch
    dims = (14, 11, 9)
    s = O.build_lognormal(dims, 3, 1.0)
    arrs = {f: torch.from_numpy(getattr(s, f).copy()) for f in ("ae", "an", "at")}
    x64 = torch.from_numpy(np.random.default_rng(2).normal(size=s.n))
    e = bench._mixed_error(arrs, x64, x64.clone(), dims)
    assert e["p"]["max"] == 0 and e["face_flux_correction"]["rms"] == 0
    e = bench._mixed_error(arrs, x64, x64 + 0.25, dims)
    assert e["p"]["max"] > 0
    assert e["face_flux_correction"]["max"] < 1e-14
    ni, nj, nk = dims
    x = x64.clone()
    c = (4 - 1) * ni * nj + (5 - 1) * nj + (6 - 1)       # cell (k, i, j) = (4, 5, 6)
    x[c] += 1e-3
    e = bench._mixed_error(arrs, x64, x, dims)
    v = lambda a: a.view(nk, ni, nj)  # noqa: E731
    ref = []
    for a, P, E in ((arrs["ae"], (slice(2, -2), slice(2, -3), slice(2, -2)), (slice(2, -2), slice(3, -2), slice(2, -2))),
                    (arrs["an"], (slice(2, -2), slice(2, -2), slice(2, -3)), (slice(2, -2), slice(2, -2), slice(3, -2))),
                    (arrs["at"], (slice(2, -3), slice(2, -2), slice(2, -2)), (slice(3, -2), slice(2, -2), slice(2, -2)))):
        ref.append((v(a)[P] * (v(x64)[E] - v(x64)[P])).reshape(-1))
    ref = torch.cat(ref)
    k, i, j = 3, 4, 5                                      # 0-based
    moved = [s.ae.reshape(nk, ni, nj)[k, i, j], s.ae.reshape(nk, ni, nj)[k, i - 1, j],
             s.an.reshape(nk, ni, nj)[k, i, j], s.an.reshape(nk, ni, nj)[k, i, j - 1],
             s.at.reshape(nk, ni, nj)[k, i, j], s.at.reshape(nk, ni, nj)[k - 1, i, j]]
    want_max = 1e-3 * max(abs(m) for m in moved) / float(ref.abs().max())
    np.testing.assert_allclose(e["face_flux_correction"]["max"], want_max, rtol=1e-9)
    want_rms = 1e-3 * np.sqrt(sum(m * m for m in moved) / ref.numel()) / float(ref.pow(2).mean().sqrt())
    np.testing.assert_allclose(e["face_flux_correction"]["rms"], want_rms, rtol=1e-9)
