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
