"""cap_courant_noalloc (static device scratch, no allocation) against
cap_courant (context-allocated scratch) through the C ABI: two calls in a
row on one GPU, the second on a smaller grid, each bit-identical to
cap_courant on the same inputs; with a warm call behind it, a call on a
fresh context leaves the device's free memory unchanged."""
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
import make_golden as MG  # noqa: E402

pytestmark = pytest.mark.gpu


def _inputs(dims, seed):
    import torch
    f = MG.fields(dims, seed)
    f1, f2, f3 = f["f123"]
    return [torch.from_numpy(a.copy()).cuda() for a in (f1, f2, f3, f["vol"])]


def _both(gpu, dims, seed, fix):
    import torch
    f1, f2, f3, vol = _inputs(dims, seed)
    with gpu.Context(*dims, device_pointers=True) as ctx:
        su_a = torch.full_like(f1, float("nan"))
        su_b = torch.full_like(f1, float("nan"))
        print(f"courant noalloc {dims} fix {fix}", flush=True)
        b = ctx.courant(f1, f2, f3, vol, su_b, 0.01, 2.0, bool(fix), 0.8, noalloc=True)
        a = ctx.courant(f1, f2, f3, vol, su_a, 0.01, 2.0, bool(fix), 0.8)
    assert a == b, (a, b)
    assert torch.equal(su_a, su_b)
    return a


def test_noalloc_matches_courant_two_calls(gpu):
    """large grid (8192 blocks, the scratch's full extent) then a smaller one
    (183 blocks), each against cap_courant on the same inputs"""
    big = _both(gpu, (132, 132, 132), 5, 1)
    small = _both(gpu, (36, 36, 36), 6, 0)
    assert big[0] > 0 and small[0] > 0


def test_noalloc_allocates_nothing(gpu):
    import torch
    dims = (40, 36, 30)
    f1, f2, f3, vol = _inputs(dims, 7)
    su = torch.zeros_like(f1)
    with gpu.Context(*dims, device_pointers=True) as warm:
        warm.courant(f1, f2, f3, vol, su, 0.01, 0.0, noalloc=True)
    torch.cuda.synchronize()
    with gpu.Context(*dims, device_pointers=True) as ctx:
        torch.cuda.synchronize()
        free0 = torch.cuda.mem_get_info()[0]
        ctx.courant(f1, f2, f3, vol, su, 0.01, 0.0, noalloc=True)
        torch.cuda.synchronize()
        free1 = torch.cuda.mem_get_info()[0]
    assert free1 == free0


def test_noalloc_refuses_host_pointer_context(gpu):
    dims = (12, 10, 9)
    n = dims[0] * dims[1] * dims[2]
    a = [np.ones(n) for _ in range(5)]
    with gpu.Context(*dims) as ctx:
        with pytest.raises(gpu.CapError, match="DEVICE_POINTERS"):
            ctx.courant(*a, 0.01, 0.0, noalloc=True)
