"""The seeded problem generator (cap_build_lognormal) runs on the device and
must reproduce the oracle's orc_build_lognormal bit for bit: on small grids
against the oracle directly (device- and host-pointer contexts), and at
128^3 and 256^3 (the benchmark's system) against SHA-256 digests of the
oracle's arrays committed in tests/golden/lognormal_digests.json."""
import hashlib
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
FIELDS = ("ae", "aw", "an", "as_", "at", "ab", "ap", "su")


def _digest(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype="<f8").tobytes()).hexdigest()


@pytest.mark.parametrize("dims,seed,sigma", [((12, 10, 9), 3, 1.0), ((20, 17, 11), 4, 0.6),
                                             ((7, 7, 7), 5, 1.0), ((36, 36, 36), 2024, 1.0)])
@pytest.mark.parametrize("device", [False, True])
def test_generator_bitwise(gpu, O, dims, seed, sigma, device):
    import torch
    print(f"generator {dims} seed {seed} sigma {sigma} device {device}", flush=True)
    ref = O.build_lognormal(dims, seed, sigma)
    n = ref.n
    if device:
        arr = {f: torch.full((n,), float("nan"), dtype=torch.float64, device="cuda") for f in FIELDS}
    else:
        arr = {f: np.full(n, np.nan) for f in FIELDS}
    coef = gpu.Coef(**arr)
    with gpu.Context(*dims, device_pointers=device) as ctx:
        ctx.build_lognormal(coef, seed, sigma)
    for f in FIELDS:
        got = arr[f].cpu().numpy() if device else arr[f]
        np.testing.assert_array_equal(got, getattr(ref, f), err_msg=f)


def test_generator_large_digests(gpu):
    import torch
    with open(os.path.join(HERE, "golden", "lognormal_digests.json")) as fh:
        cases = json.load(fh)
    for case in cases:
        dims = tuple(case["dims"])
        print(f"generator digests {dims} seed {case['seed']}", flush=True)
        n = dims[0] * dims[1] * dims[2]
        arr = {f: torch.empty(n, dtype=torch.float64, device="cuda") for f in FIELDS}
        with gpu.Context(*dims, device_pointers=True) as ctx:
            ctx.build_lognormal(gpu.Coef(**arr), case["seed"], case["sigma"])
        for f in FIELDS:
            assert _digest(arr[f].cpu().numpy()) == case["sha256"][f], (dims, f)
        del arr
        torch.cuda.empty_cache()
