"""The drivers' fused three-direction passes (cap_fluxmass3,
cap_piso_assemble_pressure_eq3, cap_fluxuvw3) against the three per-direction
calls they replace (east, north, top), bit for bit, every array they can touch.

Both sides start from the same state, in which every output array holds random
values (so an entry one side writes and the other does not shows up), the
fluxuvw scratch arrays included (their values on entry reach su/sv/sw through
the east call's second loop)."""
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
import flowcase as FC  # noqa: E402

DIMS = [(6, 7, 6), (12, 10, 9), (20, 17, 11), (37, 29, 23), (132, 100, 68)]
OUT = ("ae", "aw", "an", "as", "at", "ab", "f1", "f2", "f3", "be", "bn", "bt", "su", "sv", "sw",
       "bp", "bb")   # "bt" is both PISO's top bp and fluxuvw's v scratch


def _state(dims, seed):
    f = FC.flow_case(dims, seed)
    rng = np.random.default_rng(seed + 1)
    n = f["u"].size
    for q in ("gradu", "gradv", "gradw", "gradp"):
        f[q] = rng.normal(size=3 * n)
    for q in OUT:
        f[q] = rng.normal(size=n)
    for q in ("f1", "f2", "f3"):
        f[q] = 0.1 * rng.normal(size=n)
    f["vis"] = rng.uniform(1e-3, 2e-3, n)
    return f


def _to_gpu(f):
    import torch
    return {q: torch.from_numpy(a.copy()).cuda() for q, a in f.items()}


def _fields(d):
    from cappuccino_amd.host import FIELD_NAMES
    return {q: d[q] for q in FIELD_NAMES}


def _compare(a, b, what):
    import torch
    torch.cuda.synchronize()
    bad = [q for q in OUT if not torch.equal(a[q], b[q])]
    assert not bad, f"{what}: {bad} differ between the fused pass and the three calls"


@pytest.mark.parametrize("dims", DIMS)
def test_fluxmass3_bitwise(gpu, dims):
    f = _state(dims, 5)
    a, b = _to_gpu(f), _to_gpu(f)
    with gpu.Context(*dims, device_pointers=True) as ctx:
        for d, (ae, aw, fc) in enumerate((("ae", "aw", "f1"), ("an", "as", "f2"), ("at", "ab", "f3"))):
            ctx.fluxmass(_fields(a), d, a[ae], a[aw], a[fc])
        ctx.fluxmass3(_fields(b), [b[q] for q in ("ae", "aw", "an", "as", "at", "ab")],
                      [b[q] for q in ("f1", "f2", "f3")])
        _compare(a, b, f"fluxmass {dims}")


@pytest.mark.parametrize("dims", DIMS)
@pytest.mark.parametrize("ipc", [0, 1, 2])
def test_piso_assemble3_bitwise(gpu, dims, ipc):
    f = _state(dims, 7 + ipc)
    a, b = _to_gpu(f), _to_gpu(f)
    with gpu.Context(*dims, device_pointers=True) as ctx:
        for d, (ae, aw, fc, bp) in enumerate((("ae", "aw", "f1", "be"), ("an", "as", "f2", "bn"),
                                              ("at", "ab", "f3", "bt"))):
            ctx.piso_assemble_pressure_eq(_fields(a), d, a[ae], a[aw], a[fc], ipc, a[bp], a["su"])
        ctx.piso_assemble_pressure_eq3(_fields(b), [b[q] for q in ("ae", "aw", "an", "as", "at", "ab")],
                                       [b[q] for q in ("f1", "f2", "f3")], ipc,
                                       [b[q] for q in ("be", "bn", "bt")], b["su"])
        _compare(a, b, f"piso ipc {ipc} {dims}")


@pytest.mark.parametrize("dims", DIMS)
@pytest.mark.parametrize("gds", [1.0, 0.7])
def test_fluxuvw3_bitwise(gpu, dims, gds):
    f = _state(dims, 11)
    a, b = _to_gpu(f), _to_gpu(f)
    with gpu.Context(*dims, device_pointers=True) as ctx:
        for d, (ae, aw, fc) in enumerate((("ae", "aw", "f1"), ("an", "as", "f2"), ("at", "ab", "f3"))):
            ctx.fluxuvw(_fields(a), d, a["vis"], a[ae], a[aw], a[fc], a["su"], a["sv"], a["sw"],
                        a["bp"], a["bt"], a["bb"], 1, gds)
        ctx.fluxuvw3(_fields(b), b["vis"], [b[q] for q in ("ae", "aw", "an", "as", "at", "ab")],
                     [b[q] for q in ("f1", "f2", "f3")], [b[q] for q in ("su", "sv", "sw")],
                     [b[q] for q in ("bp", "bt", "bb")], 1, gds)
        _compare(a, b, f"fluxuvw gds {gds} {dims}")


def test_fused_passes_refuse_what_they_do_not_cover(gpu):
    dims = (12, 10, 9)
    f = _state(dims, 3)
    with gpu.Context(*dims) as ctx:   # host pointers
        with pytest.raises(gpu.CapError):
            ctx.fluxmass3(f, [f[q] for q in ("ae", "aw", "an", "as", "at", "ab")],
                          [f[q] for q in ("f1", "f2", "f3")])
    b = _to_gpu(f)
    with gpu.Context(*dims, device_pointers=True) as ctx:
        with pytest.raises(gpu.CapError):   # a scheme other than CDS alone
            ctx.fluxuvw3(_fields(b), b["vis"], [b[q] for q in ("ae", "aw", "an", "as", "at", "ab")],
                         [b[q] for q in ("f1", "f2", "f3")], [b[q] for q in ("su", "sv", "sw")],
                         [b[q] for q in ("bp", "bt", "bb")], 8, 1.0)
