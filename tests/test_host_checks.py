"""CPU tests of the binding's argument checks: every array handed to the C
ABI must have the dtype, length, contiguity and side (host or device) that
the library reads and writes, and anything else raises before a call.

No GPU is needed: the checks run in Python, before any library call, and the
context object below is never initialised by cap_init."""
import numpy as np
import pytest
import torch


def _ctx(cap, dims=(8, 7, 6), device=False):
    ctx = object.__new__(cap.Context)     # no cap_init: only the checks are exercised
    ctx.dims = dims
    ctx.device_pointers = device
    ctx._h = None
    return ctx


def test_accepts_matching_host_array(cap):
    ctx = _ctx(cap)
    a = np.zeros(8 * 7 * 6)
    assert ctx._f(a) == a.ctypes.data
    assert ctx._f(None) is None


@pytest.mark.parametrize("bad", [
    np.zeros(8 * 7 * 6 - 1),                      # one element short: the C side would overrun it
    np.zeros(8 * 7 * 6 + 5),                      # too long
    np.zeros(8 * 7 * 6, np.float32),              # wrong dtype
    np.zeros((2, 8 * 7 * 6))[:, 0],               # right length would not matter: not contiguous
    np.zeros(2 * 8 * 7 * 6)[::2],                 # strided view of the right length
    [0.0] * (8 * 7 * 6),                          # a list
])
def test_rejects_bad_host_array(cap, bad):
    ctx = _ctx(cap)
    with pytest.raises(cap.CapError):
        ctx._f(bad)


def test_rejects_host_array_in_device_context(cap):
    ctx = _ctx(cap, device=True)
    with pytest.raises(cap.CapError, match="host array"):
        ctx._f(np.zeros(8 * 7 * 6))


def test_rejects_cpu_tensor_in_device_context(cap):
    ctx = _ctx(cap, device=True)
    with pytest.raises(cap.CapError, match="host tensor"):
        ctx._f(torch.zeros(8 * 7 * 6, dtype=torch.float64))


def test_rejects_tensor_of_wrong_length_or_dtype(cap):
    ctx = _ctx(cap, device=False)
    with pytest.raises(cap.CapError, match="elements"):
        ctx._f(torch.zeros(10, dtype=torch.float64))
    with pytest.raises(cap.CapError, match="dtype"):
        ctx._f(torch.zeros(8 * 7 * 6, dtype=torch.float32))
    with pytest.raises(cap.CapError, match="contiguous"):
        ctx._f(torch.zeros(2 * 8 * 7 * 6, dtype=torch.float64)[::2])


def test_coef_struct_checks_every_field(cap):
    ctx = _ctx(cap)
    n = 8 * 7 * 6
    good = {f: np.zeros(n) for f in ("ae", "aw", "an", "as_", "at", "ab", "ap", "su")}
    cap.Coef(**good)._struct(ctx)                 # res and bb..bt may be NULL
    for f in good:
        bad = dict(good)
        bad[f] = np.zeros(n - 1)
        with pytest.raises(cap.CapError, match=f):
            cap.Coef(**bad)._struct(ctx)


def test_gradient_fields_are_three_components(cap):
    ctx = _ctx(cap)
    n = 8 * 7 * 6
    ctx._fields({"gradu": np.zeros(3 * n)})
    with pytest.raises(cap.CapError, match="gradu"):
        ctx._fields({"gradu": np.zeros(n)})
    ctx._flow({"gradp": np.zeros(3 * n), "u": np.zeros(n)})
    with pytest.raises(cap.CapError, match="u"):
        ctx._flow({"u": np.zeros(3 * n)})


def test_boundary_type_arrays_are_int32_of_face_size(cap):
    ni, nj, nk = 8, 7, 6
    ctx = _ctx(cap, (ni, nj, nk))
    sizes = ctx._lb_sizes()
    assert sizes == (nj * nk, nj * nk, ni * nk, ni * nk, ni * nj, ni * nj)
    from importlib import import_module
    host = import_module(type(ctx).__module__)
    host._iptr(np.zeros(sizes[0], np.int32), sizes[0], False)
    with pytest.raises(cap.CapError):
        host._iptr(np.zeros(sizes[0], np.int64), sizes[0], False)
    with pytest.raises(cap.CapError):
        host._iptr(np.zeros(sizes[0] - 1, np.int32), sizes[0], False)
