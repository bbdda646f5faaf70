"""Independent pin of the U-Net oracle (oracle/unet_ref.py): the same network
written with torch's own conv3d, avg_pool3d, nearest-neighbour interpolate and
channel concatenation, in float64 on the CPU, must give the oracle's forward
pass, and torch's autograd the oracle's hand-written backward pass.

A transposed tap order, a mirrored kernel, a wrong padding, pooling or
upsampling rule, or a skip connection concatenated in the other order moves
these by O(1), far beyond the 1e-12 bound."""
import os
import sys

import numpy as np
import pytest
import torch
import torch.nn.functional as F

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import unet_ref as U  # noqa: E402


def torch_weights(p):
    """unet_ref's per-layer w[cout, tap, cin] (tap = (dz+1)*9 + (dy+1)*3 +
    (dx+1), input read at v + (dz, dy, dx)) as torch conv3d weights
    [cout, cin, kd, kh, kw] (cross-correlation, kd = dz + 1)"""
    out = []
    for w, b in U.split(np.asarray(p, dtype=np.float64)):
        co, _, ci = w.shape
        W = torch.from_numpy(np.ascontiguousarray(w.reshape(co, 3, 3, 3, ci).transpose(0, 4, 1, 2, 3)))
        out.append((W, torch.from_numpy(b.copy())))
    return out


def torch_forward(L, x):
    """x: [1, 3, D, H, W]; the U-Net of unet_ref.forward"""
    def cr(h, l):
        return F.relu(F.conv3d(h, L[l][0], L[l][1], padding=1))

    def up(h):
        return F.interpolate(h, scale_factor=2, mode="nearest")
    a2 = cr(cr(x, 0), 1)
    a4 = cr(cr(F.avg_pool3d(a2, 2), 2), 3)
    a6 = cr(cr(F.avg_pool3d(a4, 2), 4), 5)
    a8 = cr(cr(F.avg_pool3d(a6, 2), 6), 7)
    a10 = cr(cr(torch.cat([up(a8), a6], dim=1), 8), 9)
    a12 = cr(cr(torch.cat([up(a10), a4], dim=1), 10), 11)
    a14 = cr(cr(torch.cat([up(a12), a2], dim=1), 12), 13)
    return F.conv3d(a14, L[14][0], L[14][1], padding=1)[0, 0]


def _case(seed, shape):
    rng = np.random.default_rng(seed)
    p = U.init_params(seed)
    # non-zero biases and a last layer of ordinary size exercise every term
    for l, (w, b) in enumerate(U.split(p)):
        b[:] = rng.normal(0, 0.1, b.shape)
    x = rng.normal(size=shape + (3,))
    return p, x


@pytest.mark.parametrize("shape", [(8, 8, 8), (16, 8, 8), (8, 16, 24)])
def test_forward_matches_torch(shape):
    p, x = _case(7, shape)
    out, _ = U.forward(p, x)
    xt = torch.from_numpy(np.ascontiguousarray(x.transpose(3, 0, 1, 2)))[None]
    with torch.no_grad():
        ref = torch_forward(torch_weights(p), xt).numpy()
    scale = np.max(np.abs(ref))
    assert scale > 0
    np.testing.assert_allclose(out, ref, rtol=0, atol=1e-12 * scale)


def test_gradient_matches_torch_autograd():
    shape = (8, 8, 16)
    p, x = _case(11, shape)
    target = np.random.default_rng(12).normal(size=shape)
    loss, g = U.loss_and_grad(p, x, target)
    L = torch_weights(p)
    params = []
    for W, b in L:
        W.requires_grad_(True)
        b.requires_grad_(True)
        params += [W, b]
    xt = torch.from_numpy(np.ascontiguousarray(x.transpose(3, 0, 1, 2)))[None]
    out = torch_forward(L, xt)
    lt = torch.mean((out - torch.from_numpy(target)) ** 2)
    lt.backward()
    assert loss == pytest.approx(lt.item(), rel=1e-12)
    # back to unet_ref's layout: per layer dW [cout, tap, cin] then db
    parts = []
    for W, b in L:
        co, ci = W.shape[0], W.shape[1]
        parts.append(W.grad.numpy().transpose(0, 2, 3, 4, 1).reshape(co, 27, ci).ravel())
        parts.append(b.grad.numpy())
    ref = np.concatenate(parts)
    np.testing.assert_allclose(g, ref, rtol=0, atol=1e-10 * np.max(np.abs(ref)))
