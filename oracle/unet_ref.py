"""CPU oracle of the learned initial guess -- TEST INFRASTRUCTURE ONLY.

A numpy restatement of the 3-D U-Net of cappuccino-ghostcells-mpi_amd/csrc/
cap_unet.hip (forward and backward), used by the tests to check the GPU's
MFMA convolutions, pooling, upsampling and gradients.  The network has no
counterpart in the reference; this oracle is pinned by finite differences
(tests/test_oracle.py) rather than by reference outputs.

Tensors are channels-last [D, H, W, C]; the parameter vector has, per layer,
the weights [cout][27][cin] (tap = (dz+1)*9 + (dy+1)*3 + (dx+1)) and then the
bias [cout], in the layer order of LAYERS.  With bf16=True the values that the
GPU stores in BF16 (inputs, weights, activations) are rounded to BF16 the same
way (round to nearest even); arithmetic is float64 otherwise.
"""
from __future__ import annotations

import numpy as np

# (cin, cout, level, relu)
LAYERS = [(3, 16, 0, 1), (16, 16, 0, 1), (16, 32, 1, 1), (32, 32, 1, 1), (32, 48, 2, 1),
          (48, 48, 2, 1), (48, 64, 3, 1), (64, 64, 3, 1), (112, 48, 2, 1), (48, 48, 2, 1),
          (80, 32, 1, 1), (32, 32, 1, 1), (48, 16, 0, 1), (16, 16, 0, 1), (16, 1, 0, 0)]


def num_params():
    return sum(co * 27 * ci + co for ci, co, _, _ in LAYERS)


def split(p):
    """Per-layer (W [cout, 27, cin], b [cout]) views of the parameter vector."""
    out, o = [], 0
    for ci, co, _, _ in LAYERS:
        w = p[o:o + co * 27 * ci].reshape(co, 27, ci)
        o += co * 27 * ci
        b = p[o:o + co]
        o += co
        out.append((w, b))
    return out


def bf16(x):
    """Round to the nearest BF16 (ties to even), returned as float64."""
    f = np.asarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


TAPS = [(dz, dy, dx) for dz in (-1, 0, 1) for dy in (-1, 0, 1) for dx in (-1, 0, 1)]


def conv(x, w, b):
    """y[v, co] = sum_tap,ci x[v + off(tap), ci] * w[co, tap, ci] + b[co], zero padding"""
    D, H, W, _ = x.shape
    xp = np.pad(x, ((1, 1), (1, 1), (1, 1), (0, 0)))
    y = np.zeros((D, H, W, w.shape[0]))
    for t, (dz, dy, dx) in enumerate(TAPS):
        xs = xp[1 + dz:1 + dz + D, 1 + dy:1 + dy + H, 1 + dx:1 + dx + W]
        y += xs @ w[:, t, :].T
    return y + b


def conv_back(x, w, dy):
    """(dx, dw, db) of conv for the output gradient dy"""
    D, H, W, ci = x.shape
    xp = np.pad(x, ((1, 1), (1, 1), (1, 1), (0, 0)))
    dxp = np.zeros_like(xp)
    dw = np.zeros_like(w)
    g2 = dy.reshape(-1, dy.shape[-1])
    for t, (dz, dy_, dx_) in enumerate(TAPS):
        sl = (slice(1 + dz, 1 + dz + D), slice(1 + dy_, 1 + dy_ + H), slice(1 + dx_, 1 + dx_ + W))
        dw[:, t, :] = g2.T @ xp[sl].reshape(-1, ci)
        dxp[sl] += dy @ w[:, t, :]
    return dxp[1:-1, 1:-1, 1:-1], dw, dy.reshape(-1, dy.shape[-1]).sum(0)


def pool(x):
    D, H, W, C = x.shape
    return x.reshape(D // 2, 2, H // 2, 2, W // 2, 2, C).mean(axis=(1, 3, 5))


def pool_back(g):
    return np.repeat(np.repeat(np.repeat(g, 2, 0), 2, 1), 2, 2) / 8.0


def up(x):
    return np.repeat(np.repeat(np.repeat(x, 2, 0), 2, 1), 2, 2)


def up_back(g):
    D, H, W, C = g.shape
    return g.reshape(D // 2, 2, H // 2, 2, W // 2, 2, C).sum(axis=(1, 3, 5))


def to_grid(x3, shape):
    """network input: 3 channel-major arrays (x fastest) -> [D, H, W, 3]"""
    nz, ny, nx = shape
    return np.asarray(x3, dtype=np.float64).reshape(3, nz, ny, nx).transpose(1, 2, 3, 0)


def forward(p, x, bf=False):
    """Returns (out [D, H, W], cache) for input x [D, H, W, 3]."""
    r = bf16 if bf else (lambda a: a)
    L = [(r(w), b) for w, b in split(np.asarray(p, dtype=np.float64))]
    relu = lambda a: np.maximum(a, 0.0)  # noqa: E731
    c = {"x": r(x)}
    c["a1"] = r(relu(conv(c["x"], *L[0])))
    c["a2"] = r(relu(conv(c["a1"], *L[1])))
    c["p1"] = r(pool(c["a2"]))
    c["a3"] = r(relu(conv(c["p1"], *L[2])))
    c["a4"] = r(relu(conv(c["a3"], *L[3])))
    c["p2"] = r(pool(c["a4"]))
    c["a5"] = r(relu(conv(c["p2"], *L[4])))
    c["a6"] = r(relu(conv(c["a5"], *L[5])))
    c["p3"] = r(pool(c["a6"]))
    c["a7"] = r(relu(conv(c["p3"], *L[6])))
    c["a8"] = r(relu(conv(c["a7"], *L[7])))
    c["c3"] = np.concatenate([up(c["a8"]), c["a6"]], axis=-1)
    c["a9"] = r(relu(conv(c["c3"], *L[8])))
    c["a10"] = r(relu(conv(c["a9"], *L[9])))
    c["c2"] = np.concatenate([up(c["a10"]), c["a4"]], axis=-1)
    c["a11"] = r(relu(conv(c["c2"], *L[10])))
    c["a12"] = r(relu(conv(c["a11"], *L[11])))
    c["c1"] = np.concatenate([up(c["a12"]), c["a2"]], axis=-1)
    c["a13"] = r(relu(conv(c["c1"], *L[12])))
    c["a14"] = r(relu(conv(c["a13"], *L[13])))
    out = conv(c["a14"], *L[14])[..., 0]
    c["L"] = L
    return out, c


def loss_and_grad(p, x, target, bf=False):
    """loss = mean((out - target)^2) and its gradient with respect to p.
    bf=False: float64 throughout.  bf=True: the forward pass rounds to BF16
    where the GPU stores BF16, and so does the backward pass for the stored
    gradient tensors (loss gradient, input gradients of each convolution,
    combined gradients); weight gradients accumulate in float64."""
    out, c = forward(p, x, bf=bf)
    r = bf16 if bf else (lambda a: a)
    L = c["L"]
    n = out.size
    loss = float(np.mean((out - target) ** 2))
    grads = [None] * len(LAYERS)
    g = r((2.0 * (out - target) / n)[..., None])
    mask = lambda a: (a > 0).astype(np.float64)  # noqa: E731

    def back(l, xin, gout):
        dx, dw, db = conv_back(xin, L[l][0], gout)
        grads[l] = (dw, db)
        return dx

    g14 = r(back(14, c["a14"], g) * mask(c["a14"]))
    g13 = r(back(13, c["a13"], g14) * mask(c["a13"]))
    gc1 = r(back(12, c["c1"], g13))
    g12 = r(up_back(gc1[..., :32]) * mask(c["a12"]))
    g11 = r(back(11, c["a11"], g12) * mask(c["a11"]))
    gc2 = r(back(10, c["c2"], g11))
    g10 = r(up_back(gc2[..., :48]) * mask(c["a10"]))
    g9 = r(back(9, c["a9"], g10) * mask(c["a9"]))
    gc3 = r(back(8, c["c3"], g9))
    g8 = r(up_back(gc3[..., :64]) * mask(c["a8"]))
    g7 = r(back(7, c["a7"], g8) * mask(c["a7"]))
    gp3 = r(back(6, c["p3"], g7))
    g6 = r((gc3[..., 64:] + pool_back(gp3)) * mask(c["a6"]))
    g5 = r(back(5, c["a5"], g6) * mask(c["a5"]))
    gp2 = r(back(4, c["p2"], g5))
    g4 = r((gc2[..., 48:] + pool_back(gp2)) * mask(c["a4"]))
    g3 = r(back(3, c["a3"], g4) * mask(c["a3"]))
    gp1 = r(back(2, c["p1"], g3))
    g2 = r((gc1[..., 32:] + pool_back(gp1)) * mask(c["a2"]))
    g1 = r(back(1, c["a1"], g2) * mask(c["a1"]))
    back(0, c["x"], g1)
    flat = np.concatenate([np.concatenate([dw.ravel(), db]) for dw, db in grads])
    return loss, flat


def init_params(seed):
    """He-normal weights and zero biases (any seeded initialisation; the GPU
    has its own counter-based one)."""
    rng = np.random.default_rng(seed)
    parts = []
    for l, (ci, co, _, _) in enumerate(LAYERS):
        sd = (0.1 if l == len(LAYERS) - 1 else 1.0) * np.sqrt(2.0 / (27 * ci))
        parts.append(rng.normal(0, sd, co * 27 * ci))
        parts.append(np.zeros(co))
    return np.concatenate(parts)
