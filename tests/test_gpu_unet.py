"""The learned initial guess on the GPU: MFMA convolutions, forward and
backward against the CPU oracle (oracle/unet_ref.py, pinned by finite
differences), training on solver-generated pairs, and what the trained guess
gains the unchanged PCG-SIP solve.

Tolerances: the GPU stores activations, weights and gradient tensors in BF16
(8 significant bits) and accumulates in FP32; the forward output and the
gradients are compared with the oracle run with the same BF16 rounding points
(forward: relative L2 error < 2e-2; gradients: 1 - cosine similarity per layer
within 4x the oracle's own change under a weight perturbation far below BF16
resolution -- rounding flips make a BF16 network's gradients move by a few
percent under such perturbations); the distance to the float64 oracle is
printed beside them."""
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import unet_ref as U  # noqa: E402


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.fixture(scope="module")
def small(gpu):
    shape = (16, 16, 24)          # nz, ny, nx
    nz, ny, nx = shape
    ctx = gpu.Context(ny + 4, nx + 4, nz + 4)
    net = gpu.UNet(ctx, nx, ny, nz, seed=3)
    yield ctx, net, shape
    net.close()
    ctx.close()


def test_parameter_layout(small):
    ctx, net, shape = small
    assert net.num_params == U.num_params()
    p = U.init_params(11).astype(np.float32)
    net.set_params(p)
    np.testing.assert_array_equal(net.get_params(), p)


def test_forward_matches_oracle(small):
    ctx, net, shape = small
    rng = np.random.default_rng(1)
    p = U.init_params(7)
    for w, b in U.split(p):
        b[:] = rng.normal(0, 0.05, b.size)
    net.set_params(p.astype(np.float32))
    x3 = rng.normal(size=(3,) + shape).astype(np.float32)
    out = np.zeros(int(np.prod(shape)), np.float32)
    net.forward(x3.ravel(), out)
    ref, _ = U.forward(p.astype(np.float32).astype(np.float64), U.to_grid(x3, shape), bf=True)
    err = rel_l2(out.reshape(shape), ref)
    print(f"forward rel L2 vs BF16 oracle {err:.2e}")
    assert err < 2e-2


def test_gradients_match_oracle(small):
    ctx, net, shape = small
    _check_gradients(net, shape, 2)


def test_gradients_match_oracle_many_bricks(gpu):
    """A grid whose full-resolution level has 640 weight-gradient bricks of
    16 x 4 x 4 voxels, more than the blocks of one launch (so a block sums
    several bricks), and whose second level ends in a partial brick."""
    shape = (32, 64, 80)          # nz, ny, nx
    nz, ny, nx = shape
    with gpu.Context(ny + 4, nx + 4, nz + 4) as ctx:
        net = gpu.UNet(ctx, nx, ny, nz, seed=5)
        try:
            _check_gradients(net, shape, 4)
        finally:
            net.close()


def _check_gradients(net, shape, seed):
    rng = np.random.default_rng(seed)
    p = U.init_params(8)
    for w, b in U.split(p):
        b[:] = rng.normal(0, 0.05, b.size)
    p = p.astype(np.float32).astype(np.float64)
    net.set_params(p.astype(np.float32))
    x3 = rng.normal(size=(3,) + shape).astype(np.float32)
    t = rng.normal(size=shape).astype(np.float32)
    g = np.zeros(U.num_params(), np.float32)
    loss = net.gradients(x3.ravel(), t.ravel(), g)
    xg = U.to_grid(x3, shape)
    t64 = t.astype(np.float64)
    loss_bf, g_bf = U.loss_and_grad(p, xg, t64, bf=True)
    loss_64, g_64 = U.loss_and_grad(p, xg, t64)
    # the BF16 network's own sensitivity: the oracle's gradients after a
    # change of the weights far below BF16 resolution (1e-6 relative)
    pp = p * (1 + 1e-6 * rng.normal(size=p.size))
    _, g_pp = U.loss_and_grad(pp, xg, t64, bf=True)
    assert loss == pytest.approx(loss_bf, rel=1e-2)
    cos = lambda u, v: float(u @ v / (np.linalg.norm(u) * np.linalg.norm(v)))  # noqa: E731
    o = 0
    for l, (ci, co, _, _) in enumerate(U.LAYERS):
        n = co * 27 * ci + co
        a = g[o:o + n].astype(np.float64)
        gap, sens = 1 - cos(a, g_bf[o:o + n]), 1 - cos(g_pp[o:o + n], g_bf[o:o + n])
        print(f"layer {l:2d} ({ci:3d}->{co:2d}): 1 - cosine vs BF16 oracle {gap:.2e} "
              f"(oracle's own sensitivity {sens:.2e}); cosine vs float64 oracle "
              f"{cos(a, g_64[o:o + n]):.5f}")
        assert gap <= 4 * sens + 1e-4, l
        o += n


def test_gradients_are_the_same_every_call(small):
    """The weight gradients are summed without atomics (one partial per
    block, added in block order), so the same inputs give the same bits."""
    ctx, net, shape = small
    rng = np.random.default_rng(6)
    net.set_params(U.init_params(10).astype(np.float32))
    x3 = rng.normal(size=(3,) + shape).astype(np.float32).ravel()
    t = rng.normal(size=shape).astype(np.float32).ravel()
    g1 = np.zeros(U.num_params(), np.float32)
    g2 = np.zeros(U.num_params(), np.float32)
    l1 = net.gradients(x3, t, g1)
    l2 = net.gradients(x3, t, g2)
    assert l1 == l2
    assert np.array_equal(g1, g2)
    # not a trivial pass: most weights get a gradient (taps that see only
    # the zero padding of the 2 x 2 x 3 coarsest level, and dead ReLUs, do not)
    assert np.count_nonzero(g1) > 0.5 * g1.size


def test_training_lowers_the_loss(small):
    ctx, net, shape = small
    rng = np.random.default_rng(4)
    net.set_params(U.init_params(9).astype(np.float32))
    x3 = rng.normal(size=(3,) + shape).astype(np.float32)
    # a smooth target the network can fit
    z, y, x = np.meshgrid(*[np.linspace(0, 1, s) for s in shape], indexing="ij")
    t = (np.sin(3 * x) * np.cos(2 * y) + z).astype(np.float32)
    losses = [net.train_step(x3.ravel(), t.ravel(), 2e-3) for _ in range(60)]
    print(f"loss {losses[0]:.4e} -> {losses[-1]:.4e}")
    assert losses[-1] < 0.2 * losses[0]


def test_learned_guess_saves_iterations(gpu, O):
    """calcp's pressure solve (pp = 0 before each solve,
    calcp-multiple_correction_SIMPLE.f90:86) for a slowly varying source on
    one log-normal operator.  The network is trained on the (su, ap, previous
    p') -> p' pairs that the GPU solver itself produces, then held-out steps
    are solved from zero and from the guess to the same absolute L1 residual:
    the guess is accepted, lowers the initial residual and saves iterations."""
    import torch
    n = 32
    dims = (n + 4, n + 4, n + 4)
    npad = dims[0] * dims[1] * dims[2]
    s = O.build_lognormal(dims, 5, 1.0)
    ctx = gpu.Context(*dims, device_pointers=True)
    dev = {f: torch.from_numpy(getattr(s, f).copy()).cuda() for f in gpu.COEF_FIELDS}
    coef = gpu.Coef(**dev)
    kk, ii, jj = np.meshgrid(*[np.arange(d) for d in (dims[2], dims[0], dims[1])], indexing="ij")
    inner = (kk >= 2) & (kk < dims[2] - 2) & (ii >= 2) & (ii < dims[0] - 2) & (jj >= 2) & (jj < dims[1] - 2)
    pin = (kk == 2) & (ii == 2) & (jj == 2)
    interior = lambda a: a.reshape(dims[2], dims[0], dims[1])[2:-2, 2:-2, 2:-2].ravel()  # noqa: E731

    def source(step):
        ph = 0.05 * step
        f = (np.sin(2 * np.pi * (ii / n + ph)) * np.cos(2 * np.pi * (jj / n - 0.5 * ph)) +
             0.5 * np.sin(2 * np.pi * (kk / n + 0.3 * ph)))
        f = np.where(inner & ~pin, f, 0.0)
        f[inner & ~pin] -= f[inner & ~pin].mean()
        return f.ravel()

    sols, srcs = [], []
    for step in range(48):
        srcs.append(source(step))
        dev["su"].copy_(torch.from_numpy(srcs[-1]))
        fi = torch.zeros(npad, dtype=torch.float64, device="cuda")
        rep = ctx.pcgsip(fi, coef, alfa=0.0, sor=1e-8, nsw=2000)
        assert rep["converged"]
        sols.append(fi.cpu().numpy())
    net = gpu.UNet(ctx, n, n, n, seed=2)
    # training pairs, normalised the way cap_unet_guess normalises its
    # channels (r = the residual of the previous p' for the new source); the
    # target is the change from the previous p' in units of its own RMS --
    # the guess fits the weights of previous p' and of the network's output
    ap = interior(s.ap)
    X, T = [], []
    for step in range(1, 40):
        s.su[:] = srcs[step]
        r = interior(s.su - O.spmv(s, 3, dims[2] - 2, sols[step - 1]))
        prev, sol = interior(sols[step - 1]), interior(sols[step])
        X.append(np.stack([r / np.sqrt(np.mean(r ** 2)), ap / ap.mean(),
                           prev / np.sqrt(np.mean(prev ** 2))]).astype(np.float32))
        T.append(((sol - prev) / np.sqrt(np.mean((sol - prev) ** 2))).astype(np.float32))
    Xd = [torch.from_numpy(x.ravel()).cuda() for x in X]
    Td = [torch.from_numpy(t).cuda() for t in T]
    losses = []
    for it in range(40 * len(Xd)):
        q = it % len(Xd)
        losses.append(net.train_step(Xd[q], Td[q], 5e-4 if it < 30 * len(Xd) else 1e-4))
    first, last = np.mean(losses[:len(Xd)]), np.mean(losses[-len(Xd):])
    print(f"training loss (mean over an epoch) {first:.3e} -> {last:.3e}")
    assert last < 0.5 * first
    base_it, guess_it, warm_it = [], [], []
    for step in range(40, 48):
        dev["su"].copy_(torch.from_numpy(srcs[step]))
        prev = torch.from_numpy(sols[step - 1]).cuda()
        f0 = torch.zeros(npad, dtype=torch.float64, device="cuda")
        r0 = ctx.pcgsip(f0, coef, alfa=0.0, sor=1e-8, nsw=2000)
        f1 = torch.zeros(npad, dtype=torch.float64, device="cuda")
        r1 = net.pcgsip(f1, coef, prev=prev, alfa=0.0, sor=1e-8, nsw=2000)
        # for comparison: the warm start from the previous p' alone, to the
        # same absolute residual
        s.su[:] = srcs[step]
        l1_prev = np.abs(interior(s.su - O.spmv(s, 3, dims[2] - 2, sols[step - 1]))).sum()
        f2 = prev.clone()
        r2 = ctx.pcgsip(f2, coef, alfa=0.0, sor=1e-8 * r0["resor"] / l1_prev, nsw=2000)
        base_it.append(r0["iters"])
        guess_it.append(r1["iters"])
        warm_it.append(r2["iters"])
        print(f"step {step}: from zero {r0['iters']} iterations (L1 {r0['resor']:.3e} -> "
              f"{r0['resl']:.3e}), from the guess {r1['iters']} (L1 {r1['resor']:.3e} -> "
              f"{r1['resl']:.3e}, accepted {r1['accepted']}), from previous p' {r2['iters']} "
              f"(L1 {l1_prev:.3e})")
        assert r1["converged"] and r1["accepted"]
        assert r1["resor"] < r0["resor"]
        # the same absolute stopping point (1 % slack for the last iteration)
        assert r1["resl"] < 1e-8 * r0["resor"] * 1.01
    print(f"mean iterations: from zero {np.mean(base_it):.1f}, from the learned guess "
          f"{np.mean(guess_it):.1f}, from previous p' alone {np.mean(warm_it):.1f}")
    assert np.mean(guess_it) < np.mean(base_it)

    # learned + mixed: the guess as the start of the mixed-precision solve,
    # which stops at the absolute L1 residual of the FP64 solve from zero
    # (its test is relative to its own start, so sor is scaled)
    step = 40
    dev["su"].copy_(torch.from_numpy(srcs[step]))
    s.su[:] = srcs[step]
    f3 = torch.zeros(npad, dtype=torch.float64, device="cuda")
    net.guess(f3, coef, prev=torch.from_numpy(sols[step - 1]).cuda())
    l1 = lambda x: np.abs(interior(s.su - O.spmv(s, 3, dims[2] - 2, x))).sum()  # noqa: E731
    r0_l1, rg_l1 = np.abs(interior(s.su)).sum(), l1(f3.cpu().numpy())
    assert rg_l1 < r0_l1
    rm = ctx.pcgsip_mixed(f3, coef, alfa=0.0, sor=1e-8 * r0_l1 / rg_l1, nsw=4000, inner_rtol=3e-4)
    fm = f3.cpu().numpy()
    err = np.max(np.abs(interior(fm - sols[step]))) / np.max(np.abs(interior(sols[step])))
    print(f"learned + mixed, step {step}: guess L1 {rg_l1 / r0_l1:.3f} of r0, {rm['iters']} inner "
          f"iterations, true L1 {l1(fm) / r0_l1:.2e} of r0, max err vs fp64 {err:.2e}")
    assert rm["converged"]
    assert l1(fm) < 1e-8 * r0_l1
    assert err < 1e-5
    net.close()
    ctx.close()
