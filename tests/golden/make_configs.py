"""Fixtures of the 128^3 stretched-grid configuration (C2), from the oracle.

The oracle at 128^3 takes minutes for the three Krylov solves, too long for a
GPU test to repeat; this script runs it once in the build container and
writes tests/golden/c2_128.json: SHA-256 digests of the assembled p' system,
of each solver's p' with the compensated sums (the GPU's arithmetic: the GPU
must match them bit for bit) with its iteration count and residuals, the
largest difference between that p' and the p' of the reference's plain
running sums (so a bit-identical GPU result is within it of the reference's
summation order) next to how far the solve moves between two summation
orders of the reference's own loop (per-plane against running sums), and
digests of GRAD_GAUSS and fluxuvw (CDS, gds = 1) on the
same field.  usage: python tests/golden/make_configs.py [n]
"""
import hashlib
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.dirname(HERE))
import oracle as O  # noqa: E402
import configcase as CC  # noqa: E402

SOLVERS = (("pcgsip", dict(alfa=0.0)), ("iccg", dict()), ("cgstab_sip", dict(alfa=0.92)))
UVW_OUT = ("acfe", "acfw", "su", "sv", "sw", "bp", "bt", "bb")


def digest(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype="<f8").tobytes()).hexdigest()


def grad_and_fluxuvw(dims, f):
    """GRAD_GAUSS of u, v, w into gradu.. (3, n), then fluxuvw (CDS, gds 1)
    for the three directions into zeroed outputs; returns the digests."""
    g = dict(f)
    out = {}
    geom = {q: f[q] for q in O.GEOM_FIELDS}
    for q in ("u", "v", "w"):
        dx, dy, dz = O.grad_gauss(dims, f[q], geom)
        g["grad" + q] = np.ascontiguousarray(np.stack([dx, dy, dz], axis=1).ravel())
        out["grad" + q] = digest(g["grad" + q])
    n = f["u"].size
    for d in range(3):
        o = {q: np.zeros(n) for q in UVW_OUT}
        O.fluxuvw(dims, g, f["vis"], d, o, 1, 1.0)
        for q in UVW_OUT:
            out[f"fluxuvw_{q}_{d}"] = digest(o[q])
    return out


def main():
    n = int(sys.argv[1]) if len(sys.argv) > 1 else 128
    t0 = time.time()
    dims, f, lb, inl = CC.state(O, "stretched", n, 2)
    rec = dict(n=n, seed=2, dims=list(dims), state={q: digest(f[q]) for q in
                                                    ("u", "v", "w", "f1", "f2", "f3", "fx", "ar1x")})
    rec["stencils"] = grad_and_fluxuvw(dims, f)
    CC.momentum_and_assembly(O, dims, f)
    s = CC.p_system(O, dims, f)
    rec["system"] = {q: digest(getattr(s, "as_" if q == "as" else q)) for q in CC.SYSTEM}
    print(f"state + assembly {time.time() - t0:.1f} s", flush=True)
    rec["solvers"] = {}
    for name, kw in SOLVERS:
        fn = getattr(O, name)
        t = time.time()
        rep_acc, _, fi_acc = O.accurate_sums(fn, s.copy(), np.zeros(s.n), sor=1e-8, nsw=4000, **kw)
        s2, fi_plain = s.copy(), np.zeros(s.n)
        rep_plain = fn(s2, fi_plain, sor=1e-8, nsw=4000, **kw)
        diff = float(np.max(np.abs(fi_acc - fi_plain)) / np.max(np.abs(fi_plain)))
        # the solve's own sensitivity to the summation order: per-plane sums
        # against the plain running sums (both orders the reference could have)
        s3, fi_plane = s.copy(), np.zeros(s.n)
        with O.plane_sums(1):
            fn(s3, fi_plane, sor=1e-8, nsw=4000, **kw)
        sens = float(np.max(np.abs(fi_plane - fi_plain)) / np.max(np.abs(fi_plain)))
        rec["solvers"][name] = dict(kw, iters=rep_acc["iters"], resor=rep_acc["resor"],
                                    resl=rep_acc["resl"], converged=rep_acc["converged"],
                                    fi=digest(fi_acc), fi_l2=float(np.linalg.norm(fi_acc)),
                                    iters_plain=rep_plain["iters"], max_rel_vs_plain=diff,
                                    order_sensitivity=sens)
        print(f"{name}: {rep_acc['iters']} iterations (plain {rep_plain['iters']}), "
              f"vs plain sums {diff:.2e} (order sensitivity {sens:.2e}), "
              f"{time.time() - t:.1f} s", flush=True)
    with open(os.path.join(HERE, f"c2_{n}.json"), "w") as fh:
        json.dump(rec, fh, indent=1)
    print(f"c2_{n}.json written ({time.time() - t0:.0f} s)")


if __name__ == "__main__":
    main()
