#!/usr/bin/env python3
"""Benchmark: ms per p' solve to 1e-8 L1 relative residual at 256^3.

Workload: the 256^3-cell log-normal (sigma = 1) Neumann p' operator on a
uniform grid with the reference-pressure cell pinned, RHS N(0,1), fi0 = 0
(generated by the library's seeded generator, resident in HBM before the
timed region).  One step = one complete p' solve from fi = 0 to
sum|r| / sum|r0| < 1e-8 (the reference's stop test, pcg-sip.f90:186).

The headline `value` is the plain FP64 drop-in, PCG with the SIP
preconditioner (alfa = 0: with the reference's alfa = 0.92 CG does not reach
1e-8 on this symmetric operator).  The other FP64 solvers and the
mixed-precision path are timed beside it under "solvers".

roofline: the dominant kernel of one PCG-SIP iteration, timed live on the
library's stream with HIP events; algorithmic bytes per launch = array
streams x 8 B x interior cells.  cpu_baseline: the single-threaded C
restatement of pcgsip (the oracle) on the host, on a bounded number of
iterations of the same 256^3 system, scaled to the GPU's iteration count.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("ms per p' solve to 1e-8 rel. residual at 256^3 (plain FP64 vs learned+mixed); %HBM")
HBM_PEAK_GBS = 8000.0
# algorithmic FP64 bytes per interior cell of the PCG-SIP kernels, in the order
# cap_time_kernels reports them.  Every iteration after the first splits the
# previous one's update: sip_fwd_res (reads res, the packed factors, A pk;
# writes z, res) updates the residual and sums its L1 norm before the forward
# sweep, and pk_update (reads z, pk, fi; writes pk, fi) updates fi first.
# sip_fwd and update_l1 run once per solve.  spmv_dot reads p, ap, ae, an, at
# and writes z (48 B): on a matrix that is symmetric entry for entry, as the
# p' matrix is (a check per solve finds the few blocks where it is not), the
# west, south and bottom coefficients are the neighbours' ae, an, at; the
# reference's loop streams all seven arrays (72 B).
KERNEL_BYTES = {"sip_fwd": 48, "sip_bwd_dot": 48, "pk_update": 40, "spmv_dot": 48,
                "update_l1": 48, "sip_fwd_res": 64}
KERNEL_NAMES = list(KERNEL_BYTES)
ITER_KERNELS = ("sip_fwd_res", "sip_bwd_dot", "pk_update", "spmv_dot")
ITER_BYTES = sum(KERNEL_BYTES[q] for q in ITER_KERNELS)   # 200


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=256, help="interior cells per direction")
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--cpu-iters", type=int, default=24,
                    help="PCG-SIP iterations of the CPU baseline sample (0: skip)")
    ap.add_argument("--others",
                    default="cgstab_sip,iccg,cgstab,pcg_jacobi,mixed,fcg,learned,unet,stencils,channel,"
                            "channel_mixed,channel_mixed_warm,learned_channel,channel_mixed_learned,"
                            "channel_fp32_flux,host",
                    help="comma list of extra solves to time once each (fcg: PCG-SIP with "
                         "the preconditioner in FP32; learned: the U-Net "
                         "guess, trained on solver-generated 64^3 pairs, then the mixed and the "
                         "FP64 solve from it to the same absolute residual; unet: the learned "
                         "guess's forward pass and training step on 128x96x64; stencils: the fused "
                         "three-direction assembly passes against the per-direction calls and "
                         "one Gauss gradient on a 256^3 channel state; channel: time "
                         "steps (Courant, LES, calcuvw, PISO) of the 384x256x192 channel with "
                         "the FP64 p' solve; channel_mixed: the same with the mixed-precision "
                         "p' solve; channel_mixed_warm: the same with the last corrector "
                         "started from the previous step's p; learned_channel: the learned "
                         "guess trained on and evaluated against the previous step's p on the "
                         "128x96x64 channel; channel_mixed_learned: the 384x256x192 channel with "
                         "the mixed p' solve and the last corrector started from that network's "
                         "guess; channel_fp32_flux: 10 steps of the 384x256x192 channel with the "
                         "momentum fluxes' face arithmetic in FP32 against FP64; host: the same "
                         "p' solve through host arrays, PCIe copies included)")
    ap.add_argument("--c4-pairs", type=int, default=600,
                    help="training pairs of the learned guess on the 128x96x64 channel")
    ap.add_argument("--c4-epochs", type=int, default=10,
                    help="training epochs over those pairs")
    ap.add_argument("--channel-steps", type=int, default=100,
                    help="time steps of the 384x256x192 channel timed with the mixed-precision "
                         "p' solve (the FP64 comparison runs a tenth of them)")
    ap.add_argument("--pmc-sample", action="store_true",
                    help="short run for hardware-counter passes: 3 PCG-SIP iterations, then "
                         "a 1 GiB device copy as the byte-counter calibration; prints one line")
    ap.add_argument("--pmc-stencils", action="store_true",
                    help="with --pmc-sample: also the stencil leg once (one call of each pass "
                         "on the 256^3 channel state) before the copy")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(the p' path runs on one MI355X)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    import __graft_entry__ as GE
    cap = GE.load_package()
    torch.cuda.set_device(0)
    n = args.n
    dims = (n + 4, n + 4, n + 4)
    ncell = n ** 3
    npad = dims[0] * dims[1] * dims[2]

    t0 = time.time()
    arrs = {f: torch.zeros(npad, dtype=torch.float64, device="cuda")
            for f in ("ae", "aw", "an", "as_", "at", "ab", "ap", "su")}
    coef = cap.Coef(**arrs)
    ctx = cap.Context(*dims, device_pointers=True)
    ctx.build_lognormal(coef, args.seed, 1.0)
    torch.cuda.synchronize()
    log(f"[bench] generated {n}^3 system in {time.time() - t0:.1f}s")
    fi = torch.zeros(npad, dtype=torch.float64, device="cuda")

    def solve_pcg():
        fi.zero_()
        return ctx.pcgsip(fi, coef, alfa=0.0, sor=1e-8, nsw=20000)

    if args.pmc_sample:
        fi.zero_()
        r = ctx.pcgsip(fi, coef, alfa=0.0, sor=1e-8, nsw=3)
        if args.pmc_stencils:
            _bench_stencils(cap, torch, reps=1)
        src = torch.ones(1 << 27, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"pmc_sample": True, "iters": r["iters"], "copy_bytes": 1 << 30}))
        ctx.close()
        return

    for w in range(args.warmup):
        rep = solve_pcg()
        log(f"[bench] warmup {w}: iters {rep['iters']} conv {rep['converged']} {rep['ms']:.1f} ms")
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    reps = []
    for k in range(args.steps):
        reps.append(solve_pcg())
    torch.cuda.synchronize()
    ms_step = (time.perf_counter() - t1) * 1e3 / args.steps
    rep = reps[-1]
    fi64 = fi.clone()      # the headline's FP64 solution, the mixed solve's reference
    log(f"[bench] pcgsip: {ms_step:.1f} ms/solve, iters {rep['iters']}, conv {rep['converged']}")
    iters = rep["iters"]

    # live per-kernel timing on the library's stream (HIP events)
    kms = ctx.time_kernels(fp32=False, reps=10)
    kern = {}
    for name, t in zip(KERNEL_NAMES, kms):
        gbs = KERNEL_BYTES[name] * ncell / (t * 1e-3) / 1e9
        kern[name] = {"ms": round(t, 4), "GBps": round(gbs, 1),
                      "bytes_per_launch": KERNEL_BYTES[name] * ncell}
    iter_ms = sum(kern[q]["ms"] for q in ITER_KERNELS)
    dom = max(ITER_KERNELS, key=lambda q: kern[q]["ms"])
    achieved = kern[dom]["GBps"]
    log(f"[bench] kernels: {json.dumps(kern)}")

    solvers = {"pcgsip_fp64_alfa0": {"ms": round(ms_step, 2), "iters": iters,
                                     "converged": rep["converged"],
                                     "ms_per_iter": round(ms_step / max(iters, 1), 4),
                                     "effective_GBps": round(ITER_BYTES * ncell * iters /
                                                             (ms_step * 1e-3) / 1e9, 1)}}
    c4_params = None
    tight, tight_iters, fp64_err = None, None, None
    for other in [o for o in args.others.split(",") if o]:
        fi.zero_()
        torch.cuda.synchronize()
        t = time.perf_counter()
        if other == "cgstab_sip":
            r = ctx.cgstab_sip(fi, coef, alfa=0.92, sor=1e-8, nsw=20000)
        elif other == "iccg":
            r = ctx.iccg(fi, coef, sor=1e-8, nsw=20000)
        elif other == "cgstab":
            r = ctx.cgstab(fi, coef, sor=1e-8, nsw=20000)
        elif other == "pcg_jacobi":
            r = ctx.pcg_jacobi(fi, coef, sor=1e-8, nsw=40000)
        elif other == "mixed":
            r = ctx.pcgsip_mixed(fi, coef, alfa=0.0, sor=1e-8, nsw=40000, inner_rtol=3e-4)
        elif other == "fcg":
            r = ctx.pcgsip_fcg(fi, coef, alfa=0.0, sor=1e-8, nsw=20000)
        elif other == "learned":
            solvers["learned"] = _bench_learned(cap, torch, ctx, coef, arrs, dims, fi, args.seed,
                                                rep["resor"], solvers.get("mixed"))
            continue
        elif other == "unet":
            solvers["unet"] = _bench_unet(cap, torch)
            continue
        elif other == "stencils":
            solvers["stencils"] = _bench_stencils(cap, torch)
            continue
        elif other == "channel":
            solvers["channel_fp64"] = _bench_channel(cap, torch, max(1, args.channel_steps // 10), 4)
            continue
        elif other == "channel_mixed":
            solvers["channel_mixed"] = _bench_channel(cap, torch, args.channel_steps, 6)
            continue
        elif other == "channel_mixed_warm":
            solvers["channel_mixed_warm"] = _bench_channel(cap, torch, args.channel_steps, 6,
                                                           start="warm")
            continue
        elif other == "channel_fp32_flux":
            solvers["channel_fp32_flux"] = _bench_fp32_flux(cap, torch, 10)
            continue
        elif other == "learned_channel":
            solvers["learned_channel"], c4_params = _bench_learned_channel(
                cap, torch, args.c4_pairs, args.c4_epochs)
            continue
        elif other == "channel_mixed_learned":
            if c4_params is None:
                log("[bench] channel_mixed_learned needs learned_channel before it; skipped")
                continue
            solvers["channel_mixed_learned"] = _bench_channel(cap, torch, args.channel_steps, 6,
                                                              start="learned", params=c4_params)
            continue
        elif other == "host":
            solvers["pcgsip_host_pointers"] = _bench_host(cap, torch, arrs, dims)
            continue
        else:
            continue
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t) * 1e3
        solvers[other] = {"ms": round(dt, 2), "iters": r["iters"], "converged": r["converged"]}
        log(f"[bench] {other}: {dt:.1f} ms, iters {r['iters']}, conv {r['converged']}")
        if other in ("mixed", "fcg"):
            solvers[other]["error_vs_fp64"] = _mixed_error(arrs, fi64, fi, dims)
            # what the 1e-8 stop itself leaves: both solutions against an FP64
            # solve to 1e-12 of the same system (solved once)
            if tight is None:
                tight = torch.zeros_like(fi)
                rt = ctx.pcgsip(tight, coef, alfa=0.0, sor=1e-12, nsw=40000)
                tight_iters = rt["iters"]
                fp64_err = _mixed_error(arrs, tight, fi64, dims)
            solvers[other]["error_vs_fp64_1e-12"] = _mixed_error(arrs, tight, fi, dims)
            solvers[other]["fp64_error_vs_fp64_1e-12"] = fp64_err
            solvers[other]["fp64_1e-12_iters"] = tight_iters
            log(f"[bench] {other} vs FP64: {json.dumps(solvers[other]['error_vs_fp64'])}; against "
                f"FP64 to 1e-12 ({tight_iters} iterations): {other} "
                f"{json.dumps(solvers[other]['error_vs_fp64_1e-12'])}, FP64 to 1e-8 "
                f"{json.dumps(fp64_err)}")

    cpu = None
    if args.cpu_iters > 0:
        O = GE.load_oracle()
        t = time.time()
        s = O.build_lognormal(dims, args.seed, 1.0)
        fi_c = np.zeros(s.n)
        sec = O.pcgsip_time_iters(s, fi_c, args.cpu_iters, alfa=0.0)
        per_it = sec / args.cpu_iters
        cpu = {"value": round(per_it * iters * 1e3, 1), "unit": "ms per p' solve",
               "cores": 1, "kind": "port",
               "sample": (f"{args.cpu_iters} PCG-SIP iterations of the same {n}^3 system "
                          f"({sec:.1f} s, {per_it * 1e3:.1f} ms/iteration, single thread, "
                          f"C restatement -O2), scaled to the GPU's {iters} iterations"),
               "cpu": _cpu_model(), "nproc": os.cpu_count()}
        log(f"[bench] cpu baseline {per_it * 1e3:.1f} ms/iter (setup {time.time() - t:.1f}s)")

    out = {
        "metric": METRIC,
        "value": round(ms_step, 2),
        "unit": "ms",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 2),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": (f"{n}^3 log-normal (sigma=1) Neumann p' operator, uniform "
                                "grid, reference cell pinned, N(0,1) RHS, fi0=0; "
                                "PCG-SIP (alfa=0) to 1e-8 L1 relative residual"),
                   "cells": ncell, "padded": list(dims), "seed": args.seed},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": _pmc_traffic(dom, n),
                     "bytes_per_launch": KERNEL_BYTES[dom] * ncell,
                     "avg_launch_ms": kern[dom]["ms"]},
        "pct_hbm_iteration": round(100 * ITER_BYTES * ncell / (iter_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 1),
        "cell_updates_per_s": round(ncell * iters / (ms_step * 1e-3), 1),
        "kernels": kern,
        # the metric's pair: the plain FP64 drop-in (value) against the learned
        # guess + mixed-precision solve, both to the same absolute residual
        "plain_vs_learned_mixed": {
            "plain_fp64_ms": round(ms_step, 2),
            "learned_mixed_ms": solvers.get("learned", {}).get("mixed", {}).get("ms"),
            "mixed_ms": solvers.get("mixed", {}).get("ms"),
            # the learned guess on its training distribution: ms per time step
            # of the 384x256x192 channel by the start of the last corrector
            "channel_ms_per_step": {
                "fp64": solvers.get("channel_fp64", {}).get("ms_per_step"),
                "mixed": solvers.get("channel_mixed", {}).get("ms_per_step"),
                "mixed_warm": solvers.get("channel_mixed_warm", {}).get("ms_per_step"),
                "mixed_learned": solvers.get("channel_mixed_learned", {}).get("ms_per_step")}},
        "solvers": solvers,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    ctx.close()


def _bench_host(cap, torch, arrs, dims):
    """The same p' solve through a host-pointer context (the drop-in's mode:
    the coefficients go up and fi, res and the factors come back over PCIe
    on every call): ms per call, copies included."""
    host = {f: a.cpu().numpy() for f, a in arrs.items()}
    npad = dims[0] * dims[1] * dims[2]
    import numpy as np
    extra = {f: np.zeros(npad) for f in ("res", "bb", "bs", "bw", "bp", "bn", "be", "bt")}
    coef = cap.Coef(**host, **extra)
    fi = np.zeros(npad)
    with cap.Context(*dims) as hctx:
        hctx.pcgsip(fi, coef, alfa=0.0, sor=1e-8, nsw=20000)     # first call allocates
        fi[:] = 0.0
        t = time.perf_counter()
        r = hctx.pcgsip(fi, coef, alfa=0.0, sor=1e-8, nsw=20000)
        ms = (time.perf_counter() - t) * 1e3
    log(f"[bench] pcgsip through host arrays: {ms:.1f} ms (device time {r['ms']:.1f} ms), "
        f"iters {r['iters']}")
    return {"ms": round(ms, 2), "device_ms": round(r["ms"], 2), "iters": r["iters"],
            "note": "host-pointer context: 8 coefficient arrays up, fi, res and 7 factor arrays "
                    "down per call"}


def _l1_residual(arrs, x, dims):
    """sum |su - M x| over the solve range k, i, j = 3..n-2 (1-based) of a
    padded reference-layout array x, with M x = ap x - sum a_nb x_nb; the
    index is (k-1) ni nj + (i-1) nj + j, so the view is (k, i, j)."""
    ni, nj, nk = dims
    v = lambda a: a.view(nk, ni, nj)  # noqa: E731
    c = slice(2, -2)
    X = v(x)
    r = (v(arrs["su"])[c, c, c] - v(arrs["ap"])[c, c, c] * X[c, c, c]
         + v(arrs["ae"])[c, c, c] * X[c, 3:-1, c] + v(arrs["aw"])[c, c, c] * X[c, 1:-3, c]
         + v(arrs["an"])[c, c, c] * X[c, c, 3:-1] + v(arrs["as_"])[c, c, c] * X[c, c, 1:-3]
         + v(arrs["at"])[c, c, c] * X[3:-1, c, c] + v(arrs["ab"])[c, c, c] * X[1:-3, c, c])
    return float(r.abs().sum())


def _mixed_error(arrs, x64, x, dims):
    """Errors of a solution x against the FP64 solution x64 of the same
    system: p' on the solve range, and the face mass-flux corrections the
    solution makes, a_f (p'_nb - p'_P) on the east, north and top faces
    between solve-range cells (calcp-multiple_correction_SIMPLE.f90:173-
    190); max-norm and RMS, each relative to the same norm of the FP64
    values."""
    ni, nj, nk = dims
    v = lambda a: a.view(nk, ni, nj)  # noqa: E731
    c = slice(2, -2)
    X, X64 = v(x), v(x64)
    out = {}

    def rel(d, ref):
        return {"max": float(d.abs().max() / ref.abs().max()),
                "rms": float(d.pow(2).mean().sqrt() / ref.pow(2).mean().sqrt())}
    out["p"] = rel(X[c, c, c] - X64[c, c, c], X64[c, c, c])
    lo, hi = slice(2, -3), slice(3, -2)
    faces = {"east": (arrs["ae"], (c, lo, c), (c, hi, c)),
             "north": (arrs["an"], (c, c, lo), (c, c, hi)),
             "top": (arrs["at"], (lo, c, c), (hi, c, c))}
    d_all, r_all = [], []
    for name, (a, P, E) in faces.items():
        A = v(a)[P]
        f = A * (X[E] - X[P])
        f64 = A * (X64[E] - X64[P])
        d_all.append((f - f64).reshape(-1))
        r_all.append(f64.reshape(-1))
    import torch
    out["face_flux_correction"] = rel(torch.cat(d_all), torch.cat(r_all))
    return out


def _bench_learned(cap, torch, ctx, coef, arrs, dims, fi, seed, r0, mixed, n_small=64,
                   ntrain=16, epochs=30):
    """The learned initial guess on the benchmark system.  The U-Net is
    trained on ntrain solver-generated pairs of the same operator family at
    n_small^3 (other seeds: other coefficients and right-hand sides; inputs
    su / rms, ap / mean, no previous p', target p' / rms), its weights are
    moved to the full grid, and from its guess (kept only when it lowers the
    L1 residual of fi = 0) the mixed-precision solve and the FP64 PCG-SIP run
    to the same absolute residual as the plain solve from zero: the stop test
    is relative to the solve's own start, so sor is scaled by r0 / r_guess.
    Times are host wall clock around each part, synchronised."""
    n = dims[0] - 4
    ds = (n_small + 4,) * 3
    npad_s = ds[0] * ds[1] * ds[2]
    t0 = time.perf_counter()
    sarr = {f: torch.zeros(npad_s, dtype=torch.float64, device="cuda")
            for f in ("ae", "aw", "an", "as_", "at", "ab", "ap", "su")}
    scoef = cap.Coef(**sarr)
    sctx = cap.Context(*ds, device_pointers=True)
    interior = lambda a, d: a.view(d[2], d[0], d[1])[2:-2, 2:-2, 2:-2].reshape(-1)  # noqa: E731
    rms = lambda a: a.pow(2).mean().sqrt()  # noqa: E731
    X, T = [], []
    for q in range(ntrain):
        sctx.build_lognormal(scoef, seed + 1 + q, 1.0)
        sfi = torch.zeros(npad_s, dtype=torch.float64, device="cuda")
        rs = sctx.pcgsip(sfi, scoef, alfa=0.0, sor=1e-8, nsw=20000)
        assert rs["converged"]
        su, ap, sol = interior(sarr["su"], ds), interior(sarr["ap"], ds), interior(sfi, ds)
        X.append(torch.stack([su / rms(su), ap / ap.mean(), torch.zeros_like(su)])
                 .float().reshape(-1).contiguous())
        T.append((sol / rms(sol)).float().contiguous())
    torch.cuda.synchronize()
    t_data = time.perf_counter() - t0
    net = cap.UNet(sctx, n_small, n_small, n_small, seed=1)
    t0 = time.perf_counter()
    losses = [net.train_step(X[it % ntrain], T[it % ntrain], 5e-4 if it < 20 * ntrain else 1e-4)
              for it in range(epochs * ntrain)]
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t0
    params = net.get_params()
    net.close()
    sctx.close()
    loss0, loss1 = sum(losses[:ntrain]) / ntrain, sum(losses[-ntrain:]) / ntrain
    log(f"[bench] learned: {ntrain} pairs at {n_small}^3 in {t_data:.1f}s, {epochs * ntrain} "
        f"training steps in {t_train:.1f}s, loss {loss0:.3f} -> {loss1:.3f}")

    big = cap.UNet(ctx, n, n, n, seed=1)
    big.set_params(params)
    res = {"train": {"grid": [n_small] * 3, "pairs": ntrain, "steps": epochs * ntrain,
                     "loss_first_epoch": round(loss0, 4), "loss_last_epoch": round(loss1, 4),
                     "data_s": round(t_data, 2), "train_s": round(t_train, 2)}}
    r_zero = _l1_residual(arrs, torch.zeros_like(fi), dims)
    res["r0_check"] = {"torch_l1": r_zero, "solver_resor": r0}
    for solver in ("mixed", "fp64"):
        fi.zero_()
        big.guess(fi, coef)           # warm-up of the full-grid network
        fi.zero_()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        big.guess(fi, coef)
        torch.cuda.synchronize()
        t_guess = time.perf_counter() - t0
        r_guess = _l1_residual(arrs, fi, dims)
        accepted = r_guess < r0
        if not accepted:
            fi.zero_()
        start = r_guess if accepted else r0
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        if solver == "mixed":
            r = ctx.pcgsip_mixed(fi, coef, alfa=0.0, sor=1e-8 * r0 / start, nsw=40000,
                                 inner_rtol=3e-4)
        else:
            r = ctx.pcgsip(fi, coef, alfa=0.0, sor=1e-8 * r0 / start, nsw=20000)
        torch.cuda.synchronize()
        t_solve = time.perf_counter() - t0
        r_end = _l1_residual(arrs, fi, dims)
        ms = (t_guess + t_solve) * 1e3
        res[solver] = {"ms": round(ms, 2), "guess_ms": round(t_guess * 1e3, 2),
                       "solve_ms": round(t_solve * 1e3, 2), "iters": r["iters"],
                       "converged": r["converged"], "accepted": accepted,
                       "r_guess_over_r0": round(r_guess / r0, 4),
                       "final_l1_over_r0": float(f"{r_end / r0:.3e}")}
        log(f"[bench] learned+{solver}: {ms:.1f} ms (guess {t_guess * 1e3:.1f} ms, accepted "
            f"{accepted}, L1 {r_guess / r0:.3f} of r0; solve {t_solve * 1e3:.1f} ms, "
            f"{r['iters']} iterations), final L1 {r_end / r0:.2e} of r0")
    if mixed:
        res["mixed_from_zero"] = {"ms": mixed["ms"], "iters": mixed["iters"]}
    big.close()
    return res


# (cin, cout, level) of the U-Net's 3x3x3 convolutions, encoder then decoder
UNET_LAYERS = [(3, 16, 0), (16, 16, 0), (16, 32, 1), (32, 32, 1), (32, 48, 2), (48, 48, 2),
               (48, 64, 3), (64, 64, 3), (112, 48, 2), (48, 48, 2), (80, 32, 1), (32, 32, 1),
               (48, 16, 0), (16, 16, 0), (16, 1, 0)]
MFMA_BF16_PEAK_TFLOPS = 2500.0


def _bench_unet(cap, torch, nx=128, ny=96, nz=64, reps=5):
    """Forward pass and training step of the learned guess's U-Net on the
    128 x 96 x 64 channel box: ms and algorithmic TFLOP/s (2 x 27 x cin x
    cout multiply-adds per voxel and layer, padding not counted)."""
    ctx = cap.Context(ny + 4, nx + 4, nz + 4, device_pointers=True)
    net = cap.UNet(ctx, nx, ny, nz, seed=1)
    nvox = nx * ny * nz
    x = torch.randn(3 * nvox, dtype=torch.float32, device="cuda")
    t = torch.randn(nvox, dtype=torch.float32, device="cuda")
    out = torch.empty(nvox, dtype=torch.float32, device="cuda")
    flops = sum(2 * 27 * ci * co * (nvox >> (3 * lv)) for ci, co, lv in UNET_LAYERS)
    net.forward(x, out)
    net.train_step(x, t, 1e-4)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        net.forward(x, out)
    torch.cuda.synchronize()
    fwd = (time.perf_counter() - t0) * 1e3 / reps
    t0 = time.perf_counter()
    for _ in range(reps):
        net.train_step(x, t, 1e-4)
    torch.cuda.synchronize()
    trn = (time.perf_counter() - t0) * 1e3 / reps
    net.close()
    ctx.close()
    tf = flops / (fwd * 1e-3) / 1e12
    log(f"[bench] unet {nx}x{ny}x{nz}: forward {fwd:.2f} ms ({tf:.1f} TFLOP/s), "
        f"train step {trn:.2f} ms")
    return {"grid": [nx, ny, nz], "forward_ms": round(fwd, 3), "forward_gflop": round(flops / 1e9, 2),
            "forward_TFLOPs": round(tf, 2), "mfma_frac": round(tf / MFMA_BF16_PEAK_TFLOPS, 4),
            "train_step_ms": round(trn, 3),
            "note": "host-side wall time per call including launch and synchronisation"}


# bytes per interior cell of the fused three-direction passes (the distinct
# arrays one pass reads and writes, 8 B each; fluxmass3 equals the 384 B/cell
# of fluxmass2.f90's three directions) and of one Gauss gradient
STENCIL_BYTES = {"fluxmass": 384, "piso_assemble": 408, "fluxuvw": 416, "grad_gauss": 136}


def _bench_stencils(cap, torch, n=256, reps=5):
    """The p' assembly stencils on an n^3 channel-geometry state (periodic x,
    z, tanh-stretched y, after one time step): the drivers' fused
    three-direction passes against the three per-direction entry points they
    replace, and one Gauss gradient.  ms per call (host wall time around
    synchronised calls) and GB/s over each fused pass's distinct arrays."""
    CH = _channel_module()
    ch = CH.Channel(cap, torch, n, n, n, solver=6)
    try:
        ch.step(timed=False)
        ctx, f = ch.ctx, ch.f
        F = {q: f[q] for q in cap.host.FIELD_NAMES}
        a6 = [f[q] for q in ("ae", "aw", "an", "as", "at", "ab")]
        f3 = [f[q] for q in ("f1", "f2", "f3")]
        scr = [torch.zeros_like(f["u"]) for _ in range(3)]
        bp3 = [f[q] for q in ("be", "bn", "bt")]
        su3 = [f[q] for q in ("su", "sv", "sw")]
        dirs = (("ae", "aw", "f1", "be"), ("an", "as", "f2", "bn"), ("at", "ab", "f3", "bt"))
        geom = {q: f[q] for q in cap.host.GEOM_FIELDS}
        g3 = [torch.empty_like(f["u"]) for _ in range(3)]

        def per_dir(kind):
            for d, (ae, aw, fc, bp) in enumerate(dirs):
                if kind == "fluxmass":
                    ctx.fluxmass(F, d, f[ae], f[aw], f[fc])
                elif kind == "piso_assemble":
                    ctx.piso_assemble_pressure_eq(F, d, f[ae], f[aw], f[fc], 1, f[bp], f["su"])
                else:
                    ctx.fluxuvw(F, d, f["vis"], f[ae], f[aw], f[fc], *su3, *scr, 1, 1.0)

        def fused(kind):
            if kind == "fluxmass":
                ctx.fluxmass3(F, a6, f3)
            elif kind == "piso_assemble":
                ctx.piso_assemble_pressure_eq3(F, a6, f3, 1, bp3, f["su"])
            else:
                ctx.fluxuvw3(F, f["vis"], a6, f3, su3, scr, 1, 1.0)

        def timed(fn, *a):
            fn(*a)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(reps):
                fn(*a)
            torch.cuda.synchronize()
            return (time.perf_counter() - t0) * 1e3 / reps

        ncell = n ** 3
        res = {"grid": [n, n, n], "reps": reps}
        for kind in ("fluxmass", "piso_assemble", "fluxuvw"):
            t3, t1 = timed(per_dir, kind), timed(fused, kind)
            gbs = STENCIL_BYTES[kind] * ncell / (t1 * 1e-3) / 1e9
            res[kind] = {"three_calls_ms": round(t3, 3), "fused_ms": round(t1, 3),
                         "fused_GBps": round(gbs, 1), "fused_frac": round(gbs / HBM_PEAK_GBS, 3),
                         "bytes_per_cell": STENCIL_BYTES[kind]}
            log(f"[bench] {kind} {n}^3: three calls {t3:.3f} ms, fused {t1:.3f} ms "
                f"({gbs:.0f} GB/s over {STENCIL_BYTES[kind]} B/cell)")
        tg = timed(ctx.grad_gauss, f["u"], geom, *g3)
        gbs = STENCIL_BYTES["grad_gauss"] * ncell / (tg * 1e-3) / 1e9
        res["grad_gauss"] = {"ms": round(tg, 3), "GBps": round(gbs, 1),
                             "frac": round(gbs / HBM_PEAK_GBS, 3), "bytes_per_cell": 136}
        log(f"[bench] grad_gauss {n}^3: {tg:.3f} ms ({gbs:.0f} GB/s)")
        res["note"] = ("host wall time per synchronised call; each per-direction entry point "
                       "synchronises its stream once")
        return res
    finally:
        ch.close()


def _channel_module():
    import importlib.util
    spec = importlib.util.spec_from_file_location(
        "cap_channel", os.path.join(ROOT, "cappuccino-ghostcells-mpi_amd", "channel.py"))
    CH = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(CH)
    return CH


def _bench_channel(cap, torch, steps, solver, size=(384, 256, 192), start="reference",
                   params=None):
    """Time steps of the channel (cappuccino-ghostcells-mpi_amd/channel.py):
    384 x 256 x 192 cells, periodic x and z, walls at y (the reference's k),
    tanh stretching 2.0, Re_bulk 5600, Smagorinsky LES, dt from Co = 0.5,
    CDS momentum with BDF2, PISO with 2 correctors (0.1, then 1e-8) and the
    FP64 (solver 4) or mixed-precision (solver 6) p' solve; everything on the
    device.  start "warm": the last corrector's solve starts from the
    previous step's p, "learned": from the learned guess with the network
    weights `params`, each when it lowers the L1 residual of the reference's
    start, and stops at the absolute residual that start stops at.  One
    untimed step first."""
    import numpy as np
    CH = _channel_module()
    nx, ny, nz = size
    ch = CH.Channel(cap, torch, nx, ny, nz, solver=solver)
    warm = start == "warm"
    net = None
    if start == "learned":
        ni, nj, nk = ch.dims
        net = cap.UNet(ch.ctx, nj - 4, ni - 4, nk - 4, seed=1)
        net.set_params(params)
    ch.step(timed=False, warm=warm, net=net)
    iters, accepted, l1_ratio = [], [], []
    t = time.perf_counter()
    for s in range(steps):
        reps = ch.step(warm=warm, net=net)
        iters.append([r["iters"] for r in reps])
        if warm or net is not None:
            accepted.append(ch.last["accepted"])
            l1_ratio.append(ch.last["l1"][1] / ch.last["l1"][0])
        if (s + 1) % 20 == 0:
            log(f"[bench] channel step {s + 1}/{steps}")
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t) * 1e3 / steps
    per = {k: round(v, 2) for k, v in ch.phase_ms().items()}
    it = np.array(iters)
    out = {"grid": [nx, ny, nz], "padded": list(ch.dims), "steps": steps,
           "ms_per_step": round(ms, 2), "steps_per_s": round(1e3 / ms, 3), "phases_ms": per,
           "p_solve_iters_mean_per_corrector": [round(float(x), 1) for x in it.mean(axis=0)],
           "continuity_last": [float(x) for x in ch.last["cont"][-1]],
           "courant": ch.last["courant"], "dt": ch.last["dt"],
           "bulk_velocity": round(ch.bulk_velocity(), 6),
           "start": start,
           "case": f"{nx}x{ny}x{nz} channel, periodic x and z, walls at y (reference k), tanh "
                   "2.0, Re_bulk 5600, Smagorinsky LES, Co 0.5, CDS + BDF2 momentum, PISO 2 "
                   "correctors, p' by " + ("mixed-precision PCG-SIP" if solver == 6 else
                                           "PCG-SIP (alfa 0), FP64") +
                   ("; last corrector from the previous step's p" if warm else "")}
    if accepted:
        out["start_accepted_fraction"] = round(float(np.mean(accepted)), 3)
        out["l1_start_over_reference_mean"] = round(float(np.mean(l1_ratio)), 5)
    log(f"[bench] channel (solver {solver}, {start} start): {out['ms_per_step']} ms/step {per}, "
        f"iterations {out['p_solve_iters_mean_per_corrector']}")
    if net is not None:
        net.close()
    ch.close()
    return out


def _bench_fp32_flux(cap, torch, steps, size=(384, 256, 192)):
    """The channel's momentum assembly with fluxuvw's face arithmetic in
    FP32 (CAP_UVW_FP32) against FP64: `steps` steps of each from the same
    start (mixed p' solve); calcuvw's ms per step and the largest velocity
    difference at the end, relative to the bulk velocity (1)."""
    CH = _channel_module()
    res, vel = {}, []
    for fp32 in (False, True):
        ch = CH.Channel(cap, torch, *size, solver=6, fp32_flux=fp32)
        ch.step(timed=False)
        for _ in range(steps):
            ch.step()
        res["fp32" if fp32 else "fp64"] = {k: round(v, 2) for k, v in ch.phase_ms().items()}
        vel.append({q: ch.f[q].clone() for q in ("u", "v", "w")})
        ch.close()
    diff = max(float((vel[0][q] - vel[1][q]).abs().max()) for q in ("u", "v", "w"))
    out = {"grid": list(size), "steps": steps, "phases_ms_fp64_flux": res["fp64"],
           "phases_ms_fp32_flux": res["fp32"], "max_velocity_difference": float(f"{diff:.3e}")}
    log(f"[bench] fp32 momentum fluxes: {json.dumps(out)}")
    return out


def _bench_learned_channel(cap, torch, pairs, epochs, evals=8, warmup=20):
    """The learned guess on its own training distribution: the 128 x 96 x 64
    channel (the 384 x 256 x 192 case's family).  After `warmup` steps, the
    FP64 PISO steps give `pairs` (inputs from the system of the last
    corrector and the previous step's p, target its p) that train the U-Net
    on the device.  Then, on `evals` held-out steps, each step is taken from
    the same state with the last corrector's solve started three ways: the
    reference's start (the previous corrector's p), the previous step's p
    (warm), and the learned guess -- all stopping at the absolute residual
    the reference's start stops at -- with the FP64 and the mixed-precision
    solve.  Times are host wall clock around each step, synchronised.
    Returns the report and the trained weights."""
    import numpy as np
    CH = _channel_module()
    ch = CH.Channel(cap, torch, 128, 96, 64, solver=4)
    t0 = time.perf_counter()
    for _ in range(warmup):
        ch.step(timed=False)
    X, T = [], []
    for _ in range(pairs):
        prev = ch.f["p"].clone()
        ch.step(timed=False)
        x, t = ch.pressure_pair(prev)
        X.append(x)
        T.append(t)
    torch.cuda.synchronize()
    t_data = time.perf_counter() - t0
    ni, nj, nk = ch.dims
    net = cap.UNet(ch.ctx, nj - 4, ni - 4, nk - 4, seed=1)
    g = torch.Generator().manual_seed(0)
    t0 = time.perf_counter()
    losses = []
    for ep in range(epochs):
        lr = 5e-4 if ep < int(0.7 * epochs) else 1e-4
        losses.append(float(np.mean([net.train_step(X[q], T[q], lr)
                                     for q in torch.randperm(pairs, generator=g).tolist()])))
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t0
    del X, T
    params = net.get_params()
    log(f"[bench] learned channel: {pairs} pairs in {t_data:.1f}s, {epochs} epochs in "
        f"{t_train:.1f}s, loss {losses[0]:.4f} -> {losses[-1]:.4f}")
    res = {k: {"iters": [], "ms": [], "l1_start_over_ref": []}
           for k in ("reference", "warm", "learned", "mixed_reference", "mixed_warm", "mixed_learned")}
    for _ in range(evals):
        snap = ch.snapshot()
        for solver, pre in ((4, ""), (6, "mixed_")):
            ch.opts.solver = solver
            for kind in ("reference", "warm", "learned"):
                ch.restore(snap)
                torch.cuda.synchronize()
                t = time.perf_counter()
                reps = ch.step(timed=False, net=net if kind == "learned" else None,
                               warm=kind == "warm")
                torch.cuda.synchronize()
                r = res[pre + kind]
                r["ms"].append((time.perf_counter() - t) * 1e3)
                r["iters"].append(reps[-1]["iters"])
                if kind != "reference":
                    l1 = ch.last["l1"]
                    r["l1_start_over_ref"].append(l1[1] / l1[0])
                    assert ch.last["accepted"]
        ch.opts.solver = 4
        ch.restore(snap)
        ch.step(timed=False)
    net.close()
    ch.close()
    out = {"grid": [128, 96, 64], "pairs": pairs, "epochs": epochs, "data_s": round(t_data, 1),
           "train_s": round(t_train, 1), "loss_first_epoch": round(losses[0], 4),
           "loss_last_epoch": round(losses[-1], 4), "eval_steps": evals}
    for k, r in res.items():
        out[k] = {"iters_mean": round(float(np.mean(r["iters"])), 2),
                  "step_ms_mean": round(float(np.mean(r["ms"])), 2)}
        if r["l1_start_over_ref"]:
            out[k]["l1_start_over_reference"] = round(float(np.mean(r["l1_start_over_ref"])), 5)
    log(f"[bench] learned channel: {json.dumps(out)}")
    return out, params


def _pmc_traffic(kernel, n):
    """HBM bytes per launch of a kernel at an n^3 grid from the committed
    counter summary (profiles/pmc_traffic.txt, lines `traffic <kernel> <bytes>`
    measured at the grid named by its `grid <n>` line), or None: the counters
    were read at one grid size, so any other size reports no traffic."""
    try:
        grid, found = None, None
        for line in open(os.path.join(ROOT, "profiles", "pmc_traffic.txt")):
            f = line.split()
            if len(f) == 2 and f[0] == "grid":
                grid = int(f[1])
            if len(f) == 3 and f[0] == "traffic" and f[1] == kernel:
                found = int(f[2])
        return found if grid == n else None
    except (OSError, ValueError):
        return None


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


if __name__ == "__main__":
    main()
