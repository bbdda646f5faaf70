"""Time the C restatement (liboracle.so) beside the reference's own solvers
compiled from its sources (oracle/_ref/ref_driver: sipsol.f90, cgstab_sip.f90
built with amdflang -O2, the reference Makefile's -O2), on the same seeded
log-normal p' systems, one host core each.  Test infrastructure: it shows
whether the CPU baseline bench.py reports (the restatement) runs at the speed
of the compiled reference.  Build container only (the reference build never
leaves it).

usage: python oracle/time_against_ref.py [--sizes 32,128] [--iters 40] [--reps 3]

Both run a fixed number of iterations (sor far below reach), and the script
checks that the two report the same iteration count and final residual.
The reference is timed inside its driver around the solver call alone
(SOLVE_SECONDS, system_clock), the restatement around its ctypes call.
"""
import argparse
import os
import re
import subprocess
import sys
import tempfile
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
import oracle as O  # noqa: E402

_SECS = re.compile(r"SOLVE_SECONDS\s+([-+0-9.Ee]+)")


def _big_stack():
    import resource
    _, hard = resource.getrlimit(resource.RLIMIT_STACK)
    resource.setrlimit(resource.RLIMIT_STACK, (hard, hard))


def ref_solve(mode, s, fi, alfa, sor, nsw):
    ni, nj, nk = s.dims
    arrays = [fi] + [getattr(s, f) for f in O.SYS_FIELDS]
    with tempfile.TemporaryDirectory() as td:
        fin, fout, flog = (os.path.join(td, x) for x in ("in.bin", "out.bin", "log.txt"))
        with open(fin, "wb") as f:
            np.array([ni, nj, nk, nsw], dtype=np.int32).tofile(f)
            np.array([alfa, sor], dtype=np.float64).tofile(f)
            for a in arrays:
                np.ascontiguousarray(a, dtype=np.float64).tofile(f)
        # the reference's solvers keep their work arrays on the stack
        # (automatic arrays of nxyza elements): give the child a large stack
        r = subprocess.run([O.REF_DRIVER, mode, fin, fout, flog], check=True,
                           capture_output=True, text=True, preexec_fn=_big_stack)
        log = open(flog).read()
    secs = float(_SECS.search(r.stdout).group(1))
    return secs, O.parse_log(log)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--sizes", default="32,128")
    ap.add_argument("--iters", type=int, default=40)
    ap.add_argument("--reps", type=int, default=3)
    args = ap.parse_args()
    if not O.ref_available():
        sys.exit("oracle/_ref/ref_driver is not built (make -C oracle ref)")
    sor = 1e-300
    print("| solver | cells | iterations | reference s | restatement s | restatement / reference |")
    print("|---|---|---|---|---|---|")
    for n in [int(x) for x in args.sizes.split(",")]:
        dims = (n + 4, n + 4, n + 4)
        s = O.build_lognormal(dims, 7, 1.0)
        for mode, fn in (("sipsol", O.sipsol), ("cgstab_sip", O.cgstab_sip)):
            t_ref, t_orc = [], []
            for _ in range(args.reps):
                secs, rl = ref_solve(mode, s.copy(), np.zeros(s.n), O.ALFA, sor, args.iters)
                t_ref.append(secs)
                s2, fi = s.copy(), np.zeros(s.n)
                t0 = time.perf_counter()
                rep = fn(s2, fi, alfa=O.ALFA, sor=sor, nsw=args.iters)
                t_orc.append(time.perf_counter() - t0)
            # the reference's log prints the residual to four digits
            if rl is not None and (rl["iters"] != rep["iters"] or
                                   abs(rl["resl"] - rep["resl"]) > 1e-3 * abs(rep["resl"])):
                sys.exit(f"{mode} {n}^3: reference {rl} and restatement {rep} differ")
            a, b = min(t_ref), min(t_orc)
            print(f"| {mode} | {n}^3 | {rep['iters']} | {a:.4f} | {b:.4f} | {b / a:.3f} |",
                  flush=True)


if __name__ == "__main__":
    main()
