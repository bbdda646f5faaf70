"""The grid geometry on the device (cap_face_areas, cap_intfac) against the
oracle's restatements of init.f90:49-144 and intfac.f90 (pinned to the
reference build and to closed forms in test_oracle_geometry.py): bit-
identical on distorted grids, through host and device pointers, and on the
128^3 grid stretched 1.05 toward every wall."""
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
import gridcase as GC  # noqa: E402

pytestmark = pytest.mark.gpu


def _grid(kind, dims, seed):
    if kind == "distorted":
        g = GC.lattice(dims, (1.0, 0.1, 0.0), (0.05, 0.9, 0.1), (0.0, 0.2, 1.1), jitter=0.15,
                       seed=seed)
    else:
        g, dims = GC.rectilinear(*(GC.stretched_widths(m - 4, 1.05) for m in dims))
    n = dims[0] * dims[1] * dims[2]
    rng = np.random.default_rng(seed)
    for q in ("fx", "fy", "fz", "ar1x", "ar1y", "ar1z", "ar2x", "ar2y", "ar2z", "ar3x", "ar3y",
              "ar3z"):
        g[q] = rng.uniform(-1, 1, n)      # values outside the loop ranges must survive
    return g, dims


CASES = [("distorted", (9, 8, 7), 1, False), ("distorted", (20, 17, 11), 2, False),
         ("stretched", (36, 20, 28), 3, False), ("distorted", (9, 8, 7), 1, True),
         ("distorted", (20, 17, 11), 2, True), ("stretched", (36, 20, 28), 3, True),
         ("stretched", (132, 132, 132), 4, True)]


@pytest.mark.parametrize("kind,dims,seed,device", CASES)
def test_geometry_bitwise(gpu, O, kind, dims, seed, device):
    import torch
    print(f"geometry {kind} {dims} device {device}", flush=True)
    g, dims = _grid(kind, dims, seed)
    ref = {q: a.copy() for q, a in g.items()}
    O.face_areas(dims, ref["x"], ref["y"], ref["z"], ref)
    ref["fx"], ref["fy"], ref["fz"] = O.intfac(dims, ref)
    if device:
        d = {q: torch.from_numpy(a.copy()).cuda() for q, a in g.items()}
    else:
        d = {q: a.copy() for q, a in g.items()}
    with gpu.Context(*dims, device_pointers=device) as ctx:
        ctx.face_areas(d)
        ctx.intfac(d)
    for q in O.AR_FIELDS + ("fx", "fy", "fz"):
        got = d[q].cpu().numpy() if device else d[q]
        np.testing.assert_array_equal(got, ref[q], err_msg=q)
