"""cap_read_grid (host code, no GPU) against grid files that a Fortran
program writes with the compiler's own unformatted records in the layout
read_grid.f90:126-161 reads; a truncated file and mismatched extents are
refused with an error, not read."""
import os
import subprocess
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
import gridcase as GC  # noqa: E402

WRITER = os.path.join(HERE, "fortran", "_build", "write_grid")
needs_writer = pytest.mark.skipif(not os.path.exists(WRITER), reason="grid writer not built")


def _write(tmp_path, dims, seed):
    ni, nj, nk = dims
    n = ni * nj * nk
    rng = np.random.default_rng(seed)
    arrays = {q: rng.normal(size=n) for q in ("x", "y", "z", "xc", "yc", "zc", "vol", "fx", "fy",
                                               "fz", "ar1x", "ar1y", "ar1z", "ar2x", "ar2y",
                                               "ar2z", "ar3x", "ar3y", "ar3z")}
    lens = (nj * nk,) * 2 + (ni * nk,) * 2 + (ni * nj,) * 2
    lb = [rng.integers(1, 5, m).astype(np.int32) for m in lens]
    flags = rng.integers(0, 2, 9).astype(np.int32)
    fin = tmp_path / "in.bin"
    with open(fin, "wb") as f:
        np.array(dims, np.int32).tofile(f)
        for a in lb:
            a.tofile(f)
        flags.tofile(f)
        for q in arrays:
            arrays[q].tofile(f)
    grid = tmp_path / "grid-0"
    subprocess.run([WRITER, str(fin), str(grid)], check=True, capture_output=True)
    return grid, arrays, lb, flags


@needs_writer
@pytest.mark.parametrize("dims,seed", [((7, 6, 5), 1), ((20, 17, 11), 2)])
def test_read_grid_roundtrip(cap, tmp_path, dims, seed):
    grid, arrays, lb, flags = _write(tmp_path, dims, seed)
    got_dims, got, got_lb, got_flags = cap.read_grid(str(grid))
    assert got_dims == dims
    for q, a in arrays.items():
        np.testing.assert_array_equal(got[q], a, err_msg=q)
    for q, a in zip(("lbw", "lbe", "lbs", "lbn", "lbb", "lbt"), lb):
        np.testing.assert_array_equal(got_lb[q], a, err_msg=q)
    assert [int(v) for v in got_flags.values()] == [int(v) for v in flags]


@needs_writer
def test_read_grid_refuses_bad_files(cap, tmp_path):
    import ctypes as C
    grid, _, _, _ = _write(tmp_path, (7, 6, 5), 3)
    raw = grid.read_bytes()
    cut = tmp_path / "cut"
    cut.write_bytes(raw[: len(raw) - 100])
    with pytest.raises(cap.CapError):
        cap.read_grid(str(cut))
    with pytest.raises(cap.CapError):
        cap.read_grid(str(tmp_path / "missing"))
    # extents other than the file's
    L = cap.host.load()
    arrs = (C.c_void_p * 19)()
    assert L.cap_read_grid(os.fsencode(str(grid)), 7, 6, 6, arrs, None, None) != 0


@needs_writer
def test_read_grid_of_a_generated_mesh(cap, tmp_path):
    """a rectilinear stretched mesh written and read back is the same mesh"""
    g, dims = GC.rectilinear(GC.stretched_widths(6, 1.1), GC.uniform_widths(5),
                             GC.stretched_widths(4, 1.2))
    ni, nj, nk = dims
    n = ni * nj * nk
    for q in ("fx", "fy", "fz", "ar1x", "ar1y", "ar1z", "ar2x", "ar2y", "ar2z", "ar3x", "ar3y",
              "ar3z"):
        g[q] = np.zeros(n)
    fin = tmp_path / "in.bin"
    with open(fin, "wb") as f:
        np.array(dims, np.int32).tofile(f)
        for m in (nj * nk,) * 2 + (ni * nk,) * 2 + (ni * nj,) * 2:
            np.full(m, 4, np.int32).tofile(f)
        np.zeros(9, np.int32).tofile(f)
        for q in cap.host.GRID_ARRAYS:
            g[q].tofile(f)
    grid = tmp_path / "mesh-0"
    subprocess.run([WRITER, str(fin), str(grid)], check=True, capture_output=True)
    got_dims, got, lb, flags = cap.read_grid(str(grid))
    assert got_dims == dims and not any(flags.values())
    for q in ("x", "y", "z", "xc", "yc", "zc", "vol"):
        np.testing.assert_array_equal(got[q], g[q])
