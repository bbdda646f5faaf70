"""Pins of the oracle's geometry routines (test infrastructure):

* find_intersection_point: the reference's own known answers
  (test_intersection_point.f90:8-38), points on both the plane and the line
  for random inputs, and bit-identity with the reference build;
* intfac: bit-identity with the reference build on a distorted grid, and the
  closed form (X_i - xc_i) / (xc_E - xc_i) on rectilinear grids;
* face areas (init.f90:49-144; not buildable alone, the loop sits inside
  init's large setup routine): on a parallelepiped lattice every face vector
  is the cross product of the lattice vectors that span it, and on any
  distorted grid the six outward face vectors of a cell sum to zero.
"""
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
import gridcase as GC  # noqa: E402

REF = os.path.join(os.path.dirname(HERE), "oracle", "_ref", "ref_driver")
needs_ref = pytest.mark.skipif(not os.path.exists(REF),
                               reason="reference build only exists where the reference sources are")


def test_intersection_known_answers(O):
    # test_intersection_point.f90: rectilinear, then the same plane spanned
    # by three other points; both cross the line of centres at (1, .5, .5)
    p = O.find_intersection_point([1, 1, 1], [1, 0, 1], [1, 1, 0], [.5, .5, .5], [1.5, .5, .5])
    np.testing.assert_allclose(p, [1.0, 0.5, 0.5], rtol=0, atol=1e-15)
    p = O.find_intersection_point([1, 1.25, 1], [1, .25, 1], [1, 1.25, 0], [.5, .5, .5],
                                  [1.5, .5, .5])
    np.testing.assert_allclose(p, [1.0, 0.5, 0.5], rtol=0, atol=1e-15)


def test_intersection_on_plane_and_line(O):
    rng = np.random.default_rng(3)
    for _ in range(200):
        p1, p2, p3, p4, p5 = rng.uniform(-1, 1, (5, 3))
        q = O.find_intersection_point(p1, p2, p3, p4, p5)
        nrm = np.cross(p2 - p1, p3 - p1)
        scale = np.linalg.norm(nrm) * max(1.0, np.linalg.norm(q - p1))
        assert abs(np.dot(q - p1, nrm)) < 1e-11 * scale
        d = p5 - p4
        r = np.cross(q - p4, d)
        assert np.linalg.norm(r) < 1e-11 * np.linalg.norm(d) * max(1.0, np.linalg.norm(q - p4))


@needs_ref
def test_intersection_bitwise_vs_reference_build(O):
    rng = np.random.default_rng(4)
    m = 300
    pts = rng.uniform(-2, 2, (m, 15))
    dims = (20, 20, 20)                    # x and y of the driver hold the point sets
    x = np.zeros(dims[0] * dims[1] * dims[2])
    x[:pts.size] = pts.ravel()
    outs, _, _ = O.run_ref("fip", dims, [x], nsw=m)
    got = np.array([O.find_intersection_point(*p.reshape(5, 3)) for p in pts])
    np.testing.assert_array_equal(got.ravel(), outs[0])


def _distorted(dims, seed):
    g = GC.lattice(dims, (1.0, 0.1, 0.0), (0.05, 0.9, 0.1), (0.0, 0.2, 1.1), jitter=0.15, seed=seed)
    n = dims[0] * dims[1] * dims[2]
    rng = np.random.default_rng(seed + 1)
    for q in ("fx", "fy", "fz"):
        g[q] = rng.uniform(0, 1, n)        # values outside the loop range must survive
    return g


@needs_ref
@pytest.mark.parametrize("dims,seed", [((9, 8, 7), 1), ((14, 11, 12), 2)])
def test_intfac_bitwise_vs_reference_build(O, dims, seed):
    g = _distorted(dims, seed)
    outs, _, _ = O.run_ref("intfac", dims, [g[q] for q in ("x", "y", "z", "xc", "yc", "zc",
                                                            "fx", "fy", "fz")])
    for a, b, q in zip(O.intfac(dims, g), outs, ("fx", "fy", "fz")):
        np.testing.assert_array_equal(a, b, err_msg=q)


def test_intfac_rectilinear_closed_form(O):
    wx, wy, wz = GC.stretched_widths(10, 1.1), GC.uniform_widths(7), GC.stretched_widths(6, 1.3)
    g, dims = GC.rectilinear(wx, wy, wz)
    n = dims[0] * dims[1] * dims[2]
    for q in ("fx", "fy", "fz"):
        g[q] = np.full(n, -1.0)
    fx, fy, fz = O.intfac(dims, g)
    ni, nj, nk = dims
    sh = (nk, ni, nj)
    X = np.cumsum(wx) - wx[0] - wx[1]
    xc = X - 0.5 * wx
    want = (X[1:-1] - xc[1:-1]) / (xc[2:] - xc[1:-1])       # cells i = 2..nim
    got = fx.reshape(sh)[1:-1, 1:-1, 1:-1]
    # exact in exact arithmetic; the determinants round (relative 5e-14 seen)
    np.testing.assert_allclose(got, np.broadcast_to(want[None, :, None], got.shape), rtol=1e-12)
    assert np.all(fx.reshape(sh)[0] == -1.0) and np.all(fx.reshape(sh)[:, :, 0] == -1.0)
    np.testing.assert_allclose(fy.reshape(sh)[1:-1, 1:-1, 1:-1], 0.5, rtol=1e-12)


def _areas(O, dims, g, fill=np.nan):
    n = dims[0] * dims[1] * dims[2]
    ar = {q: np.full(n, fill) for q in O.AR_FIELDS}
    O.face_areas(dims, g["x"], g["y"], g["z"], ar)
    return ar


def test_face_areas_parallelepiped(O):
    """east face: its corner's j step b and k step c -> b x c; north: i step
    a and k step c -> a x c; top: j step b and i step a -> b x a (the
    orientations init.f90's three loops give: on a right-handed grid the
    north and top vectors point to -j and -k)"""
    a, b, c = np.array([1.0, 0.2, -0.1]), np.array([0.15, 0.8, 0.05]), np.array([-0.1, 0.3, 1.2])
    dims = (9, 8, 7)
    g = GC.lattice(dims, a, b, c)
    ar = _areas(O, dims, g)
    ni, nj, nk = dims
    sh = (nk, ni, nj)
    rng_e = (slice(1, nk - 1), slice(0, ni - 1), slice(1, nj - 1))
    rng_n = (slice(1, nk - 1), slice(1, ni - 1), slice(0, nj - 1))
    rng_t = (slice(0, nk - 1), slice(1, ni - 1), slice(1, nj - 1))
    for pre, rg, want in (("ar1", rng_e, np.cross(b, c)), ("ar2", rng_n, np.cross(a, c)),
                          ("ar3", rng_t, np.cross(b, a))):
        for d, q in enumerate("xyz"):
            got = ar[pre + q].reshape(sh)[rg]
            np.testing.assert_allclose(got, want[d], rtol=0, atol=1e-13, err_msg=pre + q)
    # outside the loop ranges nothing is written
    assert np.all(np.isnan(ar["ar1x"].reshape(sh)[0])) and np.all(np.isnan(ar["ar3x"].reshape(sh)[:, 0]))


def test_face_areas_closed_cells(O):
    """the six face vectors of every interior cell of a distorted grid, each
    taken outward (east +ar1, north -ar2, top -ar3 by the orientation of the
    loops), sum to zero: exact for the mean-edge face vectors of a
    hexahedron with bilinear faces, rounding only"""
    dims = (10, 9, 8)
    g = GC.lattice(dims, (1.0, 0.0, 0.1), (0.1, 1.0, 0.0), (0.0, 0.1, 1.0), jitter=0.2, seed=5)
    ar = _areas(O, dims, g, fill=0.0)
    ni, nj, nk = dims
    sh = (nk, ni, nj)
    for d in "xyz":
        e, nn, t = (ar[f"ar{m}{d}"].reshape(sh) for m in (1, 2, 3))
        s = (e[2:-1, 2:-1, 2:-1] - e[2:-1, 1:-2, 2:-1] - nn[2:-1, 2:-1, 2:-1] + nn[2:-1, 2:-1, 1:-2] -
             t[2:-1, 2:-1, 2:-1] + t[1:-2, 2:-1, 2:-1])
        assert np.max(np.abs(s)) < 1e-14


def test_face_areas_rectilinear(O):
    wx, wy, wz = GC.stretched_widths(8, 1.05), GC.stretched_widths(6, 1.2), GC.uniform_widths(5)
    g, dims = GC.rectilinear(wx, wy, wz)
    ar = _areas(O, dims, g)
    ni, nj, nk = dims
    sh = (nk, ni, nj)
    e = ar["ar1x"].reshape(sh)[1:-1, :-1, 1:-1]
    np.testing.assert_allclose(e, (wz[1:-1, None, None] * wy[None, None, 1:-1]) *
                               np.ones((1, ni - 1, 1)), rtol=1e-13)
    for q in ("ar1y", "ar1z", "ar2x", "ar2z", "ar3x", "ar3y"):
        v = ar[q]
        assert np.all(v[np.isfinite(v)] == 0.0), q
