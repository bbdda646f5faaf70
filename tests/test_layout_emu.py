"""The GPU wavefront schedule and tile-sheared layout, emulated on the host.

tests/emu_wave.cpp runs the same index functions and per-step lane schedule
as the HIP kernels (csrc/cap_layout.h, csrc/cap_wave.h) sequentially; the
results must be bit-identical to the oracle's sequential loops, across tile
boundaries, both k-ranges and halo data in res / bn / be / bt.
"""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "emu_wave.cpp")
LIB = os.path.join(HERE, "libemu_wave.so")
dp = C.POINTER(C.c_double)


@pytest.fixture(scope="module")
def E():
    # built under a per-process name and moved into place, so that parallel
    # test workers never load a library another one is still writing
    tmp = f"{LIB}.{os.getpid()}"
    subprocess.run(["g++", "-O2", "-ffp-contract=off", "-fPIC", "-shared", "-std=c++17",
                    "-o", tmp, SRC], check=True)
    os.replace(tmp, LIB)
    L = C.CDLL(LIB)
    L.emu_wave.restype = C.c_double
    L.emu_wave.argtypes = [C.c_int] * 6 + [C.c_double, C.POINTER(dp), C.POINTER(dp),
                                           C.POINTER(dp)]
    return L


def P(a):
    return a.ctypes.data_as(dp)


def arr(lst):
    return (dp * len(lst))(*[P(a) for a in lst])


class Emu:
    def __init__(self, E, dims, klo, khi):
        self.E, self.dims, self.k = E, dims, (klo, khi)
        o = (C.c_int * 5)()
        E.emu_lay(*dims, klo, khi, o)
        self.n = o[4]

    def gather(self, a):
        t = np.zeros(self.n)
        self.E.emu_gather(*self.dims, *self.k, P(a), P(t))
        return t

    def scatter(self, t, base):
        out = base.copy()
        self.E.emu_scatter(*self.dims, *self.k, P(t), P(out))
        return out

    def wave(self, op, alfa, ins, deps, xos):
        return self.E.emu_wave(op, *self.dims, *self.k, alfa, arr(ins), arr(deps),
                               arr(xos if xos else [deps[0]]))


CASES = [((12, 10, 9), 3, 7), ((12, 10, 9), 2, 8), ((21, 38, 7), 3, 5), ((36, 36, 36), 2, 35),
         ((40, 23, 19), 3, 17)]


def halo_system(O, dims, seed):
    s, fi = O.build_general(dims, seed)
    rng = np.random.default_rng(seed)
    for f in ("bn", "be", "bt", "bb", "bw", "bs", "bp", "res"):
        getattr(s, f)[:] = rng.uniform(-0.2, 0.2, s.n)
    return s, fi


@pytest.mark.parametrize("dims,klo,khi", CASES)
def test_sip_factor_and_sweeps(E, O, dims, klo, khi):
    s, _ = halo_system(O, dims, 3)
    ref = s.copy()
    O.sip_factor(ref, klo, khi, alfa=0.92)
    em = Emu(E, dims, klo, khi)
    T = {f: em.gather(getattr(s, f)) for f in ("ab", "aw", "as_", "ap", "an", "ae", "at",
                                                "bn", "be", "bt")}
    xo = {f: np.zeros(em.n) for f in ("bb", "bw", "bs", "bp")}
    em.wave(0, 0.92, [T[f] for f in ("ab", "aw", "as_", "ap", "an", "ae", "at")],
            [T["bn"], T["be"], T["bt"]], [xo["bb"], xo["bw"], xo["bs"], xo["bp"]])
    for f, t in (("bn", T["bn"]), ("be", T["be"]), ("bt", T["bt"]), ("bb", xo["bb"]),
                 ("bw", xo["bw"]), ("bs", xo["bs"]), ("bp", xo["bp"])):
        np.testing.assert_array_equal(em.scatter(t, getattr(s, f)), getattr(ref, f), err_msg=f)
    rhs = np.random.default_rng(1).normal(size=s.n)
    zref = O.sip_apply(ref, klo, khi, rhs)
    F = {f: em.gather(getattr(ref, f)) for f in ("bb", "bw", "bs", "bp", "bn", "be", "bt")}
    tr, z = em.gather(rhs), np.zeros(em.n)
    em.wave(1, 0.0, [tr, F["bb"], F["bw"], F["bs"], F["bp"]], [z], None)
    dot = em.wave(2, 0.0, [z, F["bn"], F["be"], F["bt"], tr], [z], None)
    np.testing.assert_array_equal(em.scatter(z, np.zeros(s.n)), zref)
    mask = zref != 0
    assert dot == pytest.approx(np.dot(rhs[mask], zref[mask]), rel=1e-12)


@pytest.mark.parametrize("dims,klo,khi", CASES)
def test_spmv(E, O, dims, klo, khi):
    s, _ = halo_system(O, dims, 4)
    ni, nj, nk = dims
    p = np.random.default_rng(2).normal(size=s.n)
    inside = np.zeros((nk, ni, nj), bool)
    inside[klo - 1:khi, 2:ni - 2, 2:nj - 2] = True
    p[~inside.ravel()] = 0.0
    zref = O.spmv(s, klo, khi, p)
    em = Emu(E, dims, klo, khi)
    A = [em.gather(getattr(s, f)) for f in ("ae", "aw", "an", "as_", "at", "ab", "ap")]
    tz = np.zeros(em.n)
    E.emu_spmv(*dims, klo, khi, arr(A), P(em.gather(p)), P(tz))
    np.testing.assert_array_equal(em.scatter(tz, np.zeros(s.n)), zref)


def test_sipsol_sweep_with_res_halo(E, O):
    """sipsol's sweeps read res on the boundary planes (sipsol.f90:70-71,86-87)."""
    dims, klo, khi = (20, 37, 23), 3, 21
    s, fi = halo_system(O, dims, 6)
    ref, fref = s.copy(), fi.copy()
    O.sipsol(ref, fref, sor=0.0, nsw=1)
    em = Emu(E, dims, klo, khi)
    s2 = s.copy()
    O.sip_factor(s2, klo, khi, alfa=0.92)
    # residual of the starting fi in the operand order of sipsol.f90:66-68
    ni, nj, nk = dims
    nij = ni * nj
    inside = np.zeros((nk, ni, nj), bool)
    inside[klo - 1:khi, 2:ni - 2, 2:nj - 2] = True
    c = np.flatnonzero(inside.ravel())
    r = np.zeros(s.n)
    r[c] = (s2.ae[c] * fi[c + nj] + s2.aw[c] * fi[c - nj] + s2.an[c] * fi[c + 1] +
            s2.as_[c] * fi[c - 1] + s2.at[c] * fi[c + nij] + s2.ab[c] * fi[c - nij] + s2.su[c] -
            s2.ap[c] * fi[c])
    F = {f: em.gather(getattr(s2, f)) for f in ("bb", "bw", "bs", "bp", "bn", "be", "bt")}
    tres, tfi = em.gather(s.res), em.gather(fi)
    em.wave(1, 0.0, [em.gather(r), F["bb"], F["bw"], F["bs"], F["bp"]], [tres], None)
    em.wave(7, 0.0, [tres, F["bn"], F["be"], F["bt"], tfi], [tres], [tfi])
    np.testing.assert_array_equal(em.scatter(tfi, fi), fref)
    np.testing.assert_array_equal(em.scatter(tres, s.res), ref.res)


@pytest.mark.parametrize("dims,klo,khi", CASES)
@pytest.mark.parametrize("nblocks", [1, 7, 64, 2048])
def test_flat_kernel_walk(E, dims, klo, khi, nblocks):
    # the flat kernels' per-block slot walk and wrap-based neighbour indices
    assert E.emu_check_walk(*dims, klo, khi, nblocks) == 0


@pytest.mark.parametrize("dims,klo,khi", CASES + [((68, 52, 12), 3, 10), ((100, 132, 6), 2, 5)])
@pytest.mark.parametrize("nblocks", [8, 56, 512, 2048])
def test_skewed_walk(E, dims, klo, khi, nblocks):
    # the SpMV's rotated per-tile walk: coverage, planes, neighbours, and the
    # neighbour tiles' edge rows visited at the same step
    assert E.emu_check_skewed_walk(*dims, klo, khi, nblocks) == 0


MOVE_CASES = CASES + [((68, 132, 100), 3, 98), ((196, 388, 62), 3, 60), ((21, 17, 40), 2, 39)]


@pytest.mark.parametrize("dims,klo,khi", MOVE_CASES)
def test_block_moves(E, dims, klo, khi):
    """k_gather / k_scatter block by block (LDS rows of MOVE_SLOTS slots, the
    neighbouring-range rows shared): the same tiled vector and the same
    solve-range cells as the element-wise moves, every tiled element written
    once, every access inside the arrays"""
    E.emu_gather_blocks.restype = C.c_long
    E.emu_scatter_blocks.restype = C.c_long
    em = Emu(E, dims, klo, khi)
    n = dims[0] * dims[1] * dims[2]
    src = np.arange(1, n + 1, dtype=np.float64)
    ref = em.gather(src)
    got = np.full(em.n, -7.0)
    assert E.emu_gather_blocks(*dims, klo, khi, P(src), P(got)) == 0
    np.testing.assert_array_equal(got, ref)
    base = np.full(n, -3.0)
    t = np.random.default_rng(1).normal(size=em.n)
    want = em.scatter(t, base)
    back = base.copy()
    assert E.emu_scatter_blocks(*dims, klo, khi, P(t), P(back)) == 0
    np.testing.assert_array_equal(back, want)
