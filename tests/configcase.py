"""Module state of the configurations the benchmark names, for their tests:

* C1 -- the lid-driven cavity: unit cube of n^3 uniform cells, Re = 100
  (rho = 1, mu = 0.01), u, v, w ~ U[-0.1, 0.1] from a seed, walls on all six
  faces and the lid (u = 1) on the top (k = nk) face;
* C2 -- the same box with n^3 cells stretched geometrically by 1.05 toward
  every wall.

The grid's interpolation factors and face areas come from the oracle's
restatements of intfac.f90 and init.f90:49-144 (the device versions are
bit-identical, test_gpu_geometry.py).  init.f90's loops point the north and
top area vectors to -j and -k on a right-handed grid; the state takes every
face vector outward (+i, +j, +k: the north and top vectors negated), which is
what the flux routines' face coefficients and Rhie-Chow fluxes assume.

The lid is the inlet boundary type (1) on the top face with uin = 1: the
ghost fill then sets u_ghost = (1 - u_P f_P) / f_E there
(update_values_at_ghostcells_nonperiodic.f90:466-471), the lid applied as a
ghost-cell step, identically on both sides of a comparison.
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
import flowcase as FC  # noqa: E402
import gridcase as GC  # noqa: E402

GEO12 = ("fx", "fy", "fz", "ar1x", "ar1y", "ar1z", "ar2x", "ar2y", "ar2z", "ar3x", "ar3y", "ar3z")
GHOST7 = ("u", "v", "w", "te", "ed", "t", "vis")


def _geometry(O, wx, wy, wz):
    g, dims = GC.rectilinear(wx, wy, wz)
    n = dims[0] * dims[1] * dims[2]
    for q in GEO12:
        g[q] = np.zeros(n)
    O.face_areas(dims, g["x"], g["y"], g["z"], g)
    for q in ("ar2x", "ar2y", "ar2z", "ar3x", "ar3y", "ar3z"):
        g[q] = -g[q]
    g["fx"], g["fy"], g["fz"] = O.intfac(dims, g)
    return g, dims


def _interior(dims):
    ni, nj, nk = dims
    m = np.zeros((nk, ni, nj), bool)
    m[2:-2, 2:-2, 2:-2] = True
    return m.ravel()


def initial_fluxes(dims, f):
    """Mass fluxes of the interior faces from the linearly interpolated
    velocity, rho (u_f . S); faces on the boundary carry none (walls)."""
    ni, nj, nk = dims
    sh = (nk, ni, nj)
    U = {q: f[q].reshape(sh) for q in ("u", "v", "w", "den", "fx", "fy", "fz", "ar1x", "ar1y",
                                       "ar1z", "ar2x", "ar2y", "ar2z", "ar3x", "ar3y", "ar3z")}
    out = {}
    for name, ax, lam, ar in (("f1", 1, "fx", "ar1"), ("f2", 2, "fy", "ar2"), ("f3", 0, "fz", "ar3")):
        # axis of the direction in (k, i, j) order, and its extent
        size = (nk, ni, nj)[ax]
        F = np.zeros(sh)
        Pi = [slice(2, nk - 2), slice(2, ni - 2), slice(2, nj - 2)]
        E = list(Pi)
        Pi[ax] = slice(2, size - 3)          # P and its neighbour E both interior
        E[ax] = slice(3, size - 2)
        E, Pi = tuple(E), tuple(Pi)
        fe = U[lam][Pi]
        face = {q: fe * U[q][E] + (1.0 - fe) * U[q][Pi] for q in ("u", "v", "w", "den")}
        F[Pi] = face["den"] * (face["u"] * U[ar + "x"][Pi] + face["v"] * U[ar + "y"][Pi] +
                               face["w"] * U[ar + "z"][Pi])
        out[name] = F.ravel()
    return out


def state(O, kind, n, seed, fill=True):
    """(dims, state dict, boundary types lb [6], inlet values) of C1 (kind
    "cavity") or C2 (kind "stretched") with n^3 interior cells; fill: the
    ghost cells already set by the oracle's ghost fill (else left for the
    caller's own fill)."""
    if kind == "cavity":
        w = GC.uniform_widths(n)
        g, dims = _geometry(O, w, w, w)
    else:
        w = GC.stretched_widths(n, 1.05)
        g, dims = _geometry(O, w, w, w)
    f = FC.flow_case(dims, seed)              # every array the drivers take, with its shape
    for q in g:
        f[q] = g[q]
    N = dims[0] * dims[1] * dims[2]
    rng = np.random.default_rng(seed)
    inside = _interior(dims)
    for q in ("u", "v", "w"):
        a = np.zeros(N)
        a[inside] = rng.uniform(-0.1, 0.1, int(inside.sum()))
        f[q] = a
        for o in ("o", "oo"):
            f[q + o] = a.copy()
    f["den"] = np.ones(N)
    f["vis"] = np.full(N, 0.01)
    f["p"] = np.zeros(N)
    f["t"] = np.ones(N)
    f["te"] = np.full(N, 1e-4)
    f["ed"] = np.full(N, 1e-4)
    for q in ("pp", "ae", "aw", "an", "as", "at", "ab", "su", "sv", "sw", "sp", "res", "bb", "bs",
              "bw", "bp", "bn", "be", "bt", "apu", "apv", "apw"):
        f[q] = np.zeros(N)
    f["ap"] = np.ones(N)
    for q in ("gradu", "gradv", "gradw", "gradp"):
        f[q] = np.zeros(3 * N)
    ni, nj, nk = dims
    lens = (nj * nk,) * 2 + (ni * nk,) * 2 + (ni * nj,) * 2
    lb = [np.full(m, 4, np.int32) for m in lens]
    inl = np.zeros(6)
    if kind == "cavity":
        lb[5][:] = 1                          # the lid: inlet type on the top face ...
        inl[0] = 1.0                          # ... with uin = 1
    if fill:
        O.ghostcells(dims, [f[q] for q in GHOST7], [f[q] for q in GEO12], lb, inl)
    f.update(initial_fluxes(dims, f))
    return dims, f, lb, inl


def uvw_opts(cls, dims, schemes=1, gds=1.0):
    """calcuvw settings of a steady SIMPLE iteration: urf 0.7, sipsol to 0.2
    of the initial residual in at most 5 sweeps (the reference's usual
    momentum settings)"""
    q = cls()
    q.schemes, q.gds = schemes, gds
    for m in range(3):
        q.urfr[m], q.urfm[m], q.sor[m], q.nsw[m] = 1 / 0.7, 0.3, 0.2, 5
    q.nx, q.ny, q.nz = dims[0] - 2, dims[1] - 2, dims[2] - 2
    q.gradpcmf = 0.0
    return q


def drv_opts(cls, dims, **kw):
    """steady SIMPLE pressure correction: urf(p) = 0.3, the reference
    pressure cell (3, 3, 3), one corrector"""
    base = dict(ncorr=1, npcor=1, solver=1, nsw=2000, urf=0.3, sor=1e-8, alfa=0.92,
                ijkpr=FC.ref_cell(dims), const_mflux=1, densit=1.0, gravz=0.0, timestep=1.0)
    base.update(kw)
    return FC.opts(cls, **base)


def momentum_and_assembly(O, dims, f):
    """C2's p' system from the oracle: calcuvw (momentum, sipsol), then
    calcp's assembly with one PCG-SIP iteration (the assembled ae..ap, su
    stay in f; the solve itself is run separately)."""
    spv = np.zeros_like(f["u"])
    with O.plane_sums(2):
        O.calcuvw(dims, f, spv, drv_opts(O.DrvOpts, dims), uvw_opts(O.UvwOpts, dims))
        O.calcp(dims, f, drv_opts(O.DrvOpts, dims, solver=4, alfa=0.0, nsw=1))
    return spv


SYSTEM = ("ae", "aw", "an", "as", "at", "ab", "ap", "su")


def p_system(O, dims, f):
    """The assembled p' system as an oracle System, with identity rows
    (a_nb = 0, ap = 1, su = 0) on the ghost planes k = 2 and nkm that iccg
    and cgstab_sip include in their range: there calcp assembles nothing and
    the arrays hold calcuvw's w-equation, so each solver sees p' = 0 on the
    boundary planes, as pcgsip and sipsol do."""
    s = O.System(dims)
    for q in SYSTEM:
        getattr(s, "as_" if q == "as" else q)[:] = f[q]
    ni, nj, nk = dims
    sh = (nk, ni, nj)
    for q in ("ae", "aw", "an", "as_", "at", "ab", "su"):
        a = getattr(s, q).reshape(sh)
        a[1] = 0.0
        a[nk - 2] = 0.0
    ap = s.ap.reshape(sh)
    ap[1] = 1.0
    ap[nk - 2] = 1.0
    return s
