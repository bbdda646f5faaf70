"""Seeded module state for the pressure-correction driver tests.

A rectilinear, stretched and slightly jittered mesh (i -> x, j -> y, k -> z in
the reference's padded j-fastest layout) with consistent node and centre
coordinates, face areas (plus small non-orthogonal components), interpolation
factors and volumes, and a smooth velocity field with noise.  Coefficient and
work arrays start at zero, except ap = 1e3: the solvers whose range includes
the boundary planes k = 2 and nkm (iccg, cgstab_sip) see there the face
coefficients fluxmass writes and the ap the drivers do not assemble, and a
large ap keeps those rows diagonally dominant.
"""
from __future__ import annotations

import numpy as np

# names of the driver state arrays that the tests compare after a call
OUTPUTS = ("u", "v", "w", "p", "pp", "f1", "f2", "f3", "ae", "aw", "an", "as", "at", "ab", "ap",
           "su", "sv", "sw", "sp", "res", "bb", "bs", "bw", "bp", "bn", "be", "bt",
           "he", "hw", "hn", "hs", "ht", "hb", "gradu", "gradv", "gradw", "gradp",
           "t", "te", "ed", "vis")


def _axis(n, rng, stretch):
    """node coordinates X[0..n-1] (node m closes cell m) and centres"""
    h = stretch ** np.arange(n) * (1.0 + 0.15 * rng.uniform(-1, 1, n))
    X = np.cumsum(h)
    xc = X - 0.5 * h
    return X, xc, h


def flow_case(dims, seed, stretch=1.03):
    ni, nj, nk = dims
    rng = np.random.default_rng(seed)
    X, XC, DX = _axis(ni, rng, stretch)
    Y, YC, DY = _axis(nj, rng, 1.0 / stretch)
    Z, ZC, DZ = _axis(nk, rng, stretch)
    k, i, j = np.meshgrid(np.arange(nk), np.arange(ni), np.arange(nj), indexing="ij")
    f = {}
    f["x"], f["y"], f["z"] = X[i].ravel(), Y[j].ravel(), Z[k].ravel()
    f["xc"], f["yc"], f["zc"] = XC[i].ravel(), YC[j].ravel(), ZC[k].ravel()
    f["vol"] = (DX[i] * DY[j] * DZ[k]).ravel()

    def fac(C, P, idx, n):
        nxt = np.minimum(idx + 1, n - 1)
        w = (P[idx] - C[idx]) / np.where(nxt > idx, C[nxt] - C[idx], 1.0)
        return np.where(nxt > idx, w, 0.5)

    f["fx"] = fac(XC, X, i, ni).ravel()
    f["fy"] = fac(YC, Y, j, nj).ravel()
    f["fz"] = fac(ZC, Z, k, nk).ravel()
    s = lambda: rng.uniform(-0.01, 0.01, (nk, ni, nj))  # noqa: E731
    a1, a2, a3 = DY[j] * DZ[k], DX[i] * DZ[k], DX[i] * DY[j]
    f["ar1x"], f["ar1y"], f["ar1z"] = a1.ravel(), (s() * a1).ravel(), (s() * a1).ravel()
    f["ar2x"], f["ar2y"], f["ar2z"] = (s() * a2).ravel(), a2.ravel(), (s() * a2).ravel()
    f["ar3x"], f["ar3y"], f["ar3z"] = (s() * a3).ravel(), (s() * a3).ravel(), a3.ravel()
    n = ni * nj * nk
    f["den"] = rng.uniform(0.95, 1.05, n)
    for q in ("apu", "apv", "apw"):
        f[q] = rng.uniform(0.5, 1.5, n)
    xs, ys, zs = X[i] / X[-1], Y[j] / Y[-1], Z[k] / Z[-1]
    f["u"] = (1.0 + 0.3 * np.sin(2 * np.pi * ys) * np.cos(np.pi * zs)).ravel() + 0.05 * rng.normal(size=n)
    f["v"] = (0.2 * np.sin(2 * np.pi * xs) * np.sin(np.pi * zs)).ravel() + 0.05 * rng.normal(size=n)
    f["w"] = (0.2 * np.cos(2 * np.pi * xs) * np.sin(np.pi * ys)).ravel() + 0.05 * rng.normal(size=n)
    f["p"] = (0.5 * np.cos(np.pi * xs) * np.cos(np.pi * ys)).ravel()
    for q in ("u", "v", "w"):
        f[q + "o"] = f[q] + 0.02 * rng.normal(size=n)
        f[q + "oo"] = f[q] + 0.04 * rng.normal(size=n)
    for q in ("pp", "f1", "f2", "f3", "ae", "aw", "an", "as", "at", "ab", "su", "sv", "sw", "sp",
              "res", "bb", "bs", "bw", "bp", "bn", "be", "bt"):
        f[q] = np.zeros(n)
    for q in ("he", "hw", "hn", "hs", "ht", "hb"):     # PISO's copies of a_nb (for getHbyA alone)
        f[q] = rng.uniform(0.1, 1.0, n)
    f["ap"] = np.full(n, 1e3)
    for q in ("gradu", "gradv", "gradw", "gradp"):
        f[q] = np.zeros(3 * n)
    f["t"] = 1.0 + 0.1 * rng.normal(size=n)
    f["te"], f["ed"] = rng.uniform(0.01, 0.1, n), rng.uniform(0.01, 0.1, n)
    f["vis"] = rng.uniform(1e-3, 2e-3, n)
    return f


def boundary_types(dims, east=2, others=4):
    """lbw..lbt (int32, indexed as the reference's inbc) with every east
    boundary cell of type `east` and the rest `others`"""
    ni, nj, nk = dims
    lb = {}
    for name, m in (("lbw", nj * nk), ("lbe", nj * nk), ("lbs", ni * nk), ("lbn", ni * nk),
                    ("lbb", ni * nj), ("lbt", ni * nj)):
        lb[name] = np.full(m, east if name == "lbe" else others, dtype=np.int32)
    return lb


def opts(Opts, **kw):
    """a driver settings struct with the defaults used by the tests"""
    o = Opts()
    o.ncorr, o.npcor, o.solver, o.nsw = 1, 1, 4, 400
    o.urf, o.sor, o.alfa = 0.3, 1e-8, 0.0
    o.ijkpr, o.const_mflux = 1, 1
    o.lout = o.low = o.loe = o.los = o.lon = o.lob = o.lot = 0
    o.flowin = 0.0
    o.cn = o.bdf = o.buoy = o.boussinesq = 0
    o.timestep, o.btime, o.densit, o.beta, o.tref = 0.01, 1.0, 1.0, 1e-3, 1.0
    o.gravx, o.gravy, o.gravz, o.cumulative = 0.0, 0.0, -9.81, 0.0
    for key, val in kw.items():
        setattr(o, key, val)
    return o


def ref_cell(dims, i=3, j=3, k=3):
    """1-based index lk(k) + li(i) + j"""
    ni, nj, nk = dims
    return (k - 1) * ni * nj + (i - 1) * nj + j
