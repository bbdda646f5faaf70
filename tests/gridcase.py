"""Structured grids for the geometry and configuration tests.

The reference reads its grid (nodes, centres, volumes, interpolation
factors, face areas) from a file (read_grid.f90:126-161) written by an
external generator; these helpers make such grids in the reference's padded
j-fastest layout (ijk = (k-1) ni nj + (i-1) nj + j): node ijk is the corner
that closes cell ijk (the one at the cell's largest i, j, k), and every
padded cell, the two ghost layers included, has a width of its own.
"""
from __future__ import annotations

import numpy as np


def stretched_widths(ncells, ratio, length=1.0):
    """Cell widths of an interval of `length` split into ncells cells that
    grow geometrically by `ratio` from each end toward the middle
    (symmetric), plus two ghost cells at each end mirroring the first
    interior cells."""
    m = np.arange(ncells)
    h = ratio ** np.minimum(m, ncells - 1 - m).astype(np.float64)
    h = h * (length / h.sum())
    return np.concatenate([h[1::-1], h, h[:-3:-1]])


def tanh_widths(ncells, factor, length=1.0):
    """Cell widths of an interval clustered toward both ends by a tanh map
    with the given stretching factor (the channel's wall-normal direction),
    plus two mirrored ghost cells at each end."""
    s = np.linspace(-1.0, 1.0, ncells + 1)
    y = 0.5 * length * (1.0 + np.tanh(factor * s) / np.tanh(factor))
    h = np.diff(y)
    return np.concatenate([h[1::-1], h, h[:-3:-1]])


def uniform_widths(ncells, length=1.0):
    return np.full(ncells + 4, length / ncells)


def rectilinear(wx, wy, wz):
    """Nodes, centres and volumes of the tensor-product grid whose padded
    cells (i = 1..ni etc.) have widths wx, wy, wz; the interior starts at
    coordinate 0 (the first interior cell, i = 3, spans [0, wx[2]])."""
    X = np.cumsum(wx) - wx[0] - wx[1]
    Y = np.cumsum(wy) - wy[0] - wy[1]
    Z = np.cumsum(wz) - wz[0] - wz[1]
    ni, nj, nk = len(wx), len(wy), len(wz)
    k, i, j = np.meshgrid(np.arange(nk), np.arange(ni), np.arange(nj), indexing="ij")
    g = {"x": X[i].ravel(), "y": Y[j].ravel(), "z": Z[k].ravel(),
         "xc": (X - 0.5 * wx)[i].ravel(), "yc": (Y - 0.5 * wy)[j].ravel(),
         "zc": (Z - 0.5 * wz)[k].ravel(), "vol": (wx[i] * wy[j] * wz[k]).ravel()}
    return g, (ni, nj, nk)


def lattice(dims, a, b, c, jitter=0.0, seed=0):
    """Nodes of the lattice x(i, j, k) = i a + j b + k c (a parallelepiped
    grid when jitter = 0), each node moved by a uniform random vector of up
    to `jitter` per component; centres are the means of the eight corners
    of each cell (ghost cells without eight corners: the node less half a
    lattice step)."""
    ni, nj, nk = dims
    rng = np.random.default_rng(seed)
    k, i, j = np.meshgrid(np.arange(nk), np.arange(ni), np.arange(nj), indexing="ij")
    P = (i[..., None] * np.asarray(a, float) + j[..., None] * np.asarray(b, float) +
         k[..., None] * np.asarray(c, float))
    P = P + rng.uniform(-jitter, jitter, P.shape)
    C = P - 0.5 * (np.asarray(a) + np.asarray(b) + np.asarray(c))
    C[1:, 1:, 1:] = 0.125 * (P[1:, 1:, 1:] + P[:-1, 1:, 1:] + P[1:, :-1, 1:] + P[1:, 1:, :-1] +
                             P[:-1, :-1, 1:] + P[:-1, 1:, :-1] + P[1:, :-1, :-1] + P[:-1, :-1, :-1])
    g = {q: np.ascontiguousarray(P[..., d]).ravel() for d, q in enumerate("xyz")}
    g.update({q + "c": np.ascontiguousarray(C[..., d]).ravel() for d, q in enumerate("xyz")})
    return g
