"""Seeded synthetic point clouds for the benchmark configurations.

The reference reads Medusa files; these generators make jittered unit-square
clouds of any size reproducibly (splitmix64 + IEEE double arithmetic in a
fixed order, so every implementation produces the same bits):
  node k = j*s + i at (i/(s-1), j/(s-1));
  interior nodes jittered by U(-a, a) in x then y, a = jitter*h;
  edge nodes jittered along the edge only; corners fixed;
  tags: 'inflow' = x == 0 or y == 0 edges, 'outflow' = the rest of the
  boundary; outward normals (corners: the diagonal).
"""
import numpy as np

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)
_G = np.uint64(0x9E3779B97F4A7C15)
_C1 = np.uint64(0xBF58476D1CE4E5B9)
_C2 = np.uint64(0x94D049BB133111EB)


def splitmix64_stream(seed, n):
    """n consecutive outputs of splitmix64 started at `seed`."""
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + _G * np.arange(1, n + 1, dtype=np.uint64)
        z = (z ^ (z >> np.uint64(30))) * _C1
        z = (z ^ (z >> np.uint64(27))) * _C2
        return z ^ (z >> np.uint64(31))


def uniform01(seed, n):
    return (splitmix64_stream(seed, n) >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)


def cloud_square(s, seed, jitter=0.3):
    """Returns x, y, tag (0 interior, 1 inflow, 2 outflow), nx, ny."""
    i = np.tile(np.arange(s), s)
    j = np.repeat(np.arange(s), s)
    bx = i.astype(np.float64) / float(s - 1)
    by = j.astype(np.float64) / float(s - 1)
    bl, br, bb, bt = i == 0, i == s - 1, j == 0, j == s - 1
    vert, horz = bl | br, bb | bt
    corner = vert & horz
    ndraw = np.where(corner, 0, np.where(vert | horz, 1, 2))
    start = np.concatenate([[0], np.cumsum(ndraw)[:-1]])
    u = uniform01(seed, int(ndraw.sum()))
    h = 1.0 / float(s - 1)
    a = jitter * h
    x, y = bx.copy(), by.copy()
    m = vert & ~horz
    y[m] = by[m] + (2.0 * u[start[m]] - 1.0) * a
    m = horz & ~vert
    x[m] = bx[m] + (2.0 * u[start[m]] - 1.0) * a
    m = ~vert & ~horz
    x[m] = bx[m] + (2.0 * u[start[m]] - 1.0) * a
    y[m] = by[m] + (2.0 * u[start[m] + 1] - 1.0) * a
    tag = np.zeros(s * s, np.int8)
    tag[vert | horz] = 2
    tag[bl | bb] = 1
    nx, ny = np.zeros(s * s), np.zeros(s * s)
    r2 = 1.0 / np.sqrt(2.0)
    nx[bl] = -1.0
    nx[br] = 1.0
    ny[bb & ~vert] = -1.0
    ny[bt & ~vert] = 1.0
    nx[corner] = np.where(br[corner], r2, -r2)
    ny[corner] = np.where(bt[corner], r2, -r2)
    return x, y, tag, nx, ny


def boundary_tags(tag, nx, ny):
    """{'inflow': (idx, normals), 'outflow': (idx, normals)} in node order."""
    out = {}
    for name, code in (("inflow", 1), ("outflow", 2)):
        idx = np.nonzero(tag == code)[0].astype(np.int32)
        out[name] = (idx, np.stack([nx[idx], ny[idx]], axis=1))
    return out


def ragged_sizes(n_nodes, seed, p_boundary=0.08, n_interior=30, lo=40, hi=50):
    """Per-stencil sizes: n_interior, or U{lo..hi} with probability p_boundary."""
    u = uniform01(seed, 2 * n_nodes)
    sel = u[0::2] < p_boundary
    n = np.full(n_nodes, n_interior, np.int64)
    n[sel] = lo + np.minimum((u[1::2][sel] * (hi - lo + 1)).astype(np.int64), hi - lo)
    return n


def gaussian_noise(seed, n, sigma):
    """Box-Muller N(0, sigma^2) samples."""
    u = uniform01(seed, 2 * n)
    u1 = np.maximum(u[0::2], 2.0 ** -53)
    return sigma * np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u[1::2])
