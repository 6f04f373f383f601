"""Particle orders of the library's tile sort, restated in numpy.

TEST INFRASTRUCTURE ONLY: imported by tests/, never by the product package
gample_amd/.

The tile layout is this library's own (the reference pushes its particles in
storage order and never sorts them), so these functions restate the
library's definition; tests/test_sortspec.py pins them against a stable
sort and brute force.
"""
from __future__ import annotations

import numpy as np

GP_TILE = 16


def tiles(x, y, ngx, ngy, ngb):
    """tile coordinates (tx, ty) of positions, cells clamped to the grid"""
    ci = np.clip(np.floor(x).astype(np.int64) - ngb, 0, ngx - 1)
    cj = np.clip(np.floor(y).astype(np.int64) - ngb, 0, ngy - 1)
    return ci // GP_TILE, cj // GP_TILE


def rebin_groups(x, y, tile_off, n_binned, ngx, ngy, ngb):
    """(tile, group) of every particle for the neighbour re-bin: group
    k = 3 (ey + 1) + ex + 1 for a particle of segment s now in tile t with
    e = s - t in -1..1 on both axes (tile coordinates, periodic wrap), 9 for
    the others and for the unbinned tail (index >= n_binned)."""
    ntx, nty = -(-ngx // GP_TILE), -(-ngy // GP_TILE)
    assert ntx >= 3 and nty >= 3
    n = len(x)
    tx, ty = tiles(x, y, ngx, ngy, ngb)
    t = ty * ntx + tx
    off = np.minimum(np.asarray(tile_off, np.int64), n_binned)
    idx = np.arange(n)
    seg = np.searchsorted(off, idx, side="right") - 1      # last segment starting at or before i
    seg = np.clip(seg, 0, ntx * nty - 1)
    ex = seg % ntx - tx
    ey = seg // ntx - ty
    ex = np.where(ex == ntx - 1, -1, np.where(ex == 1 - ntx, 1, ex))
    ey = np.where(ey == nty - 1, -1, np.where(ey == 1 - nty, 1, ey))
    near = (idx < n_binned) & (np.abs(ex) <= 1) & (np.abs(ey) <= 1)
    k = np.where(near, 3 * (ey + 1) + ex + 1, 9)
    return t, k


def rebin_order(x, y, tile_off, n_binned, ngx, ngy, ngb):
    """Permutation of the neighbour re-bin: new position p holds old index
    order[p].  Tiles in order; inside a tile its near arrivals grouped by
    source offset (row-major from (-1, -1)), then its far arrivals, each group
    in the old order."""
    t, k = rebin_groups(x, y, tile_off, n_binned, ngx, ngy, ngb)
    return np.lexsort((np.arange(len(x)), k, t))


def tile_offsets(x, y, ngx, ngy, ngb):
    """first position of every tile (and the count at the end) of positions
    grouped by tile"""
    ntx, nty = -(-ngx // GP_TILE), -(-ngy // GP_TILE)
    tx, ty = tiles(x, y, ngx, ngy, ngb)
    c = np.bincount(ty * ntx + tx, minlength=ntx * nty)
    return np.r_[0, np.cumsum(c)].astype(np.int64)
