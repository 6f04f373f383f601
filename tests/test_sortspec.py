"""The numpy statement of the neighbour re-bin's permutation
(oracle/sortspec.py), pinned against a stable sort where the two must agree
and against brute-force enumeration of the source tiles elsewhere."""
import numpy as np

from oracle import sortspec as ss

T = ss.GP_TILE


def binned_state(rng, n, ngx, ngy, ngb):
    """positions grouped by tile (a stable sort by tile) and their offsets"""
    x = (ngb + rng.random(n) * ngx).astype(np.float32)
    y = (ngb + rng.random(n) * ngy).astype(np.float32)
    ntx = -(-ngx // T)
    tx, ty = ss.tiles(x, y, ngx, ngy, ngb)
    o = np.argsort(ty * ntx + tx, kind="stable")
    x, y = x[o], y[o]
    return x, y, ss.tile_offsets(x, y, ngx, ngy, ngb)


def test_interior_moves_give_the_stable_sort():
    """moves of at most 6 cells that stay inside the domain: every particle is
    near, no move wraps, and the re-bin is the stable sort by tile"""
    rng = np.random.default_rng(5)
    ngx, ngy, ngb = 80, 64, 1
    x, y, off = binned_state(rng, 20000, ngx, ngy, ngb)
    x1 = np.clip(x + rng.uniform(-6, 6, x.size), ngb, ngb + ngx - 1e-3).astype(np.float32)
    y1 = np.clip(y + rng.uniform(-6, 6, y.size), ngb, ngb + ngy - 1e-3).astype(np.float32)
    ntx = -(-ngx // T)
    tx, ty = ss.tiles(x1, y1, ngx, ngy, ngb)
    ref = np.argsort(ty * ntx + tx, kind="stable")
    got = ss.rebin_order(x1, y1, off, x1.size, ngx, ngy, ngb)
    np.testing.assert_array_equal(got, ref)
    assert (ref != np.arange(ref.size)).any()


def brute_order(x, y, off, n_binned, ngx, ngy, ngb):
    """tile by tile: the nine source tiles in row-major offset order, each
    segment's particles now in the tile in index order, then the far ones"""
    ntx, nty = -(-ngx // T), -(-ngy // T)
    n = len(x)
    tile = []
    for i in range(n):
        ci = min(max(int(np.floor(x[i])) - ngb, 0), ngx - 1)
        cj = min(max(int(np.floor(y[i])) - ngb, 0), ngy - 1)
        tile.append((cj // T) * ntx + ci // T)
    seg = [-1] * n
    for s in range(ntx * nty):
        for i in range(min(off[s], n_binned), min(off[s + 1], n_binned)):
            seg[i] = s
    out = []
    for t in range(ntx * nty):
        tx, ty = t % ntx, t // ntx
        taken = set()
        for ey in (-1, 0, 1):
            for ex in (-1, 0, 1):
                s = ((ty + ey) % nty) * ntx + (tx + ex) % ntx
                for i in range(n):
                    if tile[i] == t and seg[i] == s:
                        out.append(i)
                        taken.add(i)
        out.extend(i for i in range(n) if tile[i] == t and i not in taken)
    return np.array(out)


def test_wrap_far_and_tail_against_brute_force():
    rng = np.random.default_rng(9)
    ngx, ngy, ngb = 48, 56, 2            # 3 x 4 tiles: the smallest grid the re-bin takes
    x, y, off = binned_state(rng, 3000, ngx, ngy, ngb)
    # moves up to 12 cells wrap across the periodic edges; a tenth jump far
    d = rng.uniform(-12, 12, (2, x.size))
    far = rng.random(x.size) < 0.1
    d[:, far] = rng.uniform(-40, 40, (2, far.sum()))
    x1 = (ngb + np.mod(x + d[0] - ngb, ngx)).astype(np.float32)
    y1 = (ngb + np.mod(y + d[1] - ngb, ngy)).astype(np.float32)
    # a tail of 300 unbinned particles
    tx_ = (ngb + rng.random(300) * ngx).astype(np.float32)
    ty_ = (ngb + rng.random(300) * ngy).astype(np.float32)
    x1, y1 = np.r_[x1, tx_], np.r_[y1, ty_]
    got = ss.rebin_order(x1, y1, off, 3000, ngx, ngy, ngb)
    np.testing.assert_array_equal(got, brute_order(x1, y1, off, 3000, ngx, ngy, ngb))
    t, k = ss.rebin_groups(x1, y1, off, 3000, ngx, ngy, ngb)
    assert (k == 9).sum() > 300                    # far movers besides the tail
    assert np.all(np.diff(t[got]) >= 0)            # grouped by tile
    # wrapped near moves exist (a particle of an edge segment now at the far edge)
    seg = np.searchsorted(np.minimum(off, 3000), np.arange(3000), side="right") - 1
    ntx = -(-ngx // T)
    assert ((k[:3000] < 9) & (np.abs(seg % ntx - t[:3000] % ntx) == ntx - 1)).any()


def test_compacted_segments():
    """segments cut by a compaction (n_binned below tile_off[ntiles]) and
    empty segments"""
    rng = np.random.default_rng(13)
    ngx, ngy, ngb = 64, 48, 1
    x, y, off = binned_state(rng, 2500, ngx, ngy, ngb)
    keep = 2000
    x1 = (x + rng.uniform(-3, 3, x.size)).astype(np.float32)[:keep]
    y1 = (y + rng.uniform(-3, 3, y.size)).astype(np.float32)[:keep]
    x1 = (ngb + np.mod(x1 - ngb, ngx)).astype(np.float32)
    y1 = (ngb + np.mod(y1 - ngb, ngy)).astype(np.float32)
    got = ss.rebin_order(x1, y1, off, keep, ngx, ngy, ngb)
    np.testing.assert_array_equal(got, brute_order(x1, y1, off, keep, ngx, ngy, ngb))
