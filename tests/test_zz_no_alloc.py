"""gp_reduce_noalloc (the compaction with its scratch in the library's
static device array) against gp_reduce through the C-ABI: two species hold
the same particles, take the same push across absorbing walls, and are
compacted one by each entry point; the survivors (in order), the per-wall
counts and the kinetic energy agree to the bit.  Two calls in a row, the
second with fewer particles, so the shared array's contents from the first
call must not leak into the second."""
import numpy as np
import pytest

from test_gpu_parity import EC, ME, mkctx, rand_parts, say

pytestmark = pytest.mark.gpu


def test_reduce_noalloc_equals_reduce(gp):
    ngx, ngy, dx, dt = 32, 32, 0.01, 1e-12
    ctx = mkctx(gp, ngx, ngy, 3, pbc=("Delete", "Delete", "periodic", "periodic"),
                fbc=("Dirichlet", "Dirichlet", "periodic", "periodic"), dx=dx)
    sa = ctx.add_species(-EC, ME, 1.0, 500000)
    sb = ctx.add_species(-EC, ME, 1.0, 500000)
    for call, (n, seed) in enumerate(((400000, 21), (100000, 22))):
        say(f"[noalloc] call {call}: {n} particles")
        p = rand_parts(n, ngx, ngy, 1, vs=1e10, seed=seed)
        for s in (sa, sb):
            ctx.upload(s, *p)
            ctx.sort(s)
        ctx.reset_grids()
        ctx.push(sa, dt, deposit=1)
        ctx.push(sb, dt, deposit=1)
        mid_a, mid_b = ctx.download(sa), ctx.download(sb)
        for a, b in zip(mid_a, mid_b):
            np.testing.assert_array_equal(a.view(np.uint32), b.view(np.uint32))
        nd = int((mid_a[0] < 0).sum())
        assert nd > 2048, nd          # more than one scan block: the block sums live in the scratch
        ca, ka = ctx.reduce(sa)
        cb, kb = ctx.reduce_noalloc(sb)
        assert ca == cb and sum(ca) == nd
        assert ka == kb
        ra, rb = ctx.download(sa), ctx.download(sb)
        assert len(ra[0]) == len(rb[0]) == n - nd
        for a, b in zip(ra, rb):
            np.testing.assert_array_equal(a.view(np.uint32), b.view(np.uint32))
    ctx.close()
