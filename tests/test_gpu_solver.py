"""The elliptic solver's coarse-level paths and its MG-PCG stopping logic on
given right-hand sides (gp_solve_rhs), checked with the oracle's assembly of
the reference's Poisson matrix (EMField.mm:236-405).

On a Dirichlet/periodic grid the solver's operator is S = -A (A the
reference's CSR matrix), so its solution x of S x = b satisfies A x = -b; on
an all-periodic grid b is first projected to zero mean."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _ctx(gp, ngx, ngy, fbc, dx=0.01, dy=0.012):
    cfg = gp.make_config(ngx, ngy, dx, dy, order=1, fbc=fbc, pbc=("periodic",) * 4)   # no ghost layer
    return gp.Context(cfg)


def _oracle_relres(orc, ngx, ngy, fbc, b, x, dx=0.01, dy=0.012):
    f = orc.field(ngx, ngy, 0, dx, dy, fbc=fbc)
    ptr, col, val, _ = orc.poisson_csr(f)
    rhs = -b
    if fbc[0] == "periodic" and fbc[2] == "periodic":
        rhs = rhs - rhs.mean()
    return orc.relres(ptr, col, val, rhs, x)


# grids whose hierarchy ends in a level of at most 64 x 64 unknowns (solved
# exactly by fast diagonalisation on MFMA, zero-padded to 64 x 64), and one
# whose coarsening stops at an odd interval count (124 x 50 level: the
# multi-level LDS bottom kernel with the scalar fast diagonalisation)
GRIDS = [
    ("periodic 64x64 (exact solve only)", 64, 64, ("periodic",) * 4),
    ("Dirichlet 64x64 (63x63 unknowns)", 64, 64, ("Dirichlet",) * 4),
    ("Dirichlet-x 40x48 (39x48, padded)", 40, 48, ("Dirichlet", "Dirichlet", "periodic", "periodic")),
    ("periodic 128x128 (one grid level)", 128, 128, ("periodic",) * 4),
    ("Dirichlet 96x80 (47x39 coarsest)", 96, 80, ("Dirichlet",) * 4),
    ("Dirichlet-x 500x200 (124x50 coarsest)", 500, 200, ("Dirichlet", "Dirichlet", "periodic", "periodic")),
]


@pytest.mark.parametrize("name,ngx,ngy,fbc", GRIDS)
def test_solves_meet_the_oracle_matrix(gp, orc, name, ngx, ngy, fbc):
    """Five solves in a row to 1e-10 (the MG-PCG adapts its batch size and
    residual-replacement interval from one solve to the next): each meets the
    tolerance in the oracle's FP64 residual of the reference's matrix"""
    say(f"[solver] {name}")
    with _ctx(gp, ngx, ngy, fbc) as c:
        m = (c.nkx + 1) * (c.nky + 1)
        its = []
        for seed in range(5):
            b = np.random.default_rng(seed).standard_normal(m)
            x, st, rc = c.solve_rhs(b, tol=1e-10, maxiter=200)
            rr = _oracle_relres(orc, ngx, ngy, fbc, b, x)
            its.append(st.iterations)
            assert rc == 0 and st.relres <= 1e-10 and rr <= 1.05e-10, (seed, st.iterations, st.relres, rr)
        say(f"   iterations {its}")
        assert max(its) <= 16


@pytest.mark.parametrize("name,ngx,ngy,fbc", GRIDS[:3])
def test_exact_coarse_solve(gp, orc, name, ngx, ngy, fbc):
    """Where the whole hierarchy is the one MFMA level, the V-cycle is an
    exact solve: one MG-PCG iteration reduces the residual to FP32 rounding
    (measured 3e-6 to 5e-6), and repeated solves to 1e-10 take at most 3
    iterations once the replacement interval has adapted (the first solve may
    take a few more while the FP32 recursion drifts and is restarted)"""
    say(f"[exact coarse] {name}")
    with _ctx(gp, ngx, ngy, fbc) as c:
        m = (c.nkx + 1) * (c.nky + 1)
        b = np.random.default_rng(11).standard_normal(m)
        x, st, rc = c.solve_rhs(b, tol=1e-14, maxiter=1)
        rr = _oracle_relres(orc, ngx, ngy, fbc, b, x)
        say(f"   one iteration: relres {st.relres:.2e} (oracle {rr:.2e})")
        assert st.iterations == 1 and rr < 1e-4
        its = []
        for seed in range(4):
            b = np.random.default_rng(seed).standard_normal(m)
            x, st, rc = c.solve_rhs(b, tol=1e-10, maxiter=50)
            its.append(st.iterations)
            assert rc == 0 and _oracle_relres(orc, ngx, ngy, fbc, b, x) <= 1.05e-10
        say(f"   iterations {its}")
        assert its[0] <= 8 and max(its[1:]) <= 3
