"""Pins of the oracle's field routines (src/EMField.mm:236-405, 483-623,
749-1136, 1138-1329): operator structure, manufactured discrete solutions,
the Laplace problem of testmodules/poisson/main.mm solved by dense LU,
charge conservation of the fold, reconstruction + gradient of a linear
potential, and the energy forms."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla


def csr_matrix(orc, f):
    ptr, col, val, rbc = orc.poisson_csr(f)
    n = len(ptr) - 1
    return sp.csr_matrix((val, col, ptr), shape=(n, n)), ptr, col, val, rbc


@pytest.mark.parametrize("fbc", [("periodic",) * 4, ("Dirichlet", "Dirichlet", "periodic", "periodic"),
                                 ("Dirichlet",) * 4])
def test_poisson_operator_structure(orc, fbc):
    f = orc.field(20, 16, 1, 0.1, 0.1, fbc=fbc)
    A, *_ = csr_matrix(orc, f)
    assert abs(A - A.T).max() < 1e-9 * abs(A).max()      # symmetric for Dirichlet/periodic
    if fbc[0] == "periodic" and fbc[2] == "periodic":
        assert abs(A @ np.ones(A.shape[0])).max() < 1e-9 * abs(A).max()
    nkx, nky, ikmin, jkmin = orc.poisson_dims(f)
    exp_x = {"periodic": 20, "Dirichlet": 19}[fbc[0]]
    assert nkx + 1 == exp_x and ikmin == (2 if fbc[0] == "Dirichlet" else 1)


def test_manufactured_periodic(orc):
    """cos modes are eigenvectors of the periodic 5-point operator."""
    ngx, ngy, dx, dy = 32, 24, 0.05, 0.04
    f = orc.field(ngx, ngy, 1, dx, dy)
    A, ptr, col, val, _ = csr_matrix(orc, f)
    i = np.arange(ngx)[None, :]
    j = np.arange(ngy)[:, None]
    kx, ky = 3, 2
    phi = np.cos(2 * np.pi * kx * i / ngx) * np.cos(2 * np.pi * ky * j / ngy)
    lam = -4 / (np.float64(np.float32(dx)) ** 2) * np.sin(np.pi * kx / ngx) ** 2 \
          - 4 / (np.float64(np.float32(dy)) ** 2) * np.sin(np.pi * ky / ngy) ** 2
    b = (lam * phi).ravel()
    np.testing.assert_allclose(A @ phi.ravel(), b, atol=1e-9 * abs(b).max())
    x, it, rr = orc.cg(ptr, col, val, b, tol=1e-13, project=True)
    assert rr < 1e-12
    np.testing.assert_allclose(x, phi.ravel(), atol=1e-9)


def test_manufactured_dirichlet_periodic(orc):
    ngx, ngy, dx = 32, 16, 0.1
    f = orc.field(ngx, ngy, 1, dx, dx, fbc=("Dirichlet", "Dirichlet", "periodic", "periodic"))
    A, ptr, col, val, rbc = csr_matrix(orc, f)
    nkx, nky, _, _ = orc.poisson_dims(f)
    i = np.arange(1, nkx + 2)[None, :]          # interior nodes 1..ngx-1 (walls at 0, ngx)
    j = np.arange(nky + 1)[:, None]
    phi = np.sin(np.pi * 2 * i / ngx) * np.cos(2 * np.pi * j / ngy)
    h = np.float64(np.float32(dx))
    lam = -4 / h ** 2 * (np.sin(np.pi * 2 / (2 * ngx)) ** 2 + np.sin(np.pi * 1 / ngy) ** 2)
    b = (lam * phi).ravel()
    np.testing.assert_allclose(A @ phi.ravel(), b, atol=1e-9 * abs(b).max())
    x, it, rr = orc.cg(ptr, col, val, b, tol=1e-13)
    np.testing.assert_allclose(x, phi.ravel(), atol=1e-9)


def test_laplace_testmodule(orc):
    """testmodules/poisson/main.mm: 20x20 cells, phi = 200 on x-min, 0 on the
    other walls, rho = 0; CG against a dense LU of the same operator."""
    f = orc.field(20, 20, 2, 1e-2, 1e-2, fbc=("Dirichlet",) * 4, fval=(200.0, 0.0, 0.0, 0.0))
    A, ptr, col, val, rbc = csr_matrix(orc, f)
    nx = 20 + 4
    rho = np.zeros((nx + 1, nx + 1))
    rhs = orc.poisson_rhs(f, rho, rbc)
    xd = np.linalg.solve(A.toarray(), rhs)
    x, it, rr = orc.cg(ptr, col, val, rhs, tol=1e-13)
    np.testing.assert_allclose(x, xd, atol=1e-9 * 200)
    assert x.min() > 0 and x.max() < 200                   # maximum principle
    X = x.reshape(19, 19)
    np.testing.assert_allclose(X, X[::-1, :], atol=1e-9)    # symmetric in y
    # reconstruction: wall nodes carry the Dirichlet values
    phi = orc.reconstruct(f, x)
    ngb = 2
    np.testing.assert_allclose(phi[ngb + 1 + 5, ngb + 1], 200.0)       # node (ngb, .) at phi index ngb+1
    np.testing.assert_allclose(phi[ngb + 1 + 5, ngb + 1 + 20], 0.0, atol=1e-12)


def test_fold_conserves_charge(orc):
    rng = np.random.default_rng(0)
    for fbc in (("periodic",) * 4, ("Dirichlet", "Dirichlet", "periodic", "periodic")):
        ngb, ngx, ngy = 2, 16, 12
        nx, ny = ngx + 2 * ngb, ngy + 2 * ngb
        f = orc.field(ngx, ngy, ngb, 1.0, 1.0, fbc=fbc)
        rho = rng.random((ny + 1, nx + 1))
        r = orc.fold(f, rho)
        if fbc[0] == "periodic":
            # every node's charge lands on one of the ngx x ngy unique nodes
            uniq = r[ngb:ngb + ngy, ngb:ngb + ngx]
            np.testing.assert_allclose(uniq.sum(), rho.sum(), rtol=1e-12)
        else:
            # Dirichlet x: interior unknown nodes ngb+1..ngb+ngx-1 receive all
            uniq = r[ngb:ngb + ngy, ngb + 1:ngb + ngx]
            np.testing.assert_allclose(uniq.sum(), rho.sum(), rtol=1e-12)


def test_linear_potential_gives_uniform_field(orc):
    """Dirichlet V at x-min, 0 at x-max, no charge: phi linear, Ex = V/L."""
    ngx, ngy, dx, V = 32, 16, 0.01, 3.0
    f = orc.field(ngx, ngy, 1, dx, dx, fbc=("Dirichlet", "Dirichlet", "periodic", "periodic"), fval=(V, 0, 0, 0))
    A, ptr, col, val, rbc = csr_matrix(orc, f)
    nx, ny = ngx + 2, ngy + 2
    rhs = orc.poisson_rhs(f, np.zeros((ny + 1, nx + 1)), rbc)
    x, it, rr = orc.cg(ptr, col, val, rhs, tol=1e-14)
    phi = orc.reconstruct(f, x)
    Ex, Ey = orc.gradient(f, phi)
    L = ngx * np.float64(np.float32(dx))
    np.testing.assert_allclose(Ex, V / L, rtol=1e-6)
    np.testing.assert_allclose(Ey, 0.0, atol=1e-6 * V / L)
    we_ref, we_cgs, _ = orc.field_energy(f, Ex, Ey)
    np.testing.assert_allclose(we_cgs, (V / L) ** 2 * L * ngy * np.float32(dx) / (8 * np.pi), rtol=1e-5)


def test_darwin_operator(orc):
    """Neumann x / periodic y node operator: rows sum to zero, and D*A (rows
    of Neumann boundary nodes halved) is symmetric."""
    f = orc.field(16, 12, 1, 0.1, 0.1, fbc=("Dirichlet", "Dirichlet", "periodic", "periodic"))
    ptr, col, val, (nkx, nky) = orc.darwin_csr(f)
    n = len(ptr) - 1
    A = sp.csr_matrix((val, col, ptr), shape=(n, n)).toarray()
    assert np.abs(A.sum(axis=1)).max() < 1e-9 * np.abs(A).max()
    d = np.ones((nky + 1, nkx + 1))
    d[:, 0] = d[:, -1] = 0.5
    S = d.ravel()[:, None] * A
    np.testing.assert_allclose(S, S.T, atol=1e-9 * np.abs(A).max())


@pytest.mark.parametrize("ngb", [0, 1, 2])
@pytest.mark.parametrize("perx,pery", [(True, True), (True, False), (False, True)])
def test_integer_fold_matches_reference_fold(orc, ngb, perx, pery):
    """fold_acc (the periodic ghost fold on int64 accumulators) gives the
    interior nodes of the reference-order FP64 fold (gpo_fold_rho) on
    integer charge (exact in FP64), clears the periodic ghosts, and conserves
    the total; a deposit moved to a ghost node's interior image folds to the
    same array."""
    ngx, ngy = 12, 9
    nx, ny = ngx + 2 * ngb, ngy + 2 * ngb
    rng = np.random.default_rng(10 * ngb + 2 * perx + pery)
    a = rng.integers(0, 2 ** 40, (ny + 1, nx + 1), dtype=np.int64)
    fbc = ("periodic" if perx else "Neumann",) * 2 + ("periodic" if pery else "Neumann",) * 2
    f = orc.field(ngx, ngy, ngb, 0.1, 0.1, fbc=fbc)
    ref = orc.fold(f, a.astype(np.float64))
    got = orc.fold_acc(a, ngb, ngx, ngy, perx, pery)
    ix = slice(ngb, ngb + ngx) if perx else slice(0, nx + 1)
    iy = slice(ngb, ngb + ngy) if pery else slice(0, ny + 1)
    if perx and pery:
        np.testing.assert_array_equal(got[iy, ix].astype(np.float64), ref[iy, ix])
        assert got.sum() == a.sum()
    if perx:
        assert not got[:, :ngb].any() and not got[:, ngb + ngx:].any()
    if pery:
        assert not got[:ngb, :].any() and not got[ngb + ngy:, :].any()
    # moving charge from a ghost node to its interior image leaves the fold unchanged
    if perx:
        b = a.copy()
        g = ngb + ngx            # first high ghost column -> interior column ngb
        b[:, g - ngx] += b[:, g]
        b[:, g] = 0
        np.testing.assert_array_equal(orc.fold_acc(b, ngb, ngx, ngy, perx, pery), got)
