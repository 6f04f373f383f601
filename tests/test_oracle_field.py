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


# ---------------------------------------------------------------------------
# Darwin right-hand sides (src/EMField.mm:1147-1275).  Node (i, j) of the
# Darwin grid is physical node (i + ngb, j + ngb); jx[r, c] is the x face
# between physical nodes c - 1 and c on row r, jy[r, c] the y face between
# rows r - 1 and r on column c.  The divergences and Laplacians below are
# formed from those definitions with array differences, independently of the
# oracle's index arithmetic.
# ---------------------------------------------------------------------------
CL = 2.99792458e10


def _darwin_field(orc, per_x, ngx=14, ngy=10, ngb=1, dx=0.11, dy=0.07):
    fbc = ("periodic",) * 2 if per_x else ("Dirichlet",) * 2
    f = orc.field(ngx, ngy, ngb, dx, dy, fbc=fbc + ("periodic", "periodic"))
    ptr, col, val, (nkx, nky) = orc.darwin_csr(f)
    n = len(ptr) - 1
    A = sp.csr_matrix((val, col, ptr), shape=(n, n)).toarray()
    return f, A, nkx, nky


def _div_j(f, jx, jy, nkx, nky):
    """Node divergence (jx[i+1/2] - jx[i-1/2])/dx + (jy[j+1/2] - jy[j-1/2])/dy."""
    g = f.ngb
    JX = jx[g:g + nky + 1, g:g + nkx + 2]
    JY = jy[g:g + nky + 2, g:g + nkx + 1]
    return np.diff(JX, axis=1) / f.dx + np.diff(JY, axis=0) / f.dy


def _grad_field(f, psi, per_x):
    """Staggered J = grad psi of a node function psi (nky+1, nkx+1) that is
    periodic in y, and periodic in x or independent of x."""
    g, ngx, ngy = f.ngb, f.ngx, f.ngy
    nx, ny = ngx + 2 * g, ngy + 2 * g
    ncol = psi.shape[1]
    jx = np.zeros((ny + 1, nx + 2))
    jy = np.zeros((ny + 2, nx + 1))
    for r in range(ny + 1):
        j = (r - g) % ngy
        for c in range(nx + 2):
            if per_x:
                jx[r, c] = (psi[j, (c - g) % ngx] - psi[j, (c - 1 - g) % ngx]) / f.dx
    for r in range(ny + 2):
        j1, j0 = (r - g) % ngy, (r - 1 - g) % ngy
        for c in range(nx + 1):
            i = (c - g) % ngx if per_x else min(max(c - g, 0), ncol - 1)
            jy[r, c] = (psi[j1, i] - psi[j0, i]) / f.dy
    return jx, jy


@pytest.mark.parametrize("per_x", [True, False])
def test_darwin_operator_eigenfunctions(orc, per_x):
    """cos modes are eigenvectors of the Darwin node operator: periodic,
    lambda = -4/h^2 sin^2(pi k/N); Neumann with the ghost mirror of
    EMField.mm:483-623 (2 Cx on the wall rows), lambda = -4/h^2 sin^2(pi k/(2 nkx))."""
    f, A, nkx, nky = _darwin_field(orc, per_x)
    i = np.arange(nkx + 1)[None, :]
    j = np.arange(nky + 1)[:, None]
    for kx, ky in ((1, 1), (3, 2), (5, 4)):
        if per_x:
            tx = 2 * np.pi * kx / f.ngx
        else:
            tx = np.pi * kx / nkx
        ty = 2 * np.pi * ky / f.ngy
        psi = np.cos(tx * i + 0.3) * np.cos(ty * j + 0.7) if per_x else np.cos(tx * i) * np.cos(ty * j + 0.7)
        lam = -4 / f.dx ** 2 * np.sin(tx / 2) ** 2 - 4 / f.dy ** 2 * np.sin(ty / 2) ** 2
        np.testing.assert_allclose(A @ psi.ravel(), lam * psi.ravel(), atol=1e-11 * abs(lam))


@pytest.mark.parametrize("per_x", [True, False])
def test_darwin_rhs_of_a_gradient_current(orc, per_x):
    """A purely longitudinal current J = grad psi (psi a discrete eigenmode,
    zero mean) gives the chi right-hand side A psi, so chi = psi, and with it
    zero right-hand sides for Ax and Ay: J - grad chi = 0.  Walls in x: psi
    depends on y only (no wall current)."""
    f, A, nkx, nky = _darwin_field(orc, per_x)
    i = np.arange(nkx + 1)[None, :]
    j = np.arange(nky + 1)[:, None]
    ty = 2 * np.pi * 2 / f.ngy
    if per_x:
        psi = np.cos(2 * np.pi * 3 * i / f.ngx + 0.4) * np.cos(ty * j + 0.2) + 0.5 * np.sin(2 * np.pi * i / f.ngx)
    else:
        psi = np.cos(ty * j + 0.2) * np.ones_like(i, dtype=np.float64)
    jx, jy = _grad_field(f, psi, per_x)
    rc = orc.darwin_rhs_chi(f, jx, jy)
    scale = np.abs(A @ psi.ravel()).max()
    np.testing.assert_allclose(rc, A @ psi.ravel(), atol=1e-12 * scale)
    chi, *_ = np.linalg.lstsq(A, rc, rcond=None)
    chi -= chi.mean()
    np.testing.assert_allclose(chi, psi.ravel() - psi.mean(), atol=1e-9 * np.abs(psi).max())
    rx, ry = orc.darwin_rhs_a(f, jx, jy, psi.ravel())
    jmax = max(np.abs(jx).max(), np.abs(jy).max())
    assert np.abs(rx).max() <= 1e-12 * 4 * np.pi / CL * jmax
    assert np.abs(ry).max() <= 1e-12 * 4 * np.pi / CL * jmax


@pytest.mark.parametrize("per_x", [True, False])
def test_darwin_transverse_current(orc, per_x):
    """Random J.  (1) The chi right-hand side is the node divergence of J
    minus one constant (the plain mean shift, EMField.mm:1184-1187) on every
    node off the walls; on a wall node it carries the reference's extra term
    -+2 jx/dx of the ghost face (EMField.mm:1161-1167).  (2) For any chi the
    A right-hand sides, read as a face field, have divergence
    -4 pi/c (div J - lap chi) at the nodes off the boundary faces (their mean
    shifts cancel in differences).  (3) With chi the solution of the
    symmetrised system (rows of wall nodes halved, projected to zero mean),
    the transverse current J - grad chi has one constant divergence off the
    walls, zero on a fully periodic grid.  (4) All three right-hand sides
    have zero mean."""
    f, A, nkx, nky = _darwin_field(orc, per_x)
    g = f.ngb
    nx, ny = f.ngx + 2 * g, f.ngy + 2 * g
    rng = np.random.default_rng(3 + per_x)
    jx = rng.standard_normal((ny + 1, nx + 2))
    jy = rng.standard_normal((ny + 2, nx + 1))
    # one value per physical face on the periodic axes: the wrap face's two
    # array entries (face -1/2 and face N - 1/2) agree
    jy[g, :] = jy[g + f.ngy, :]
    if per_x:
        jx[:, g] = jx[:, g + f.ngx]
    dj = _div_j(f, jx, jy, nkx, nky)
    rc = orc.darwin_rhs_chi(f, jx, jy).reshape(nky + 1, nkx + 1)
    assert abs(rc.sum()) <= 1e-12 * np.abs(rc).sum()
    off = rc - dj
    inner = off[:, 1:nkx] if not per_x else off
    np.testing.assert_allclose(inner, inner[0, 0], atol=1e-12 * np.abs(dj).max())
    if not per_x:
        np.testing.assert_allclose(off[:, 0] - inner[0, 0], -2 * jx[g:g + nky + 1, g] / f.dx, atol=1e-12 * np.abs(dj).max())
        np.testing.assert_allclose(off[:, nkx] - inner[0, 0], 2 * jx[g:g + nky + 1, g + nkx] / f.dx,
                                   atol=1e-12 * np.abs(dj).max())

    def lap(c):
        c = c.reshape(nky + 1, nkx + 1)
        ly = (np.roll(c, -1, 0) - 2 * c + np.roll(c, 1, 0)) / f.dy ** 2
        lx = np.zeros_like(c)
        lx[:, 1:-1] = (c[:, 2:] - 2 * c[:, 1:-1] + c[:, :-2]) / f.dx ** 2
        return lx + ly

    def div_faces(rx, ry):
        rx, ry = rx.reshape(nky + 1, nkx + 1), ry.reshape(nky + 1, nkx + 1)
        d = np.full_like(rx, np.nan)
        d[1:nky, 1:nkx] = (rx[1:nky, 1:nkx] - rx[1:nky, 0:nkx - 1]) / f.dx + (ry[1:nky, 1:nkx] - ry[0:nky - 1, 1:nkx]) / f.dy
        return d

    c0 = rng.standard_normal((nky + 1) * (nkx + 1))
    rx, ry = orc.darwin_rhs_a(f, jx, jy, c0)
    for r in (rx, ry):
        assert abs(r.sum()) <= 1e-12 * np.abs(r).sum()
    exp = -4 * np.pi / CL * (dj - lap(c0))
    got = div_faces(rx, ry)
    np.testing.assert_allclose(got[1:nky, 1:nkx], exp[1:nky, 1:nkx], atol=1e-11 * np.abs(exp).max())

    d = np.ones((nky + 1, nkx + 1))
    if not per_x:
        d[:, 0] = d[:, -1] = 0.5
    d = d.ravel()
    S = -(d[:, None] * A)
    b = -(d * rc.ravel())
    b -= b.mean()
    chi = np.linalg.lstsq(S, b, rcond=None)[0]
    assert np.linalg.norm(S @ chi - b) <= 1e-10 * np.linalg.norm(b)
    rx, ry = orc.darwin_rhs_a(f, jx, jy, chi)
    dt_ = div_faces(rx, ry)[1:nky, 1:nkx] * (-CL / (4 * np.pi))     # div (J - grad chi)
    np.testing.assert_allclose(dt_, dt_[0, 0], atol=1e-10 * np.abs(dj).max())
    if per_x:
        assert abs(dt_[0, 0]) <= 1e-10 * np.abs(dj).max()
