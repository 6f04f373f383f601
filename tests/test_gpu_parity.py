"""GPU parity: libgample_amd (through its C-ABI) against the CPU oracle on
identical inputs.

Bar (north_star): post-push positions/velocities and cell indices bit-exact
against the oracle's kernel-order push; int64 deposits bit-exact; Poisson to
a 1e-8 FP64 relative residual; loader positions bit-exact, velocities
(libm vs device transcendentals) within 1e-6 relative.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ME, EC, CL = 9.10938188e-28, 4.803204197e-10, 2.99792458e10


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def mkctx(gp, ngx, ngy, order, pbc=("periodic",) * 4, fbc=("periodic",) * 4, boris=2, cur=0, b0=(0, 0, 0),
          dx=0.01, em=1):
    cfg = gp.make_config(ngx, ngy, dx, dx, order=order, boris=boris, pbc=pbc, fbc=fbc, b0=b0, cur_dens=cur,
                         em_model=em)
    return gp.Context(cfg)


def rand_parts(n, ngx, ngy, ngb, vs=1e8, seed=0):
    rng = np.random.default_rng(seed)
    x = (ngb + rng.random(n) * ngx).astype(np.float32)
    y = (ngb + rng.random(n) * ngy).astype(np.float32)
    v = (vs * rng.standard_normal((3, n))).astype(np.float32)
    return [x, y, v[0], v[1], v[2]]


def rand_fields(ctx, gp, seed=0, scale=1.0, bz=None):
    rng = np.random.default_rng(seed)
    Ex = (scale * rng.standard_normal(ctx.field_shape(gp.F_EX)[0])).astype(np.float32)
    Ey = (scale * rng.standard_normal(ctx.field_shape(gp.F_EY)[0])).astype(np.float32)
    ctx.set_field(gp.F_EX, Ex)
    ctx.set_field(gp.F_EY, Ey)
    Bz = None
    if bz is not None:
        Bz = (bz * (1 + 0.1 * rng.standard_normal(ctx.field_shape(gp.F_BZ)[0]))).astype(np.float32)
        ctx.set_field(gp.F_BZ, Bz)
    return Ex, Ey, Bz


def consts(q, m, dt, dx):
    cE = np.float32(0.5 * float(np.float32(q)) * dt / float(np.float32(m)))
    cB = np.float32(0.5 * float(np.float32(q)) * dt / float(np.float32(m)) / CL)
    cX = np.float32(dt / float(np.float32(dx)))
    return float(cE), float(cB), float(cX), float(cX)


CASES = [(1, 2, False), (3, 2, True), (3, 4, False), (5, 6, False), (5, 2, True), (1, 6, True)]


@pytest.mark.parametrize("order,boris,uniform", CASES)
@pytest.mark.parametrize("sorted_", [False, True])
def test_push_deposit_bit_exact(gp, orc, order, boris, uniform, sorted_):
    say(f"[push] order={order} boris={boris} uniform={uniform} sorted={sorted_}")
    ngx, ngy, dx = 64, 48, 0.01
    ngb = order // 2
    ctx = mkctx(gp, ngx, ngy, order, boris=boris, b0=(0, 0, 300.0), dx=dx)
    sid = ctx.add_species(-EC, ME, 1.0, 200000, vth_hint=1e8)
    ctx.upload(sid, *rand_parts(150000, ngx, ngy, ngb, seed=order))
    Ex, Ey, Bz = rand_fields(ctx, gp, seed=order, scale=2.0, bz=None if uniform else 300.0)
    if sorted_:
        ctx.sort(sid)
    before = ctx.download(sid)
    dt = 2e-12
    ctx.reset_grids()
    ctx.push(sid, dt, deposit=1)
    after = ctx.download(sid)
    g = orc.grid(ngx, ngy, order, pbc=("periodic",) * 4, boris=boris)
    Bz_h = ctx.field(gp.F_BZ)
    exp = orc.push_fast(g, Ex, Ey, Bz_h, consts(-EC, ME, dt, dx), before, uniform_bz=uniform, bz0=300.0)
    for k, (a, b) in enumerate(zip(after, exp)):
        bad = np.nonzero(a.view(np.uint32) != b.view(np.uint32))[0]
        assert len(bad) == 0, f"component {k}: {len(bad)} mismatches, first {bad[:5]} {a[bad[:5]]} {b[bad[:5]]}"
    # binned particles deposit by position on the periodic axes (charge at a
    # ghost node lands on its interior image): compare after the ghost fold
    rho = orc.fold_acc(ctx.acc(sid), ngb, ngx, ngy)
    erho, _, _ = orc.deposit_fx(g, exp)
    np.testing.assert_array_equal(rho, orc.fold_acc(erho, ngb, ngx, ngy))
    ctx.close()


@pytest.mark.parametrize("order", [1, 3, 5])
def test_current_deposit_bit_exact(gp, orc, order):
    say(f"[current] order={order}")
    ngx, ngy = 40, 32
    ngb = order // 2
    ctx = mkctx(gp, ngx, ngy, order, cur=1)
    sid = ctx.add_species(-EC, ME, 1.0, 100000, vth_hint=1e8)
    p = rand_parts(80000, ngx, ngy, ngb, seed=11)
    ctx.upload(sid, *p)
    ctx.sort(sid)
    p = ctx.download(sid)
    ctx.reset_grids()
    ctx.deposit(sid)
    g = orc.grid(ngx, ngy, order)
    inv = 1.0 / 2.0 ** np.ceil(np.log2(8e8))
    rho, jx, jy = orc.deposit_fx(g, p, inv_vnorm=inv, current=True)
    np.testing.assert_array_equal(ctx.acc(sid, gp.G_RHO_ACC), rho)
    np.testing.assert_array_equal(ctx.acc(sid, gp.G_JX_ACC), jx)
    np.testing.assert_array_equal(ctx.acc(sid, gp.G_JY_ACC), jy)
    ctx.close()


GP_TILE, GP_DRIFT = 16, 12   # tile edge and drift margin of the push kernel's LDS windows (cells)


def sort_order(x, y, ngx, ngy, ngb, order):
    """Permutation gp_sort applies: grouped by tile (row-major), and inside a
    tile rank-major over the 64 LDS bank classes b = (cx + W cy) mod 64 of
    the cells (W = the window pitch), each class in its previous order."""
    T, W = GP_TILE, GP_TILE + 2 * GP_DRIFT + order + 1
    ci = np.clip(np.floor(x).astype(np.int64) - ngb, 0, ngx - 1)
    cj = np.clip(np.floor(y).astype(np.int64) - ngb, 0, ngy - 1)
    ntx = (ngx + T - 1) // T
    t = (cj // T) * ntx + ci // T
    b = ((ci % T) + W * (cj % T)) & 63
    o1 = np.argsort(t * 64 + b, kind="stable")
    ks = (t * 64 + b)[o1]
    first = np.r_[0, np.flatnonzero(np.diff(ks)) + 1]
    start = np.repeat(first, np.diff(np.r_[first, ks.size]))
    r = np.arange(ks.size) - start
    return o1[np.lexsort((b[o1], r, t[o1]))]


@pytest.mark.parametrize("order", [1, 3, 5])
def test_sort_order(gp, orc, order):
    """The sort is a fixed function of the input: tiles in order, classes
    interleaved inside a tile, stable within a class."""
    say(f"[sort] order={order}")
    ngx, ngy = 64, 40
    ngb = order // 2
    ctx = mkctx(gp, ngx, ngy, order)
    sid = ctx.add_species(-EC, ME, 1.0, 300000)
    p = rand_parts(250000, ngx, ngy, ngb, seed=3)
    p[1][:5000] = np.float32(ngb + 3.5)                 # a crowded row: uneven classes
    p[4] = np.arange(len(p[0]), dtype=np.float32)      # tag with the original index
    ctx.upload(sid, *p)
    ctx.sort(sid)
    s = ctx.download(sid)
    tag = s[4].astype(np.int64)
    np.testing.assert_array_equal(tag, sort_order(p[0], p[1], ngx, ngy, ngb, order))
    for a, b in zip(s[:4], p[:4]):
        np.testing.assert_array_equal(a, b[tag])
    # a second sort of the sorted store is the same function of its input
    ctx.sort(sid)
    s2 = ctx.download(sid)
    np.testing.assert_array_equal(s2[4].astype(np.int64), tag[sort_order(s[0], s[1], ngx, ngy, ngb, order)])
    ctx.close()


@pytest.mark.parametrize("cur", [0, 1])
def test_sorted_and_unsorted_deposits_agree(gp, orc, cur):
    """int64 accumulation: the LDS-tiled and the global path give the same bits
    (the charge after the periodic ghost fold; the current as deposited)."""
    say(f"[tiling] cur={cur}")
    ctx = mkctx(gp, 64, 64, 3, cur=cur)
    sid = ctx.add_species(-EC, ME, 1.0, 300000, vth_hint=1e8)
    ctx.upload(sid, *rand_parts(300000, 64, 64, 1, seed=5))
    which = (gp.G_RHO_ACC, gp.G_JX_ACC, gp.G_JY_ACC) if cur else (gp.G_RHO_ACC,)
    ctx.reset_grids(); ctx.deposit(sid)
    a = [ctx.acc(sid, w) for w in which]
    ctx.sort(sid)
    ctx.reset_grids(); ctx.deposit(sid)
    b = [ctx.acc(sid, w) for w in which]
    a[0], b[0] = orc.fold_acc(a[0], 1, 64, 64), orc.fold_acc(b[0], 1, 64, 64)
    for u, v in zip(a, b):
        np.testing.assert_array_equal(u, v)
    ctx.close()


def test_absorption_compaction(gp, orc):
    say("[reduce]")
    ngx, ngy, dx = 32, 32, 0.01
    ctx = mkctx(gp, ngx, ngy, 3, pbc=("Delete", "Delete", "periodic", "periodic"),
                fbc=("Dirichlet", "Dirichlet", "periodic", "periodic"), dx=dx)
    sid = ctx.add_species(-EC, ME, 1.0, 100000)
    p = rand_parts(60000, ngx, ngy, 1, vs=3e9, seed=8)
    ctx.upload(sid, *p)
    ctx.sort(sid)
    before = ctx.download(sid)
    dt = 1e-12
    ctx.reset_grids()
    ctx.push(sid, dt, deposit=1)
    mid = ctx.download(sid)
    g = orc.grid(ngx, ngy, 3, pbc=("Delete", "Delete", "periodic", "periodic"))
    z = lambda s: np.zeros(ctx.field_shape(s)[0], np.float32)
    exp = orc.push_fast(g, z(gp.F_EX), z(gp.F_EY), z(gp.F_BZ), consts(-EC, ME, dt, dx), before, uniform_bz=True)
    for a, b in zip(mid, exp):
        np.testing.assert_array_equal(a, b)
    flags = np.where(mid[0] < 0, -mid[0], 0).astype(np.int32)
    assert flags.sum() > 0
    cnt, ke = ctx.reduce(sid)
    after = ctx.download(sid)
    from test_oracle_reduce import hole_donor
    model, mcnt, n_new = hole_donor(mid, flags)
    assert cnt == mcnt and len(after[0]) == n_new
    for a, b in zip(after, model):
        np.testing.assert_array_equal(a, b)
    ke_exp = orc.kinetic(exp, np.float32(ME), np.float32(1.0))
    np.testing.assert_allclose(ke, ke_exp, rtol=1e-12)
    ctx.close()


@pytest.mark.parametrize("gt", [0, 1, 4, 7])
def test_loader(gp, orc, gt):
    say(f"[load] gen_type={gt}")
    ngx, ngy, dx = 64, 64, 0.01
    ctx = mkctx(gp, ngx, ngy, 3, dx=dx)
    se = ctx.add_species(-EC, ME, 1.0, 70000)
    si = ctx.add_species(EC, 2.18e-22, 1.0, 70000)
    G = gp.Gen()
    G.gen_type = gt
    G.xmin, G.xmax, G.ymin, G.ymax = 0.0, ngx * dx, 0.0, ngy * dx
    G.ele_vth, G.ion_vth = 1.9e8, 9e4
    G.ele_u[:] = [1e7, 0, 0]
    G.pert_delta, G.pert_mode = 0.01, 4
    ctx.load_pair(se, si, G, 0, 0, 65536)
    e, i = ctx.download(se), ctx.download(si)
    O = orc.Gen()
    O.gen_type = gt
    O.xmin, O.xmax, O.ymin, O.ymax = G.xmin, G.xmax, G.ymin, G.ymax
    O.ele_vth, O.ion_vth = G.ele_vth, G.ion_vth
    O.ele_u[:] = [1e7, 0, 0]
    O.dx = O.dy = dx
    O.ngb = 1.0
    O.pert_delta, O.pert_mode = 0.01, 4
    O.xdom, O.ydom = ngx, ngy
    oe, oi = orc.load_pair(O, 0, 0, 65536)
    if gt in (0, 1, 4, 7):
        assert e[0].min() >= 1.0 and e[0].max() < 1.0 + ngx and e[1].min() >= 1.0 and e[1].max() < 1.0 + ngy
    if gt in (0, 4):
        np.testing.assert_array_equal(e[0], oe[0])
        np.testing.assert_array_equal(e[1], oe[1])
    else:
        # asinf / sin / cos: device vs glibc differ by a few ulp (cells)
        np.testing.assert_allclose(e[0], oe[0], rtol=0, atol=2e-5)
        np.testing.assert_array_equal(e[1], oe[1])
    for a, b in zip(e[2:] + i[2:], oe[2:] + oi[2:]):
        np.testing.assert_allclose(a, b, rtol=2e-5, atol=1e-6 * np.abs(b).max())
    ctx.close()


def _poisson_case(gp, orc, order, fbc, ngx, ngy, seed=0, tol=1e-8):
    ngb = order // 2
    dx = 0.005
    ctx = mkctx(gp, ngx, ngy, order, fbc=fbc, pbc=("periodic",) * 4, dx=dx)
    se = ctx.add_species(-EC, ME, 1000.0, 400000)
    si = ctx.add_species(EC, 2.18e-22, 1000.0, 400000)
    rng = np.random.default_rng(seed)
    pe = rand_parts(300000, ngx, ngy, ngb, seed=seed)
    pi = rand_parts(300000, ngx, ngy, ngb, seed=seed + 100)
    ctx.upload(se, *pe)
    ctx.upload(si, *pi)
    ctx.reset_grids()
    ctx.deposit(se)
    ctx.deposit(si)
    acc_e, acc_i = ctx.acc(se), ctx.acc(si)
    st, rc = ctx.solve_fields(0.0, tol=tol, maxiter=400)
    return ctx, st, acc_e, acc_i, dx


@pytest.mark.parametrize("order,fbc,ngx,ngy", [
    (1, ("periodic",) * 4, 64, 64),
    (3, ("Dirichlet", "Dirichlet", "periodic", "periodic"), 64, 64),
    (5, ("periodic",) * 4, 128, 64),
    (3, ("Dirichlet",) * 4, 64, 32),
    (3, ("Dirichlet", "Dirichlet", "periodic", "periodic"), 512, 512),
    (1, ("periodic",) * 4, 256, 256),
    (3, ("Dirichlet",) * 4, 96, 80),        # intervals not powers of two
    (1, ("periodic",) * 4, 16, 512),        # grid level narrower than a tile window
])
def test_poisson_against_oracle(gp, orc, order, fbc, ngx, ngy):
    say(f"[poisson] order={order} fbc={fbc} {ngx}x{ngy}")
    ctx, st, acc_e, acc_i, dx = _poisson_case(gp, orc, order, fbc, ngx, ngy)
    ngb = order // 2
    f = orc.field(ngx, ngy, ngb, dx, dx, fbc=fbc)
    # Particle.mm:540: q*w in FP32, divided by dx*dy in FP64, stored FP32
    sc_e = float(np.float32(float(np.float32(-EC) * np.float32(1000.0)) / (dx * dx)))
    sc_i = float(np.float32(float(np.float32(EC) * np.float32(1000.0)) / (dx * dx)))
    rho = acc_e.astype(np.float64) * (2.0 ** -46 * sc_e) + acc_i.astype(np.float64) * (2.0 ** -46 * sc_i)
    rho = orc.fold(f, rho)
    np.testing.assert_allclose(ctx.field(gp.F_RHO), rho, rtol=1e-13, atol=1e-13 * np.abs(rho).max())
    ptr, col, val, rbc = orc.poisson_csr(f)
    rhs = orc.poisson_rhs(f, rho, rbc)
    singular = fbc[0] == "periodic" and fbc[2] == "periodic"
    if singular:
        rhs = rhs - rhs.mean()
    sol = ctx.field(gp.F_SOL).ravel()
    rr = orc.relres(ptr, col, val, rhs, sol)
    say(f"   iters={st.iterations} refine={st.refinements} relres(gpu)={st.relres:.3e} relres(oracle)={rr:.3e}")
    assert st.relres <= 1e-8 and rr <= 1.01e-8
    x, it, r2 = orc.cg(ptr, col, val, rhs, tol=1e-13, project=singular)
    assert np.linalg.norm(sol - x) / np.linalg.norm(x) < 1e-5
    phi = orc.reconstruct(f, sol)
    Ex, Ey = orc.gradient(f, phi)
    np.testing.assert_allclose(ctx.field(gp.F_PHI), phi, rtol=1e-12, atol=1e-12 * np.abs(phi).max())
    np.testing.assert_array_equal(ctx.field(gp.F_EX), Ex)
    np.testing.assert_array_equal(ctx.field(gp.F_EY), Ey)
    a, b, _ = ctx.energy(0.0)
    ea, eb, _ = orc.field_energy(f, Ex, Ey)
    np.testing.assert_allclose([a, b], [ea, eb], rtol=1e-12)
    ctx.close()


def test_moments(gp, orc):
    say("[moments]")
    ngx, ngy = 32, 32
    ctx = mkctx(gp, ngx, ngy, 3)
    sid = ctx.add_species(-EC, ME, 50.0, 100000, vth_hint=1e8)
    p = rand_parts(100000, ngx, ngy, 1, seed=21)
    ctx.upload(sid, *p)
    m = ctx.moments(sid)
    g = orc.grid(ngx, ngy, 3)
    e = orc.moments_f64(g, p, float(np.float32(50.0)) / float(np.float32(0.01)) ** 2)
    for k in range(10):
        nrm = np.linalg.norm(e[k])
        assert np.linalg.norm(m[k] - e[k]) / nrm < 1e-6, k
    ctx.close()


def test_step_loop_matches_oracle(gp, orc):
    """10 cycles of the fused step (64^2 periodic, CIC, two species) against
    the oracle loop: push_fast + deposit_fx + FP64 CG + reconstruction."""
    say("[loop]")
    ngx = ngy = 64
    dx, dt = 0.005, 5e-12
    ctx = mkctx(gp, ngx, ngy, 1, dx=dx)
    se = ctx.add_species(-EC, ME, 2000.0, 70000)
    si = ctx.add_species(EC, 2.18e-22, 2000.0, 70000)
    pe = rand_parts(40000, ngx, ngy, 0, vs=1e8, seed=31)
    pi = rand_parts(40000, ngx, ngy, 0, vs=1e5, seed=32)
    # vz is never changed by the push (Ez = Bx = By = 0): unique tags
    pe[4] = np.arange(40000, dtype=np.float32)
    pi[4] = np.arange(40000, dtype=np.float32)
    ctx.upload(se, *pe)
    ctx.upload(si, *pi)
    ctx.sort(se); ctx.sort(si)
    pe, pi = ctx.download(se), ctx.download(si)
    ctx.reset_grids(); ctx.deposit(se); ctx.deposit(si)
    for cyc in range(1, 11):
        ctx.step(dt, tol=1e-10, maxiter=500, sort_interval=5, cycle=cyc)
    ge, gi = ctx.download(se), ctx.download(si)
    # oracle loop (order preserved: compare after sorting by original content)
    g = orc.grid(ngx, ngy, 1)
    f = orc.field(ngx, ngy, 0, dx, dx)
    ptr, col, val, rbc = orc.poisson_csr(f)
    sol = None
    for cyc in range(10):
        rho = np.zeros((ngy + 1, ngx + 1))
        for parts, q in ((pe, -EC), (pi, EC)):
            acc, _, _ = orc.deposit_fx(g, parts)
            sc = float(np.float32(np.float32(q) * np.float32(2000.0)) / (dx * dx))
            rho += acc * (2.0 ** -46 * float(np.float32(sc)))
        rho = orc.fold(f, rho)
        rhs = orc.poisson_rhs(f, rho, rbc)
        rhs -= rhs.mean()
        sol, _, _ = orc.cg(ptr, col, val, rhs, x0=sol, tol=1e-12, project=True)
        phi = orc.reconstruct(f, sol)
        Ex, Ey = orc.gradient(f, phi)
        Bz = np.zeros((ngy + 2, ngx + 2), np.float32)
        pe = orc.push_fast(g, Ex, Ey, Bz, consts(-EC, ME, dt, dx), pe, uniform_bz=True)
        pi = orc.push_fast(g, Ex, Ey, Bz, consts(EC, 2.18e-22, dt, dx), pi, uniform_bz=True)
    for gpu, cpu in ((ge, pe), (gi, pi)):
        # the device store is re-sorted every 5 cycles: match particles by
        # vz, which the push leaves unchanged (Ez = Bx = By = 0)
        o = np.argsort(gpu[4]); p = np.argsort(cpu[4])
        for a, b in zip(gpu, cpu):
            np.testing.assert_allclose(a[o], b[p], rtol=1e-5, atol=1e-5 * np.abs(b).max())
    ctx.close()


def test_deterministic_rerun(gp):
    say("[determinism]")
    outs = []
    for _ in range(2):
        ctx = mkctx(gp, 64, 64, 3, cur=1, b0=(0, 0, 100.0))
        se = ctx.add_species(-EC, ME, 500.0, 300000, vth_hint=1e8)
        G = gp.Gen()
        G.gen_type = 0
        G.xmin, G.xmax, G.ymin, G.ymax = 0.0, 0.64, 0.0, 0.64
        G.ele_vth = 1.9e8
        ctx.load_pair(se, -1, G, 5, 0, 262144)
        ctx.sort(se)
        ctx.reset_grids(); ctx.deposit(se)
        for cyc in range(1, 6):
            ctx.step(5e-12, tol=1e-8, sort_interval=3, cycle=cyc)
        outs.append([ctx.download(se), ctx.acc(se, gp.G_RHO_ACC), ctx.acc(se, gp.G_JX_ACC), ctx.field(gp.F_EX)])
        ctx.close()
    for a, b in zip(outs[0][0], outs[1][0]):
        np.testing.assert_array_equal(a, b)
    for a, b in zip(outs[0][1:], outs[1][1:]):
        np.testing.assert_array_equal(a, b)


def test_drift_triggered_sort(gp, orc):
    """gp_set_sort_drift: a species whose particles leave their tile windows
    is re-sorted by gp_step before its next push, and the run is otherwise
    the same to the bit (particles matched by tag, rho, Ex)."""
    say("[drift sort]")
    ngx, dx, dt = 64, 0.005, 5e-12
    cX = float(np.float32(dt / float(np.float32(dx))))
    runs = []
    for frac in (0.0, 1e-3):
        ctx = mkctx(gp, ngx, ngx, 3, dx=dx)
        sid = ctx.add_species(-EC, ME, 1.0, 300000, vth_hint=1e9)
        rng = np.random.default_rng(11)
        p = rand_parts(200000, ngx, ngx, 1, seed=12)
        p[2] = (rng.uniform(-4.0, 4.0, p[2].size) / cX).astype(np.float32)   # up to 4 cells per step
        p[3] = (rng.uniform(-4.0, 4.0, p[3].size) / cX).astype(np.float32)
        p[4] = np.arange(p[0].size, dtype=np.float32)
        ctx.upload(sid, *p)
        ctx.sort(sid)
        ctx.set_sort_drift(frac)
        ctx.reset_grids(); ctx.deposit(sid)
        for cyc in range(1, 9):
            ctx.step(dt, tol=1e-10, maxiter=400, sort_interval=0, cycle=cyc)
        s = ctx.download(sid)
        o = np.argsort(s[4])
        runs.append(([a[o] for a in s], ctx.acc(sid, gp.G_RHO_ACC), ctx.field(gp.F_EX), ctx.sort_stats(sid)))
        ctx.close()
    assert runs[0][3][0] == 1                       # the initial sort only
    assert runs[1][3][0] > 1 and runs[1][3][1] >= 0
    for a, b in zip(runs[0][0], runs[1][0]):
        np.testing.assert_array_equal(a, b)
    np.testing.assert_array_equal(orc.fold_acc(runs[0][1], 1, ngx, ngx), orc.fold_acc(runs[1][1], 1, ngx, ngx))
    np.testing.assert_array_equal(runs[0][2], runs[1][2])


def test_neighbour_rebin_sort(gp, orc):
    """A mostly binned species of which at least 1 % changed cell in its last
    push is re-sorted by its moves to neighbouring tiles: the permutation is
    oracle/sortspec.py's rebin_order (near arrivals grouped by source offset,
    then the far ones, the old order within each) on a grid of 4 x 3 tiles,
    with moves across the periodic edges, jumps of up to 30 cells and an
    unbinned tail appended after the last sort.  Re-sorting the result moves
    nothing."""
    from oracle import sortspec as ss
    say("[rebin sort]")
    ngx, ngy, dx, dt = 64, 48, 0.005, 5e-12
    cX = float(np.float32(dt / float(np.float32(dx))))
    ctx = mkctx(gp, ngx, ngy, 3, dx=dx)
    sid = ctx.add_species(-EC, ME, 1.0, 300000, vth_hint=1e9)
    n0, ntail = 200000, 20000
    rng = np.random.default_rng(31)
    p = rand_parts(n0, ngx, ngy, 1, seed=32)
    v = rng.uniform(-4.0, 4.0, (2, n0))
    fast = rng.random(n0) < 0.02
    v[:, fast] = rng.uniform(-15.0, 15.0, (2, int(fast.sum())))
    p[2] = (v[0] / cX).astype(np.float32)
    p[3] = (v[1] / cX).astype(np.float32)
    ctx.upload(sid, *p)
    ctx.sort(sid)
    s0 = ctx.download(sid)
    off0 = ss.tile_offsets(s0[0], s0[1], ngx, ngy, 1)
    G = gp.Gen()
    G.gen_type = 0
    G.xmin, G.xmax = 0.0, ngx * float(np.float32(dx))
    G.ymin, G.ymax = 0.0, ngy * float(np.float32(dx))
    G.ele_vth, G.ion_vth = 1e8, 1e7
    ctx.load_pair(sid, -1, G, 7, n0, ntail)
    assert ctx.count(sid) == n0 + ntail
    ctx.set_sort_drift(0.0)
    for cyc in (1, 2):          # the second step reads back the first one's cell changes
        ctx.step(dt, tol=1e-10, maxiter=400, sort_interval=0, cycle=cyc)
    pre = ctx.download(sid)
    t, k = ss.rebin_groups(pre[0], pre[1], off0, n0, ngx, ngy, 1)
    assert (k[:n0] == 9).sum() > 100 and (k[:n0] != 4).mean() > 0.1
    order = ss.rebin_order(pre[0], pre[1], off0, n0, ngx, ngy, 1)
    ctx.sort(sid)
    post = ctx.download(sid)
    for a, b in zip(post, pre):
        np.testing.assert_array_equal(a.view(np.uint32), b[order].view(np.uint32))
    ctx.sort(sid)
    again = ctx.download(sid)
    for a, b in zip(again, post):
        np.testing.assert_array_equal(a.view(np.uint32), b.view(np.uint32))
    ctx.close()


def test_hot_push_kinetic_energy(gp, orc):
    """The particles-periodic order-3 push with uniform B and charge does not
    read vz (no push changes it): its kinetic energy adds the sum of vz^2
    kept per species.  Against the oracle's per-particle FP32 |v|^2 the two
    differ only in the last FP32 rounding of each particle's sum, so the
    totals agree to 2^-24 relative; the positions and velocities stay
    bit-exact and vz untouched.  The kept sum must follow every change of the
    particles: an appended load, then an upload of new vz at the same count."""
    say("[hot push kinetic energy]")
    ngx, ngy, dx, dt = 64, 64, 0.01, 1e-12
    ctx = mkctx(gp, ngx, ngy, 3, b0=(0.0, 0.0, 500.0), dx=dx)
    sid = ctx.add_species(-EC, ME, 1.0, 300000)
    Ex, Ey, _ = rand_fields(ctx, gp, seed=4, scale=2.0)
    g = orc.grid(ngx, ngy, 3)
    bz = np.zeros(ctx.field_shape(gp.F_BZ)[0], np.float32)
    cst = consts(-EC, ME, dt, dx)
    tol = 2.0 ** -24 + 1e-12

    def check(before):
        ctx.push(sid, dt, deposit=1)
        after = ctx.download(sid)
        exp = orc.push_fast(g, Ex, Ey, bz, cst, before, uniform_bz=True, bz0=float(np.float32(500.0)))
        for a, b in zip(after, exp):
            np.testing.assert_array_equal(a.view(np.uint32), b.view(np.uint32))
        np.testing.assert_array_equal(after[4], before[4])
        _, ke = ctx.reduce(sid)
        ke_exp = orc.kinetic(exp, np.float32(ME), np.float32(1.0))
        assert abs(ke - ke_exp) <= tol * ke_exp, (ke, ke_exp)
        return after

    p = rand_parts(150000, ngx, ngy, 1, vs=3e8, seed=41)
    ctx.upload(sid, *p)
    ctx.sort(sid)
    s = check(ctx.download(sid))
    s = check(s)                                 # the kept sum reused
    G = gp.Gen()
    G.gen_type = 0
    G.xmin, G.xmax = 0.0, ngx * float(np.float32(dx))
    G.ymin, G.ymax = 0.0, ngy * float(np.float32(dx))
    G.ele_vth, G.ion_vth = 4e8, 1e7
    ctx.load_pair(sid, -1, G, 9, 150000, 50000)   # appended (unbinned) particles
    s = check(ctx.download(sid))
    q = list(s)
    q[4] = (2.5 * q[4]).astype(np.float32)      # new vz, same count
    ctx.upload(sid, *q)
    ctx.sort(sid)
    check(ctx.download(sid))
    ctx.close()


def test_edge_cases(gp):
    say("[edges]")
    ctx = mkctx(gp, 16, 16, 3)
    sid = ctx.add_species(-EC, ME, 1.0, 100)
    # empty species: every stage is a no-op
    ctx.sort(sid)
    ctx.reset_grids(); ctx.deposit(sid)
    ctx.push(sid, 1e-12, deposit=1)
    assert ctx.count(sid) == 0
    # particles on cell faces and on the lower domain wall
    x = np.array([1.0, 2.0, 16.0, 9.5], np.float32)
    y = np.array([1.0, 16.999999, 3.0, 9.5], np.float32)
    z = np.zeros(4, np.float32)
    ctx.upload(sid, x, y, z, z, z)
    ctx.reset_grids(); ctx.deposit(sid)
    assert ctx.acc(sid).sum() > 0
    # capacity overflow is an error, not a fault
    with pytest.raises(gp.GampleError):
        big = np.zeros(101, np.float32)
        ctx.upload(sid, big, big, big, big, big)
    ctx.close()


@pytest.mark.parametrize("fbc,ngx,ngy", [
    (("Dirichlet", "Dirichlet", "periodic", "periodic"), 1000, 720),
    (("periodic",) * 4, 1024, 1024),
])
def test_poisson_large_level_tiles(gp, orc, fbc, ngx, ngy):
    """Grids whose finest level has more than 512 tiles of 32 x 32 run the
    level kernels on 64 x 64 tiles (partial tiles at the edges of the 999 x
    720 case).  The solution satisfies the reference's CSR system to 1e-8,
    checked with the oracle's assembly, in as few iterations as the smaller
    grids take (a wrong V-cycle would cost many more or not converge)."""
    say(f"[poisson large tiles] fbc={fbc[0]} {ngx}x{ngy}")
    ctx, st, acc_e, acc_i, dx = _poisson_case(gp, orc, 1, fbc, ngx, ngy, seed=5)
    f = orc.field(ngx, ngy, 0, dx, dx, fbc=fbc)
    sc_e = float(np.float32(float(np.float32(-EC) * np.float32(1000.0)) / (dx * dx)))
    sc_i = float(np.float32(float(np.float32(EC) * np.float32(1000.0)) / (dx * dx)))
    rho = orc.fold(f, acc_e.astype(np.float64) * (2.0 ** -46 * sc_e) + acc_i.astype(np.float64) * (2.0 ** -46 * sc_i))
    ptr, col, val, rbc = orc.poisson_csr(f)
    rhs = orc.poisson_rhs(f, rho, rbc)
    if fbc[0] == "periodic" and fbc[2] == "periodic":
        rhs = rhs - rhs.mean()
    rr = orc.relres(ptr, col, val, rhs, ctx.field(gp.F_SOL).ravel())
    say(f"   iters={st.iterations} relres(gpu)={st.relres:.3e} relres(oracle)={rr:.3e}")
    assert st.relres <= 1e-8 and rr <= 1.01e-8
    assert st.iterations <= 20
    ctx.close()


def test_darwin_solves(gp, orc):
    """Darwin chain (EMField.mm:1138-1329) on a Dirichlet-x / periodic-y grid:
    the three node-grid solves reach 1e-8 on the symmetrised system D*A of
    the reference's CSR with the reference's right-hand sides, and delEx,
    delEy and delBz follow the reference's difference formulas on the device's
    A, on a first call (A_prev = 0) and on a second one after a push."""
    say("[darwin]")
    ngx, ngy, dx, dt = 64, 64, 0.005, 5e-12
    fbc = ("Dirichlet", "Dirichlet", "periodic", "periodic")
    ctx = mkctx(gp, ngx, ngy, 3, fbc=fbc, pbc=("Reflect", "Reflect", "periodic", "periodic"), cur=1, em=2, dx=dx)
    se = ctx.add_species(-EC, ME, 1000.0, 300000, vth_hint=1e8)
    p = rand_parts(250000, ngx, ngy, 1, vs=1e8, seed=41)
    p[2] += np.float32(3e7) * np.sin(2 * np.pi * p[1] / ngy).astype(np.float32)   # a shear flow
    ctx.upload(se, *p)
    ctx.sort(se)
    ctx.reset_grids(); ctx.deposit(se)
    ctx.solve_fields(dt, tol=1e-8)
    sts, rc = ctx.solve_darwin(dt, tol=1e-8, maxiter=400)
    for st in sts:
        assert st.relres <= 1e-8, [x.relres for x in sts]
    jx, jy = ctx.field(gp.F_JX), ctx.field(gp.F_JY)
    chi, ax, ay = ctx.field(gp.F_CHI).ravel(), ctx.field(gp.F_AX).ravel(), ctx.field(gp.F_AY).ravel()
    f = orc.field(ngx, ngy, 1, dx, dx, fbc=fbc)
    ptr, col, val, (nkx, nky) = orc.darwin_csr(f)
    n = len(ptr) - 1
    import scipy.sparse as sp
    A = sp.csr_matrix((val, col, ptr), shape=(n, n))
    d = np.ones((nky + 1, nkx + 1)); d[:, 0] = d[:, -1] = 0.5
    d = d.ravel()
    S = -(sp.diags(d) @ A)

    def check(rhs, x):
        b = -(d * rhs)
        b -= b.mean()
        rr = np.linalg.norm(b - S @ x) / np.linalg.norm(b)
        assert rr <= 1.05e-8, rr
        assert abs(x.mean()) <= 1e-9 * np.abs(x).max()

    check(orc.darwin_rhs_chi(f, jx, jy), chi)
    rx, ry = orc.darwin_rhs_a(f, jx, jy, chi)
    check(rx, ax)
    check(ry, ay)
    ngb = 1
    dxf = float(np.float32(dx))

    def check_delta(ax, ay, lax, lay):
        """EMField.mm:1283-1297 on the node grid, i < nkx, j < nky: delEx at
        Ex[j+ngb, i+ngb+1], delEy at Ey[j+ngb+2, i+ngb] (the reference's row
        offset), delBz = dxAy - dyAx at Bz[j+ngb+1, i+ngb+1]; FP64, then FP32"""
        AX, AY = ax.reshape(nky + 1, nkx + 1), ay.reshape(nky + 1, nkx + 1)
        LX, LY = lax.reshape(nky + 1, nkx + 1), lay.reshape(nky + 1, nkx + 1)
        got = ctx.field(gp.F_DEL_EX)[ngb:ngb + nky, ngb + 1:ngb + 1 + nkx]
        exp = (-(AX[:nky, :nkx] - LX[:nky, :nkx]) / (CL * dt)).astype(np.float32)
        np.testing.assert_allclose(got, exp, rtol=1e-6, atol=1e-6 * np.abs(exp).max())
        got = ctx.field(gp.F_DEL_EY)[ngb + 2:ngb + 2 + nky, ngb:ngb + nkx]
        exp = (-(AY[:nky, :nkx] - LY[:nky, :nkx]) / (CL * dt)).astype(np.float32)
        np.testing.assert_allclose(got, exp, rtol=1e-6, atol=1e-6 * np.abs(exp).max())
        got = ctx.field(gp.F_DEL_BZ)[ngb + 1:ngb + 1 + nky, ngb + 1:ngb + 1 + nkx]
        exp = ((AY[:nky, 1:nkx + 1] - AY[:nky, :nkx]) / dxf - (AX[1:nky + 1, :nkx] - AX[:nky, :nkx]) / dxf)
        exp = exp.astype(np.float32)
        assert np.abs(exp).max() > 0
        np.testing.assert_allclose(got, exp, rtol=1e-6, atol=1e-6 * np.abs(exp).max())

    # first call: the previous A is zero
    check_delta(ax, ay, np.zeros_like(ax), np.zeros_like(ay))
    # second call after a step of the particles: delta E from A - A_prev
    ctx.push(se, dt, deposit=1)
    ctx.solve_fields(dt, tol=1e-8)
    sts, rc = ctx.solve_darwin(dt, tol=1e-8, maxiter=400)
    for st in sts:
        assert st.relres <= 1e-8, [x.relres for x in sts]
    ax2, ay2 = ctx.field(gp.F_AX).ravel(), ctx.field(gp.F_AY).ravel()
    assert np.abs(ax2 - ax).max() > 1e-6 * np.abs(ax).max()
    check_delta(ax2, ay2, ax, ay)
    ctx.close()


def test_wall_injection(gp, orc):
    """Half-Maxwellian flux injection through the x-min wall against the
    oracle's restatement: y bit-exact (uniform draws only); x and v within
    2e-6 relative (device vs libm log, sqrt, sin, cos); the injected
    particles' charge deposit bit-exact against the oracle deposit of the
    device positions; append after existing particles."""
    say("[inject]")
    ngx, ngy, dx, dt = 64, 48, 0.01, 5e-12
    ctx = mkctx(gp, ngx, ngy, 3, pbc=("Delete", "Delete", "periodic", "periodic"),
                fbc=("Dirichlet", "Dirichlet", "periodic", "periodic"), dx=dx)
    sid = ctx.add_species(-EC, ME, 1.0, 200000, vth_hint=1e8)
    p0 = rand_parts(1000, ngx, ngy, 1, seed=8)
    ctx.upload(sid, *p0)
    ctx.reset_grids()
    vth, u = np.float32(1.3e8), (0.0, 2e6, -1e6)
    ctx.inject_wall(sid, 150000, float(vth), u, dt, 77, 4242, deposit=1)
    got = ctx.download(sid)
    assert ctx.count(sid) == 151000
    for a, b in zip(got, p0):
        np.testing.assert_array_equal(a[:1000], b)
    dtdx = float(np.float32(dt / float(np.float32(dx))))
    exp = orc.inject_wall(77, 4242, 150000, float(vth), u, dtdx, 1.0, float(ngx), float(ngy))
    np.testing.assert_array_equal(got[1][1000:], exp[1])
    for k in (0, 2, 3, 4):
        np.testing.assert_allclose(got[k][1000:], exp[k], rtol=2e-6, atol=2e-6 * np.abs(exp[k]).max())
    g = orc.grid(ngx, ngy, 3)
    rho, _, _ = orc.deposit_fx(g, [a[1000:] for a in got])
    np.testing.assert_array_equal(ctx.acc(sid), rho)
    ctx.close()
    # stream 24,807,051 of seed 1 draws r5 = 1 - 2^-24: on a 512-cell wall its
    # y would round onto the upper face, and is kept just below it
    ctx = mkctx(gp, 512, 512, 3, pbc=("Delete", "Delete", "periodic", "periodic"),
                fbc=("Dirichlet", "Dirichlet", "periodic", "periodic"), dx=dx)
    sid = ctx.add_species(-EC, ME, 1.0, 1000, vth_hint=1e8)
    ctx.inject_wall(sid, 1, float(vth), u, dt, 1, 24807051, deposit=1)
    y = ctx.download(sid)[1]
    assert y[0] == np.nextafter(np.float32(513.0), np.float32(0))
    assert y[0] == orc.inject_wall(1, 24807051, 1, float(vth), u, dtdx, 1.0, 512.0, 512.0)[1][0]
    ctx.close()



def _gen_pair(gp, orc, gt, region, ngx, ngy, dx, ele=(1.9e8, (0.0, 0.0, 0.0)), ion=(9e4, (0.0, 0.0, 0.0))):
    G, O = gp.Gen(), orc.Gen()
    for q in (G, O):
        q.gen_type = gt
        q.xmin, q.xmax, q.ymin, q.ymax = region
        q.ele_vth, q.ion_vth = ele[0], ion[0]
        q.ele_u[:] = list(ele[1])
        q.ion_u[:] = list(ion[1])
    O.dx = O.dy = dx
    O.ngb = 1.0
    O.xdom, O.ydom = ngx, ngy
    return G, O


def test_generation_streams(gp, orc):
    """gp_generate (artificial ionization and hollow cathode, Collision.mm:
    464-678) against the oracle's persistent state array: a seeded load of
    5000 pairs, an ionization call of 7000 x-sinusoidal pairs in a
    sub-region (ids below 5000 continue their streams after the load, the
    rest start fresh), a hollow-cathode line source of 3000 electrons (every
    id continues), and a deposited call whose int64 charge equals the
    oracle deposit of the new particles.  y is bit-exact (uniform draws), x
    bit-exact for the uniform and line sources and within 2e-5 cells for asinf,
    velocities within 2e-5 relative (device vs libm transcendentals)."""
    say("[generate] load 5000, ionize 7000, cathode 3000, ionize 2000 deposited")
    ngx, ngy, dx = 64, 48, 0.01
    ctx = mkctx(gp, ngx, ngy, 3, pbc=("Delete", "Delete", "periodic", "periodic"),
                fbc=("Dirichlet", "Dirichlet", "periodic", "periodic"), dx=dx)
    se = ctx.add_species(-EC, ME, 1.0, 20000)
    si = ctx.add_species(EC, 2.18e-22, 1.0, 20000)
    full = (0.0, ngx * dx, 0.0, ngy * dx)
    Gl, Ol = _gen_pair(gp, orc, 0, full, ngx, ngy, dx)
    Ga, Oa = _gen_pair(gp, orc, 1, (0.1, 0.5, 0.0, ngy * dx), ngx, ngy, dx, ion=(9e4, (1e5, 0.0, 0.0)))
    Gh, Oh = _gen_pair(gp, orc, 3, (0.3, 0.3, 0.0, ngy * dx), ngx, ngy, dx, ele=(1.5e8, (0.0, 1e6, 0.0)),
                       ion=(0.0, (0.0, 0.0, 0.0)))
    ctx.load_pair(se, si, Gl, 0, 0, 5000)
    ctx.generate(se, si, Ga, 7000, deposit=0)
    ctx.generate(se, -1, Gh, 3000, deposit=0)
    e, i = ctx.download(se), ctx.download(si)
    assert len(e[0]) == 15000 and len(i[0]) == 12000
    st, inc = orc.stream_states(0, 0, 7000)
    oe0, oi0 = orc.generate(Ol, st, inc, 5000)
    oe1, oi1 = orc.generate(Oa, st, inc, 7000)
    oe2, _ = orc.generate(Oh, st, inc, 3000, ion=False)
    cuts = [(0, 5000, oe0, True), (5000, 12000, oe1, False), (12000, 15000, oe2, True)]
    for lo, hi, oe, exact_x in cuts:
        np.testing.assert_array_equal(e[1][lo:hi], oe[1])
        if exact_x:
            np.testing.assert_array_equal(e[0][lo:hi], oe[0])
        else:
            np.testing.assert_allclose(e[0][lo:hi], oe[0], rtol=0, atol=2e-5)
        for a, b in zip(e[2:], oe[2:]):
            np.testing.assert_allclose(a[lo:hi], b, rtol=2e-5, atol=1e-6 * np.abs(b).max())
    for lo, hi, oi in ((0, 5000, oi0), (5000, 12000, oi1)):
        np.testing.assert_array_equal(i[1][lo:hi], oi[1])
        np.testing.assert_array_equal(i[0][lo:hi], e[0][lo:hi])          # co-located (Collision.mm:150-151)
        for a, b in zip(i[2:], oi[2:]):
            np.testing.assert_allclose(a[lo:hi], b, rtol=2e-5, atol=1e-6 * np.abs(b).max())
    assert np.all(e[0][12000:] == np.float32(np.float32(0.3) / np.float32(dx)) + np.float32(1.0))
    ctx.reset_grids()
    ctx.generate(se, si, Ga, 2000, deposit=1)
    e, i = ctx.download(se), ctx.download(si)
    g = orc.grid(ngx, ngy, 3, pbc=("Delete", "Delete", "periodic", "periodic"))
    for sid, arr in ((se, e), (si, i)):
        rho, _, _ = orc.deposit_fx(g, [a[-2000:] for a in arr])
        np.testing.assert_array_equal(ctx.acc(sid), rho)
    ctx.close()


def test_cathode_ramp(gp, orc):
    """gp_set_cathode (EMField.mm:231-232, 1063-1077): the same charge solved
    without and then with the hollow-cathode entry (the ramp acts after the
    solve, so the unknowns agree to the solve tolerance); phi equals the
    oracle's ramp applied to the oracle reconstruction of the device's
    unknowns to FP64 summation order, E to a few float ulps, and the
    diagnostics report the mean."""
    say("[cathode] 64x48 Dirichlet x, cathode at 0.1234 cm")
    fbc = ("Dirichlet", "Dirichlet", "periodic", "periodic")
    ctx, st, acc_e, acc_i, dx = _poisson_case(gp, orc, 3, fbc, 64, 48)
    sol0 = ctx.field(gp.F_SOL).ravel()
    pos = 0.1234
    ctx.set_cathode(pos)
    ctx.reset_grids()
    for sid in (0, 1):              # the particles are still in the store
        ctx.deposit(sid)
    st2, dg, rc = ctx.solve_fields(0.0, tol=1e-8, maxiter=400, diag=True)
    assert st2.relres <= 1e-8
    sol = ctx.field(gp.F_SOL).ravel()
    assert np.linalg.norm(sol - sol0) <= 1e-6 * np.linalg.norm(sol0)
    f = orc.field(64, 48, 1, dx, dx, fbc=fbc)
    phi = orc.reconstruct(f, sol)
    mean = orc.cathode_ramp(f, sol, pos, phi)
    assert dg.mean_cathode == pytest.approx(mean, rel=1e-13)
    np.testing.assert_allclose(ctx.field(gp.F_PHI), phi, rtol=0, atol=1e-12 * np.abs(phi).max())
    Ex, Ey = orc.gradient(f, phi)
    for got, exp in ((ctx.field(gp.F_EX), Ex), (ctx.field(gp.F_EY), Ey)):
        np.testing.assert_allclose(got, exp, rtol=0, atol=1e-6 * np.abs(exp).max())
    ctx.set_cathode(0.0)
    ctx.close()

@pytest.mark.parametrize("order", [1, 3, 5])
def test_drifted_particles_reach_the_window_edges(gp, orc, order):
    """Particles that have drifted up to GP_DRIFT + 0.5 cells from their tile since the
    last sort: the fused push+deposit (rho and J) stays bit-exact against the
    oracle whether a particle's stencil lands on the edge nodes of the tile's
    LDS window or outside it (global path)."""
    say(f"[drift] order={order}")
    ngx, ngy, dx, dt = 64, 64, 0.01, 2e-12
    ngb = order // 2
    ctx = mkctx(gp, ngx, ngy, order, cur=1, dx=dx)
    sid = ctx.add_species(-EC, ME, 1.0, 300000, vth_hint=1e8)
    rng = np.random.default_rng(40 + order)
    p = rand_parts(200000, ngx, ngy, ngb, seed=50 + order)
    cX = float(np.float32(dt / float(np.float32(dx))))
    d = GP_DRIFT + 0.5
    p[2] = (rng.uniform(-d, d, p[2].size) / cX).astype(np.float32)
    p[3] = (rng.uniform(-d, d, p[3].size) / cX).astype(np.float32)
    ctx.upload(sid, *p)
    ctx.sort(sid)
    s0 = ctx.download(sid)
    z = lambda w: np.zeros(ctx.field_shape(w)[0], np.float32)
    ctx.set_field(gp.F_EX, z(gp.F_EX)); ctx.set_field(gp.F_EY, z(gp.F_EY))
    g = orc.grid(ngx, ngy, order)
    ctx.reset_grids()
    ctx.push(sid, dt, deposit=1)                          # drift, no re-sort
    s1 = orc.push_fast(g, z(gp.F_EX), z(gp.F_EY), z(gp.F_BZ), consts(-EC, ME, dt, dx), s0, uniform_bz=True)
    for a, b in zip(ctx.download(sid), s1):
        np.testing.assert_array_equal(a, b)
    Ex, Ey, _ = rand_fields(ctx, gp, seed=order, scale=2.0)
    ctx.reset_grids()
    ctx.push(sid, dt, deposit=1)
    s2 = orc.push_fast(g, Ex, Ey, z(gp.F_BZ), consts(-EC, ME, dt, dx), s1, uniform_bz=True)
    for a, b in zip(ctx.download(sid), s2):
        np.testing.assert_array_equal(a, b)
    inv = 1.0 / 2.0 ** np.ceil(np.log2(8e8))
    rho, jx, jy = orc.deposit_fx(g, s2, inv_vnorm=inv, current=True)
    np.testing.assert_array_equal(ctx.acc(sid, gp.G_RHO_ACC), rho)
    np.testing.assert_array_equal(ctx.acc(sid, gp.G_JX_ACC), jx)
    np.testing.assert_array_equal(ctx.acc(sid, gp.G_JY_ACC), jy)
    ctx.close()


@pytest.mark.parametrize("order", [1, 3, 5])
def test_periodic_image_gather(gp, orc, order):
    """Particles that crossed a periodic boundary after the last sort stay in
    the segment of their edge tile, one period away from its window.  With E
    from the field solve (exactly periodic) the push gathers them at their
    image inside the window; with E set by the caller it reads global memory.
    Both are bit-exact against the oracle's push at the particle's own
    position, and so is the folded charge."""
    say(f"[image gather] order={order}")
    ngx, ngy, dx, dt = 64, 64, 0.01, 2e-12
    ngb = order // 2
    ctx = mkctx(gp, ngx, ngy, order, dx=dx)
    sid = ctx.add_species(-EC, ME, 1.0, 300000, vth_hint=1e8)
    rng = np.random.default_rng(70 + order)
    p = rand_parts(200000, ngx, ngy, ngb, seed=80 + order)
    cX = float(np.float32(dt / float(np.float32(dx))))
    p[2] = (rng.uniform(-3.0, 3.0, p[2].size) / cX).astype(np.float32)
    p[3] = (rng.uniform(-3.0, 3.0, p[3].size) / cX).astype(np.float32)
    ctx.upload(sid, *p)
    ctx.sort(sid)
    s_sorted = ctx.download(sid)
    ctx.reset_grids(); ctx.deposit(sid)
    st, _ = ctx.solve_fields(dt, tol=1e-8)
    assert st.relres <= 1e-8
    for _ in range(2):                                      # drift up to 6 cells, no re-sort
        ctx.push(sid, dt, deposit=0)
    s0 = ctx.download(sid)
    crossed = (np.abs(s0[0] - s_sorted[0]) > ngx / 2) | (np.abs(s0[1] - s_sorted[1]) > ngy / 2)
    assert crossed.sum() > 2000, crossed.sum()
    Ex, Ey = ctx.field(gp.F_EX), ctx.field(gp.F_EY)
    Bz = np.zeros(ctx.field_shape(gp.F_BZ)[0], np.float32)
    g = orc.grid(ngx, ngy, order)
    cst = consts(-EC, ME, dt, dx)
    state = s0
    for user_set in (False, True):
        if user_set:                                        # the same values, set by the caller
            ctx.set_field(gp.F_EX, Ex); ctx.set_field(gp.F_EY, Ey)
        ctx.reset_grids()
        ctx.push(sid, dt, deposit=1)
        exp = orc.push_fast(g, Ex, Ey, Bz, cst, state, uniform_bz=True)
        got = ctx.download(sid)
        for a, b in zip(got, exp):
            np.testing.assert_array_equal(a.view(np.uint32), b.view(np.uint32))
        rho_exp, _, _ = orc.deposit_fx(g, exp)
        np.testing.assert_array_equal(orc.fold_acc(ctx.acc(sid), ngb, ngx, ngy), orc.fold_acc(rho_exp, ngb, ngx, ngy))
        state = exp
    ctx.close()


def test_stream_probe_leaves_particles(gp):
    """gp_stream_probe streams the species into its scratch store: the live
    particles are unchanged and the rate is a positive time per launch."""
    say("[stream probe]")
    ctx = mkctx(gp, 64, 64, 3)
    sid = ctx.add_species(-EC, ME, 1.0, 100000)
    p = rand_parts(80000, 64, 64, 1, seed=5)
    ctx.upload(sid, *p)
    ms = ctx.stream_probe(sid, 3)
    assert ms > 0.0
    for a, b in zip(ctx.download(sid), p):
        np.testing.assert_array_equal(a, b)
    ctx.close()


def test_fused_step_equals_reference_order(gp):
    """gp_step (solve -> fused push+deposit) against the reference's call
    order per cycle (reset -> deposit each species -> solve -> push each
    species), main.mm:154-254: same particles and fields to the bit."""
    say("[fused vs reference order]")
    ngx, dx, dt = 64, 0.005, 5e-12
    outs = []
    for fused in (True, False):
        ctx = mkctx(gp, ngx, ngx, 3, dx=dx, b0=(0, 0, 200.0))
        se = ctx.add_species(-EC, ME, 800.0, 200000, vth_hint=1e8)
        si = ctx.add_species(EC, 2.18e-22, 800.0, 200000, vth_hint=1e5)
        ctx.upload(se, *rand_parts(150000, ngx, ngx, 1, vs=1e8, seed=61))
        ctx.upload(si, *rand_parts(150000, ngx, ngx, 1, vs=1e5, seed=62))
        ctx.sort(se); ctx.sort(si)
        if fused:
            ctx.reset_grids(); ctx.deposit(se); ctx.deposit(si)
            for cyc in range(1, 6):
                ctx.step(dt, tol=1e-10, maxiter=500, cycle=cyc)
        else:
            for cyc in range(1, 6):
                ctx.reset_grids(); ctx.deposit(se); ctx.deposit(si)
                ctx.solve_fields(dt, tol=1e-10, maxiter=500)
                ctx.push(se, dt, deposit=0); ctx.push(si, dt, deposit=0)
        outs.append([ctx.download(se), ctx.download(si), ctx.field(gp.F_EX), ctx.field(gp.F_EY)])
        ctx.close()
    for a, b in zip(outs[0][0] + outs[0][1], outs[1][0] + outs[1][1]):
        np.testing.assert_array_equal(a, b)
    np.testing.assert_array_equal(outs[0][2], outs[1][2])
    np.testing.assert_array_equal(outs[0][3], outs[1][3])
