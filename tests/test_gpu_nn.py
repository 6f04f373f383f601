"""GPU checks of the learned warm start / preconditioner (gp_nn.hip) through
the C-ABI: forward pass and gradients against the NumPy oracle
(oracle/nn_oracle.py, BF16 rounding emulated at the device's storage points),
training on (b, x) pairs the simulator generates, and the solves that use the
trained network (all reach the same 1e-8 FP64 relative residual as MG-CG).

Tolerances: activations and gradients pass through several BF16 roundings
(spacing 2^-8 relative) whose FP32 inputs differ from the oracle's FP64 sums
in the last bits, so a stored value can land one BF16 step apart; the
comparisons are relative L2 norms per tensor, 2e-2 for the output and 5e-2
for each layer's gradient.
"""
import sys

import numpy as np
import pytest

from oracle import nn_oracle as nno

pytestmark = pytest.mark.gpu

ME, EC = 9.10938188e-28, 4.803204197e-10


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def ctx_for(gp, ngx, ngy, fbc):
    cfg = gp.make_config(ngx, ngy, 0.01, 0.01, order=1, fbc=fbc,
                         pbc=("periodic",) * 4 if fbc[0] == "periodic" else ("Reflect", "Reflect", "periodic", "periodic"))
    return gp.Context(cfg)


def rel(a, b, floor=1e-30):
    return np.linalg.norm(np.ravel(a) - np.ravel(b)) / max(np.linalg.norm(np.ravel(b)), floor)


GEOMS = [(16, 16, ("periodic",) * 4), (24, 16, ("Dirichlet", "Dirichlet", "periodic", "periodic"))]


# the forward also on a grid whose padded image is not a multiple of the
# fused level-0 kernels' 14 x 14 output tiles, periodic in y only
FWD_GEOMS = GEOMS + [(40, 24, ("Dirichlet", "Dirichlet", "periodic", "periodic"))]


@pytest.mark.parametrize("ngx,ngy,fbc", FWD_GEOMS)
def test_inference_forward_equals_training_forward(gp, ngx, ngy, fbc):
    """gp_nn_forward runs the inference chain (level 0 fused: a0 and u0 never
    stored); gp_nn_grad runs the training chain that stores them.  With the
    inference output as the target, the training chain's mean-square loss
    is exactly 0 only if both outputs agree to the bit."""
    say(f"[nn fused forward] {ngx}x{ngy} {fbc[0]}")
    ctx = ctx_for(gp, ngx, ngy, fbc)
    ctx.nn_init(32, seed=5)
    H, W = ctx.nn_dims()
    x = np.random.default_rng(3).standard_normal((3, H, W)).astype(np.float32)
    y = ctx.nn_forward(x)
    loss, _ = ctx.nn_grad(x, y, 0)
    say(f"   H={H} W={W} loss {loss!r}")
    assert np.all(np.isfinite(y)) and np.abs(y).max() > 0
    assert loss == 0.0
    ctx.close()


@pytest.mark.parametrize("ngx,ngy,fbc", FWD_GEOMS)
def test_forward_matches_oracle(gp, ngx, ngy, fbc):
    say(f"[nn forward] {ngx}x{ngy} {fbc[0]}")
    ctx = ctx_for(gp, ngx, ngy, fbc)
    ctx.nn_init(32, seed=7)
    H, W = ctx.nn_dims()
    P = ctx.nn_params()
    rng = np.random.default_rng(1)
    x = rng.standard_normal((2, H, W)).astype(np.float32)
    y = ctx.nn_forward(x)
    perx = fbc[0] == "periodic" and W == ngx
    pery = fbc[2] == "periodic" and H == ngy
    net = nno.Net(P.astype(np.float64), perx, pery, use_bf16=True)
    ye = net.forward(x.astype(np.float64))
    say(f"   H={H} W={W} rel err {rel(y, ye):.2e}")
    assert rel(y, ye) < 2e-2
    ctx.close()


@pytest.mark.parametrize("kind", [0, 1])
@pytest.mark.parametrize("ngx,ngy,fbc", GEOMS)
def test_gradients_match_oracle(gp, ngx, ngy, fbc, kind):
    say(f"[nn grad] {ngx}x{ngy} {fbc[0]} loss kind {kind}")
    ctx = ctx_for(gp, ngx, ngy, fbc)
    ctx.nn_init(32, seed=11)
    H, W = ctx.nn_dims()
    P = ctx.nn_params()
    rng = np.random.default_rng(2)
    x = rng.standard_normal((2, H, W)).astype(np.float32)
    t = rng.standard_normal((2, H, W)).astype(np.float32)
    loss, g = ctx.nn_grad(x, t, kind)
    perx = fbc[0] == "periodic" and W == ngx
    pery = fbc[2] == "periodic" and H == ngy
    mx = ngx if fbc[0] == "periodic" else ngx - 1
    my = ngy if fbc[2] == "periodic" else ngy - 1
    eg = (mx, my, fbc[0] == "periodic", fbc[2] == "periodic", 1.0)
    le, ge = nno.Net(P.astype(np.float64), perx, pery, use_bf16=True).loss_grad(x.astype(np.float64),
                                                                                   t.astype(np.float64), kind, eg)
    say(f"   loss {loss:.6f} oracle {le:.6f}")
    assert abs(loss - le) <= 1e-2 * le
    offs, n = nno.layout()
    for l, (wo, bo) in enumerate(offs):
        co = nno.CO[l]
        # floor: a gradient that vanishes in exact arithmetic (the output
        # bias under the periodic energy loss, whose S1 rows sum to zero) is
        # compared against the size of the whole gradient
        fl = 1e-4 * np.linalg.norm(ge)
        ew, eb = rel(g[wo:bo], ge[wo:bo], fl), rel(g[bo:bo + co], ge[bo:bo + co], fl)
        say(f"   layer {l}: weight grad rel err {ew:.2e}, bias {eb:.2e}")
        assert ew < 5e-2 and eb < 5e-2, l
    ctx.close()


def harvest(gp, ngx, nstep, seed=0):
    """(b, x) pairs from a small electrostatic run (two species, periodic)."""
    cfg = gp.make_config(ngx, ngx, 0.005, 0.005, order=1, b0=(0, 0, 300.0))
    ctx = gp.Context(cfg)
    n = 64 * ngx * ngx
    se = ctx.add_species(-EC, ME, 4e3, n, vth_hint=1e8)
    si = ctx.add_species(EC, 2.18e-22, 4e3, n, vth_hint=1e5)
    G = gp.Gen()
    G.gen_type = 7                      # cosine perturbation of the density
    G.xmin, G.xmax, G.ymin, G.ymax = 0.0, ngx * 0.005, 0.0, ngx * 0.005
    G.ele_vth, G.ion_vth = 1.9e8, 1.2e5
    G.pert_delta, G.pert_mode = 0.05, 3
    ctx.load_pair(se, si, G, seed, 0, n)
    ctx.sort(se); ctx.sort(si)
    ctx.reset_grids(); ctx.deposit(se); ctx.deposit(si)
    return ctx


def test_training_and_solves(gp, orc):
    """Train on (b, x) pairs from the simulation plus random power-law
    right-hand sides, then: the training loss falls; the trained network's
    warm start leaves a smaller energy-norm error than the untrained one, and
    its flexible CG needs fewer iterations; every network solver mode
    reaches the same 1e-8 FP64 residual as MG-CG."""
    from gample_amd.nn_data import add_random_pairs, random_rhs
    say("[nn train+solve] 64x64 periodic")
    ngx = 64
    ctx = harvest(gp, ngx, 0)
    ctx.nn_init(32, seed=3)
    dt = 4e-12
    for k in range(24):
        ctx.step(dt, tol=1e-8, maxiter=400, cycle=k + 1)
        ctx.nn_add_pair()
    rng = np.random.default_rng(5)
    add_random_pairs(ctx, 72, rng, True, True)
    # held-out test right-hand side and its MG-CG solution
    bt = random_rhs(ngx, ngx, True, True, np.random.default_rng(99), slope=1.5)
    xt, st0, _ = ctx.solve_rhs(bt, solver=0)
    f = orc.field(ngx, ngx, 0, 0.005, 0.005, fbc=("periodic",) * 4)
    ptr, col, val, rbc = orc.poisson_csr(f)
    import scipy.sparse as sp
    S = -sp.csr_matrix((val, col, ptr), shape=(len(ptr) - 1, len(ptr) - 1))

    def warm_err(z):
        """relative energy-norm error of the S-optimal multiple of z"""
        z = z - z.mean()
        a = (z @ bt) / (z @ (S @ z))
        e = xt - xt.mean() - a * z
        return np.sqrt(e @ (S @ e)) / np.sqrt((xt - xt.mean()) @ (S @ (xt - xt.mean())))

    e0 = warm_err(ctx.nn_infer(bt))
    ctx.nn_config(1, 2, 400)                     # no fallback: measure the flexible CG itself
    _, st_u, _ = ctx.solve_rhs(bt, solver=3)
    l0 = ctx.nn_train(epochs=1, batch=4, lr=0.0)
    lfit = ctx.nn_train(epochs=120, batch=4, lr=1e-3)
    e1 = warm_err(ctx.nn_infer(bt))
    _, st_t, _ = ctx.solve_rhs(bt, solver=3)
    say(f"   loss {l0:.4f} -> {lfit:.4f}; warm-start energy error untrained {e0:.3e} trained {e1:.3e}")
    say(f"   flexible CG iterations: untrained {st_u.iterations} (fell back {st_u.fell_back}), "
        f"trained {st_t.iterations} (fell back {st_t.fell_back}); MG-CG {st0.iterations}")
    assert lfit < 0.5 * l0
    assert e1 < 0.5 * e0
    assert st_t.fell_back == 0 and st_t.iterations < 0.5 * st_u.iterations
    ctx.nn_config(1, 2, 0)                       # default: fall back after the MG-CG count
    for solver in (0, 1, 2, 3):
        x, st, rc = ctx.solve_rhs(bt, solver=solver)
        say(f"   solver {solver}: iters {st.iterations} refine {st.refinements} relres {st.relres:.2e} "
            f"t {1e3 * st.seconds:.2f} ms nn {st.used_nn} fell_back {st.fell_back}")
        assert st.relres <= 1e-8
        assert st.used_nn == (solver != 0)
        assert np.linalg.norm((x - x.mean()) - (xt - xt.mean())) <= 1e-6 * np.linalg.norm(xt - xt.mean())
    # the simulator's own field solve with the network
    ctx.set_field(gp.F_SOL, np.zeros(ngx * ngx))
    st, rc = ctx.solve_fields(dt, tol=1e-8, maxiter=400, solver=1)
    assert st.relres <= 1e-8 and st.used_nn == 1
    ctx.close()


def test_device_increments_equal_host_increments(gp):
    """gp_nn_add_increment forms (b_n - b_{n-1}, x_n - x_{n-1}) of the last two
    solves on the device.  A network trained on those pairs sees exactly the
    pairs the caller would form from host copies: the training losses of two
    equally seeded networks, one fed each way, agree to the bit."""
    say("[nn increments]")
    ctx = harvest(gp, 64, 0, seed=3)
    ctx.nn_init(32, seed=1)
    my, mx = ctx.field_shape(gp.F_SOL)[0]
    bs, xs, added = [], [], []
    for cyc in range(1, 8):
        ctx.step(5e-12, tol=1e-10, maxiter=400, cycle=cyc)
        bs.append(ctx.field(gp.F_RHS).ravel().copy())
        xs.append(ctx.field(gp.F_SOL).ravel().copy())
        added.append(ctx.nn_add_increment())
    assert added == [False] + [True] * 6
    db = np.array([bs[k] - bs[k - 1] for k in range(1, 7)])
    dx = np.array([xs[k] - xs[k - 1] for k in range(1, 7)])
    l_dev = [ctx.nn_train(epochs=1, batch=2, lr=0.0) for _ in range(2)]
    other = harvest(gp, 64, 0, seed=4)
    other.nn_init(32, seed=1)
    other.nn_train(epochs=0, rho=db, phi=dx)
    l_host = [other.nn_train(epochs=1, batch=2, lr=0.0) for _ in range(2)]
    assert l_dev == l_host and np.isfinite(l_dev[0]) and l_dev[0] > 0, (l_dev, l_host)
    # and training moves both the same way
    a = ctx.nn_train(epochs=2, batch=2, lr=1e-3)
    b = other.nn_train(epochs=2, batch=2, lr=1e-3)
    assert a == b, (a, b)
    ctx.close()
    other.close()


def test_darwin_fine_tune(gp):
    """A network trained on Poisson increments is fine-tuned on the Darwin
    solves' (chi, Ax, Ay) increments: switching the training system keeps the
    weights and starts an empty Darwin set, the first Darwin call only records
    the systems, later ones add three pairs, training lowers the loss, and the
    Darwin solves served by the network reach the same 1e-8 residual."""
    say("[nn darwin]")
    ngx = 64
    cfg = gp.make_config(ngx, ngx, 0.005, 0.005, order=1, pbc=("Reflect", "Reflect", "periodic", "periodic"),
                         fbc=("Dirichlet", "Dirichlet", "periodic", "periodic"), b0=(0, 0, 300.0),
                         cur_dens=1, em_model=2)
    ctx = gp.Context(cfg)
    n = 64 * ngx * ngx
    se = ctx.add_species(-EC, ME, 4e3, n, vth_hint=1e8)
    si = ctx.add_species(EC, 2.18e-22, 4e3, n, vth_hint=1e5)
    G = gp.Gen()
    G.gen_type = 7
    G.xmin, G.xmax, G.ymin, G.ymax = 0.0, ngx * 0.005, 0.0, ngx * 0.005
    G.ele_vth, G.ion_vth = 1.9e8, 1.2e5
    G.pert_delta, G.pert_mode = 0.05, 3
    ctx.load_pair(se, si, G, 5, 0, n)
    ctx.sort(se); ctx.sort(si)
    ctx.reset_grids(); ctx.deposit(se); ctx.deposit(si)
    ctx.nn_init(32, seed=1)
    cyc = 0
    for _ in range(6):
        cyc += 1
        ctx.step(5e-12, tol=1e-8, maxiter=400, cycle=cyc)
        ctx.nn_add_increment()
    ctx.nn_train(epochs=3, batch=2, lr=1e-3)
    w0 = ctx.nn_params()
    with pytest.raises(gp.GampleError):
        ctx.nn_add_darwin_increments()               # the training set is still the Poisson one
    ctx.nn_set_system(1)
    np.testing.assert_array_equal(ctx.nn_params(), w0)
    with pytest.raises(gp.GampleError):
        ctx.nn_add_increment()                       # and now the Darwin one
    added = []
    for _ in range(5):
        cyc += 1
        ctx.step(5e-12, tol=1e-8, maxiter=400, cycle=cyc)
        added.append(ctx.nn_add_darwin_increments())
    assert added == [0, 3, 3, 3, 3]
    l0 = ctx.nn_train(epochs=1, batch=2, lr=0.0)
    l1 = None
    for _ in range(20):
        l1 = ctx.nn_train(epochs=1, batch=2, lr=1e-3)
    assert l1 < l0, (l0, l1)
    for solver in (2, 1):
        cyc += 1
        st, _ = ctx.step(5e-12, tol=1e-8, maxiter=400, solver=solver, cycle=cyc)
        assert st.relres <= 1e-8
        for d in ctx.darwin_stats():
            assert d.relres <= 1e-8, (solver, d.relres)
    ctx.close()
