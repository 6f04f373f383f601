"""Pins of the oracle's generators (src/Collision.h:11-16,
src/Collision.mm:21-30, 63-174, 280-336) against published known-answer
vectors (tests/golden/rng_kat.json, see make_kat.py) and distribution
properties of the seeded loader."""
import ctypes as C
import json
import os

import numpy as np
import pytest

KAT = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "rng_kat.json")))


def test_splitmix64_kat(orc):
    s = C.c_uint64(1234567)
    got = [orc.lib().gpo_splitmix64(C.byref(s)) for _ in range(5)]
    assert got == KAT["splitmix64_seed_1234567"]
    s = C.c_uint64(0)
    assert orc.lib().gpo_splitmix64(C.byref(s)) == KAT["splitmix64_seed_0_first"]


def test_pcg32_kat(orc):
    # pcg32_srandom_r(42, 54): state = 0, inc = 54*2+1, step, state += 42, step
    inc = (54 << 1) | 1
    st = C.c_uint64(0)
    orc.lib().gpo_pcg32(C.byref(st), inc)
    st.value = (st.value + 42) & ((1 << 64) - 1)
    orc.lib().gpo_pcg32(C.byref(st), inc)
    got = [orc.lib().gpo_pcg32(C.byref(st), inc) for _ in range(6)]
    assert got == KAT["pcg32_srandom_42_54"]


def test_rand01_is_the_reference_draw(orc):
    """rand01 = float(u) * 2^-32 (src/Collision.mm:28-30): NumPy's uint32 ->
    float32 conversion (round to nearest even) times 2^-32, for random u and
    for the edges of the rounding, outside the two guarded cases."""
    f = orc.lib().gpo_rand01
    rng = np.random.default_rng(5)
    top = 2 ** 32 - 128          # from here the conversion rounds to 2^32
    us = np.concatenate([rng.integers(1, top, 20000, dtype=np.uint64),
                         np.arange(1, 600, dtype=np.uint64),                  # exact small values
                         np.arange(2 ** 24 - 8, 2 ** 24 + 40, dtype=np.uint64),  # first rounding
                         np.arange(top - 400, top, dtype=np.uint64)]).astype(np.uint32)
    exp = us.astype(np.float32) * np.float32(2.0 ** -32)
    got = np.array([f(int(u)) for u in us], dtype=np.float32)
    assert np.array_equal(got.view(np.uint32), exp.view(np.uint32))
    assert f(top - 129) == 1.0 - 2.0 ** -24      # 2^32 - 256 exactly: a reference value
    # guarded: log(0), and the 128 values whose conversion rounds to 1.0
    assert f(0) == 2.0 ** -33
    for u in (top, top + 1, 2 ** 32 - 1):
        assert f(u) == 1.0 - 2.0 ** -24


def _gen(orc, gt=0, vth_e=1e8, vth_i=1e5, L=1.0):
    g = orc.Gen()
    g.gen_type = gt
    g.xmin, g.xmax, g.ymin, g.ymax = 0.0, L, 0.0, L
    g.ele_u[:] = [0.0, 0.0, 0.0]
    g.ion_u[:] = [0.0, 0.0, 0.0]
    g.ele_vth, g.ion_vth = vth_e, vth_i
    g.dx = g.dy = L / 64
    g.ngb = 1.0
    g.pert_delta, g.pert_mode = 0.0, 0
    return g


def test_loader_uniform_maxwellian(orc):
    n = 200000
    g = _gen(orc)
    e, i = orc.load_pair(g, 0, 0, n)
    for a in (e[0], e[1]):
        assert a.min() >= 1.0 and a.max() < 65.0
        h = np.histogram(a, bins=16, range=(1, 65))[0]
        assert np.all(np.abs(h / (n / 16) - 1) < 0.03)
    np.testing.assert_array_equal(e[0], i[0])      # ions co-located (Collision.mm:150-151)
    np.testing.assert_array_equal(e[1], i[1])
    # Box-Muller with sqrt(-ln r): per-component variance vth^2/2
    for v, vth in ((e[2], 1e8), (e[3], 1e8), (e[4], 1e8), (i[2], 1e5)):
        assert abs(np.mean(v)) < 0.01 * vth
        assert abs(np.var(v.astype(np.float64)) / (vth * vth / 2) - 1) < 0.02


def test_loader_counter_based(orc):
    """Particle id k draws from its own stream, so a slice [a, b) generated
    alone equals the same slice of a full generation."""
    g = _gen(orc)
    full, _ = orc.load_pair(g, 7, 0, 5000)
    part, _ = orc.load_pair(g, 7, 1200, 800)
    for a, b in zip(full, part):
        np.testing.assert_array_equal(a[1200:2000], b)


def test_loader_cosine_perturbation(orc):
    """gen_type 7: density 1 + delta cos(2 pi m (x - xmin)/L)."""
    n = 400000
    g = _gen(orc, gt=7)
    g.pert_delta, g.pert_mode = 0.2, 3
    e, _ = orc.load_pair(g, 3, 0, n, ion=False)
    xs = (e[0] - 1.0) / 64.0
    h, edges = np.histogram(xs, bins=48, range=(0, 1))
    c = 0.5 * (edges[1:] + edges[:-1])
    expect = n / 48 * (1 + 0.2 * np.cos(2 * np.pi * 3 * c))
    assert np.max(np.abs(h - expect) / expect) < 0.05


def test_loader_keeps_positions_below_the_upper_face(orc):
    """uniform positions in [0, L) cm can round onto ngb + ngx in grid units;
    with the domain extent given, the loader moves them just below the face"""
    n, ngx, dx = 400000, 64, np.float32(5.2563e-3)
    g = orc.Gen()
    g.gen_type = 0
    g.xmin, g.xmax, g.ymin, g.ymax = 0.0, float(ngx * dx), 0.0, float(ngx * dx)
    g.ele_vth, g.ion_vth = 1.0, 1.0
    g.dx = g.dy = float(dx)
    g.ngb = 0.0
    raw, _ = orc.load_pair(g, 0, 0, n, ion=False)
    g.xdom = g.ydom = ngx
    fixed, _ = orc.load_pair(g, 0, 0, n, ion=False)
    moved = raw[0] != fixed[0]
    assert np.all(raw[0][moved] == ngx) and np.all(fixed[0][moved] == np.nextafter(np.float32(ngx), 0))
    assert fixed[0].max() < ngx and fixed[1].max() < ngx


M64 = (1 << 64) - 1


def test_stream_states_follow_the_sequential_init(orc):
    """gpo_stream_states equals the reference's id-order initialisation
    (Collision.mm:300-313, 507-520): state then inc|1 drawn from one
    splitmix64 sequence, here run step by step with the pinned splitmix64"""
    for seed in (0, 987654321):
        m = C.c_uint64(seed)
        st_seq, inc_seq = [], []
        for _ in range(300):
            st_seq.append(orc.lib().gpo_splitmix64(C.byref(m)))
            inc_seq.append(orc.lib().gpo_splitmix64(C.byref(m)) | 1)
        st, inc = orc.stream_states(seed, 0, 300)
        assert st.tolist() == st_seq and inc.tolist() == inc_seq
        st, inc = orc.stream_states(seed, 137, 50)
        assert st.tolist() == st_seq[137:187] and inc.tolist() == inc_seq[137:187]


@pytest.mark.parametrize("gt,ion", [(0, True), (1, True), (2, True), (5, True), (6, False), (7, True), (3, False)])
def test_generate_on_fresh_states_is_the_load(orc, gt, ion):
    """a generation call on freshly seeded states draws exactly what the
    initial load does (both are the kernel of Collision.mm:63-174)"""
    g = _gen(orc, gt)
    g.pert_delta, g.pert_mode = 0.02, 3
    n = 3000
    st, inc = orc.stream_states(0, 0, n)
    e, i = orc.generate(g, st, inc, n, ion=ion)
    le, li = orc.load_pair(g, 0, 0, n, ion=ion)
    for a, b in zip(e, le):
        np.testing.assert_array_equal(a, b)
    if ion and gt != 3:
        for a, b in zip(i, li):
            np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("gt,ion,draws", [(0, True, 10), (1, True, 10), (5, True, 11), (6, False, 7),
                                          (3, True, 6), (3, False, 6), (4, False, 6)])
def test_generate_advances_each_stream_by_its_draws(orc, gt, ion, draws):
    """the kernel writes each thread's state back after exactly the numbers
    it drew (Collision.mm:136-139, 171-172): 2 for the position (3 for the
    Gaussian profiles), 4 for the electron velocity, 4 for the ion's unless
    the call is the hollow cathode's; stepping the PCG32 LCG that many times
    from the initial state must land on the written-back state"""
    g = _gen(orc, gt)
    n = 257
    st0, inc = orc.stream_states(11, 0, n)
    st = st0.copy()
    orc.generate(g, st, inc, n, ion=ion)
    for k in range(n):
        s = int(st0[k])
        for _ in range(draws):
            s = (s * 6364136223846793005 + int(inc[k])) & M64
        assert s == int(st[k])


def test_cathode_ramp_pins_the_cathode_column(orc):
    """EMField.mm:1063-1077 on a random solution: the ramp subtracts a
    linear function of x only, so Ey is unchanged and Ex moves by the
    uniform mean/pos; it vanishes at phi index ngb+2; and on a Dirichlet-x
    grid (ikmin = ngb + 1) the row mean of phi on the cathode column falls to
    mean (1 - col dx / pos), col = int(pos / dx)"""
    f = orc.Field()
    f.ngx, f.ngy, f.ngb, f.dx, f.dy = 40, 24, 1, 0.005, 0.005
    f.fbc[:] = [1, 1, 0, 0]
    f.fval[:] = [30.0, 0.0, 0.0, 0.0]
    nkx, nky, ikmin, _ = orc.poisson_dims(f)
    assert ikmin == f.ngb + 1
    rng = np.random.default_rng(4)
    sol = rng.standard_normal((nky + 1) * (nkx + 1)) + 5.0
    pos = 0.1234
    col = int(pos / np.float32(0.005))
    phi0 = orc.reconstruct(f, sol)
    phi1 = phi0.copy()
    mean = orc.cathode_ramp(f, sol, pos, phi1)
    assert mean == pytest.approx(sol.reshape(nky + 1, nkx + 1)[:, col].mean(), rel=1e-14)
    np.testing.assert_array_equal(phi1[:, f.ngb + 2], phi0[:, f.ngb + 2])
    ex0, ey0 = orc.gradient(f, phi0)
    ex1, ey1 = orc.gradient(f, phi1)
    np.testing.assert_allclose(ey1, ey0, rtol=0, atol=1e-9 * np.abs(ey0).max())
    np.testing.assert_allclose(ex1 - ex0, mean / pos, rtol=1e-5)
    ci = col + ikmin + 1                                     # phi index of unknown column col
    rows = slice(1 + f.ngb, 2 + f.ngb + nky)                 # the unknown rows (periodic y)
    before = phi0[rows, ci].mean()
    assert before == pytest.approx(mean, rel=1e-12)
    xc = float(np.float32(col) * np.float32(0.005))          # the float product of the reference
    assert phi1[rows, ci].mean() == pytest.approx(mean * (1 - xc / pos), rel=1e-9)
