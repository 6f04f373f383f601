"""Pins of the CPU oracle's particle routines (weights, gather, Boris, BCs).

Each check is chosen so that a plausible slip in the restatement (a swapped
weight, an index off by one, a wrong sign in the rotation) fails it:
closed-form weight values and centroid, exact gather of linear fields,
Boris invariants, and agreement of the kernel-order push with the
reference-order push (src/Particle.mm:65-278).
"""
import math

import numpy as np
import pytest


@pytest.mark.parametrize("order", [1, 3, 5])
def test_weights_closed_form(orc, order):
    rng = np.random.default_rng(1)
    hs = np.concatenate([[0.0, 0.25, 0.5, 0.75, 1 - 2 ** -24], rng.random(2000)]).astype(np.float32)
    ngb = order // 2
    for h in hs:
        wf = orc.weights(order, float(h), "fast").astype(np.float64)
        wr = orc.weights(order, float(h), "ref").astype(np.float64)
        w64 = orc.weights(order, float(h), "f64")
        assert np.all(w64 >= -1e-15)
        assert abs(w64.sum() - 1.0) < 1e-13                      # partition of unity
        k = np.arange(order + 1) - ngb
        assert abs((w64 * k).sum() - float(h)) < 1e-12            # linear precision (centroid = h)
        assert np.max(np.abs(wf - w64)) < 2e-7                    # kernel order vs FP64 closed form
        assert np.max(np.abs(wr - w64)) < 2e-6                    # reference powf form vs FP64
        assert abs(wf.sum() - 1.0) < 4e-7


def test_weights_known_values(orc):
    # cubic B-spline: h=0 -> [1/6, 2/3, 1/6, 0], h=1/2 -> [1, 23, 23, 1]/48
    np.testing.assert_allclose(orc.weights(3, 0.0, "f64"), [1 / 6, 2 / 3, 1 / 6, 0], atol=1e-15)
    np.testing.assert_allclose(orc.weights(3, 0.5, "f64"), np.array([1, 23, 23, 1]) / 48, atol=1e-15)
    # quintic B-spline: h=0 -> [1, 26, 66, 26, 1, 0]/120, h=1/2 -> [1, 237, 1682, 1682, 237, 1]/3840
    np.testing.assert_allclose(orc.weights(5, 0.0, "f64"), np.array([1, 26, 66, 26, 1, 0]) / 120, atol=1e-15)
    np.testing.assert_allclose(orc.weights(5, 0.5, "f64"), np.array([1, 237, 1682, 1682, 237, 1]) / 3840,
                               atol=1e-15)
    np.testing.assert_allclose(orc.weights(5, 0.5, "fast"), np.array([1, 237, 1682, 1682, 237, 1]) / 3840,
                               atol=3e-8)
    np.testing.assert_allclose(orc.weights(1, 0.3, "fast"), [0.7, 0.3], atol=1e-7)


def _fields(ngx, ngy, order, kind="random", seed=0):
    ngb = order // 2
    nx, ny = ngx + 2 * ngb, ngy + 2 * ngb
    rng = np.random.default_rng(seed)
    if kind == "random":
        Ex = rng.standard_normal((ny + 1, nx + 2)).astype(np.float32)
        Ey = rng.standard_normal((ny + 2, nx + 1)).astype(np.float32)
        Bz = (1.0 + 0.1 * rng.standard_normal((ny + 2, nx + 2))).astype(np.float32)
    return Ex, Ey, Bz


def _particles(n, ngx, ngy, ngb, vscale=0.3, seed=0):
    rng = np.random.default_rng(seed)
    x = (ngb + rng.random(n) * ngx).astype(np.float32)
    y = (ngb + rng.random(n) * ngy).astype(np.float32)
    v = (vscale * rng.standard_normal((3, n))).astype(np.float32)
    return [x, y, v[0], v[1], v[2]]


@pytest.mark.parametrize("order", [1, 3, 5])
def test_gather_linear_field_exact(orc, order):
    """B-splines reproduce linear fields: with B = 0 and constE = 1 the push
    returns v' = v + 2 E(x_p) exactly, so E(x_p) is read back from v'."""
    ngx, ngy = 16, 12
    ngb = order // 2
    nx, ny = ngx + 2 * ngb, ngy + 2 * ngb
    a, bx, cy = 0.5, 0.25, -0.125
    m = np.arange(nx + 2)[None, :]
    j = np.arange(ny + 1)[:, None]
    Ex = (a + bx * (m - 0.5) + cy * j).astype(np.float32)          # Ex index m at x = m - 1/2
    mm = np.arange(nx + 1)[None, :]
    jj = np.arange(ny + 2)[:, None]
    Ey = (-a + cy * mm + bx * (jj - 0.5)).astype(np.float32)       # Ey index j at y = j - 1/2
    Bz = np.zeros((ny + 2, nx + 2), np.float32)
    g = orc.grid(ngx, ngy, order, pbc=("periodic",) * 4)
    p = _particles(4000, ngx - 2, ngy - 2, ngb + 1, vscale=0.0, seed=3)
    out = orc.push_fast(g, Ex, Ey, Bz, (1.0, 0.0, 0.0, 0.0), p)
    ex_exact = a + bx * p[0].astype(np.float64) + cy * p[1].astype(np.float64)
    ey_exact = -a + cy * p[0].astype(np.float64) + bx * p[1].astype(np.float64)
    np.testing.assert_allclose(out[2] / 2.0, ex_exact, rtol=0, atol=2e-5)
    np.testing.assert_allclose(out[3] / 2.0, ey_exact, rtol=0, atol=2e-5)
    ref = orc.push_ref(g, Ex, Ey, Bz, (1.0, 0.0, 0.0, 0.0), p)
    np.testing.assert_allclose(ref[2] / 2.0, ex_exact, rtol=0, atol=2e-5)


@pytest.mark.parametrize("boris", [2, 4, 6])
def test_boris_rotation(orc, boris):
    """E = 0, uniform Bz: |v| is conserved and the velocity turns clockwise
    by 2*atan(t), t = constB*Bz (2nd order); orders 4/6 scale t by fN1."""
    ngx = ngy = 8
    g = orc.grid(ngx, ngy, 3, pbc=("periodic",) * 4, boris=boris)
    nx = ngx + 2
    Ex = np.zeros((nx + 1, nx + 2), np.float32)
    Ey = np.zeros((nx + 2, nx + 1), np.float32)
    Bz = np.full((nx + 2, nx + 2), 1.0, np.float32)
    p = _particles(500, ngx, ngy, 1, vscale=1.0, seed=5)
    cB = 0.2
    for out in (orc.push_fast(g, Ex, Ey, Bz, (0.0, cB, 0.0, 0.0), p, uniform_bz=True, bz0=1.0),
                orc.push_ref(g, Ex, Ey, Bz, (0.0, cB, 0.0, 0.0), p)):
        v0 = np.sqrt(p[2].astype(np.float64) ** 2 + p[3] ** 2)
        v1 = np.sqrt(out[2].astype(np.float64) ** 2 + out[3] ** 2)
        if boris == 2:
            np.testing.assert_allclose(v1, v0, rtol=2e-6)   # 4/6 keep btbt unscaled (Particle.mm:199-201)
        t = cB
        if boris == 4:
            t = cB * (1 + cB * cB / 3)
        if boris == 6:
            t = cB * (1 + cB * cB / 3 + 2 * cB ** 4 / 15)
        th = 2 * math.atan(t) if boris == 2 else 2 * math.atan2(t, 1 + 0 * t)
        # 2nd order uses s = 2t/(1+t^2) with |t| the scaled value in the
        # reference only when boris == 2 (orders 4/6 keep btbt unscaled)
        if boris != 2:
            s = 2 * t / (1 + cB * cB)
            # rotation with v0 = u + u x t, u+ = u + v0 x s: exact closed form
            c1 = 1 - s * t
            s1 = s
            vx = p[2] * c1 + p[3] * s1
            vy = p[3] * c1 - p[2] * s1
        else:
            vx = p[2] * math.cos(th) + p[3] * math.sin(th)
            vy = -p[2] * math.sin(th) + p[3] * math.cos(th)
        np.testing.assert_allclose(out[2], vx, atol=3e-6)
        np.testing.assert_allclose(out[3], vy, atol=3e-6)
        np.testing.assert_array_equal(out[4], p[4])


def test_boris_pure_e(orc):
    """B = 0: v' = v + 2 constE E exactly for constE a power of two."""
    ngx = ngy = 8
    g = orc.grid(ngx, ngy, 1, pbc=("periodic",) * 4)
    Ex = np.full((ngy + 1, ngx + 2), 0.75, np.float32)
    Ey = np.full((ngy + 2, ngx + 1), -1.5, np.float32)
    Bz = np.zeros((ngy + 2, ngx + 2), np.float32)
    p = _particles(300, ngx, ngy, 0, vscale=1.0, seed=2)
    out = orc.push_fast(g, Ex, Ey, Bz, (0.5, 0.0, 0.0, 0.0), p)
    np.testing.assert_allclose(out[2], p[2] + 0.75, atol=1e-6)
    np.testing.assert_allclose(out[3], p[3] - 1.5, atol=1e-6)


@pytest.mark.parametrize("order", [1, 3, 5])
@pytest.mark.parametrize("boris", [2, 4, 6])
def test_fast_vs_reference_order(orc, order, boris):
    """Kernel op order vs the reference's literal FP32 expression order: one
    step agrees to 1e-5 relative; cell indices agree except for particles
    within 1e-4 cell of a cell face (the two orders differ in the last bits,
    which can move a particle that close to a face across it)."""
    ngx, ngy = 32, 24
    ngb = order // 2
    Ex, Ey, Bz = _fields(ngx, ngy, order, seed=order + boris)
    g = orc.grid(ngx, ngy, order, pbc=("periodic",) * 4, boris=boris)
    p = _particles(20000, ngx, ngy, ngb, vscale=1.0, seed=7)
    consts = (0.05, 0.1, 0.2, 0.2)
    f = orc.push_fast(g, Ex, Ey, Bz, consts, p)
    r = orc.push_ref(g, Ex, Ey, Bz, consts, p)
    for a, b in zip(f[:5], r[:5]):
        np.testing.assert_allclose(a, b, rtol=1e-5, atol=1e-5 * np.abs(b).max())
    cf, cr = np.floor(f[0]), np.floor(r[0])
    frac = np.abs(r[0] - np.round(r[0]))
    assert np.all((cf == cr) | (frac < 1e-4))


def test_particle_bcs(orc):
    ngx, ngy, order = 16, 16, 3
    ngb = 1
    g = orc.grid(ngx, ngy, order, pbc=("Delete", "Delete", "periodic", "periodic"))
    nx = ngx + 2
    Ex = np.zeros((nx + 1, nx + 2), np.float32)
    Ey = np.zeros((nx + 2, nx + 1), np.float32)
    Bz = np.zeros((nx + 2, nx + 2), np.float32)
    x = np.array([1.05, 16.95, 8.0, 8.0], np.float32)
    y = np.array([8.0, 8.0, 1.02, 16.99], np.float32)
    vx = np.array([-1.0, 1.0, 0.0, 0.0], np.float32)
    vy = np.array([0.0, 0.0, -1.0, 1.0], np.float32)
    vz = np.zeros(4, np.float32)
    out = orc.push_fast(g, Ex, Ey, Bz, (0.0, 0.0, 0.1, 0.1), [x, y, vx, vy, vz])
    assert out[0][0] == -1.0 and out[0][1] == -2.0          # absorbed markers x = -piflag
    # periodic in y: 1.02 - 0.1 -> wraps to the top, 16.99 + 0.1 -> bottom
    np.testing.assert_allclose(out[1][2], 1.02 - 0.1 + 16, atol=1e-5)
    np.testing.assert_allclose(out[1][3], 16.99 + 0.1 - 16, atol=1e-5)
    ref = orc.push_ref(g, Ex, Ey, Bz, (0.0, 0.0, 0.1, 0.1), [x, y, vx, vy, vz])
    assert list(ref[5]) == [1, 2, 0, 0]
    g2 = orc.grid(ngx, ngy, order, pbc=("Reflect", "Reflect", "periodic", "periodic"))
    out = orc.push_fast(g2, Ex, Ey, Bz, (0.0, 0.0, 0.1, 0.1), [x, y, vx, vy, vz])
    np.testing.assert_allclose(out[0][:2], [1.0 + 1.0 - 0.95, 17 + 17 - 17.05], atol=1e-5)
    np.testing.assert_allclose(out[2][:2], [1.0, -1.0])


def test_periodic_wrap_is_exact_fmod(orc):
    """The wrap x = fmod(x + ngx - ngb, ngx) + ngb is FP32; positions on the
    upper face are kept inside the last cell."""
    ngx, ngy = 64, 8
    g = orc.grid(ngx, ngy, 1, pbc=("periodic",) * 4)
    Ex = np.zeros((ngy + 1, ngx + 2), np.float32)
    Ey = np.zeros((ngy + 2, ngx + 1), np.float32)
    Bz = np.zeros((ngy + 2, ngx + 2), np.float32)
    rng = np.random.default_rng(0)
    x = (rng.random(5000) * ngx).astype(np.float32)
    vx = (rng.standard_normal(5000) * 3).astype(np.float32)
    y = np.full(5000, 4.0, np.float32)
    z = np.zeros(5000, np.float32)
    out = orc.push_fast(g, Ex, Ey, Bz, (0.0, 0.0, 1.0, 0.0), [x, y, vx, z, z])
    raw = (x + vx).astype(np.float32)
    exp = (np.fmod((raw + np.float32(ngx)).astype(np.float32), np.float32(ngx))).astype(np.float32)
    exp = np.where(exp >= ngx, np.nextafter(np.float32(ngx), np.float32(0)), exp)
    np.testing.assert_array_equal(out[0], exp)
    assert np.all((out[0] >= 0) & (out[0] < ngx))


@pytest.mark.parametrize("kind", ["fast", "ref"])
def test_boris_exb_drift(orc, kind):
    """Uniform Ex and Bz: a particle moving at the E x B drift
    v_d = -(constE / constB) Ex / Bz y (c = constE / constB in the push's
    constants) keeps that velocity step after step, and a gyrating particle's
    velocity averages to it over whole gyro-periods (the Boris scheme keeps
    the drift exactly; a wrong sign or factor in E or in the rotation moves it)."""
    ngx = ngy = 8
    g = orc.grid(ngx, ngy, 1, pbc=("periodic",) * 4)
    nx = ngx
    E0, B0, cE, cB = 1.0, 1.0, 0.1, 0.2
    Ex = np.full((nx + 1, nx + 2), E0, np.float32)
    Ey = np.zeros((nx + 2, nx + 1), np.float32)
    Bz = np.full((nx + 2, nx + 2), B0, np.float32)
    vd = -(cE / cB) * E0 / B0
    rng = np.random.default_rng(17)
    n = 400
    p = _particles(n, ngx, ngy, 0, seed=9)
    p[2] = np.zeros(n, np.float32)
    p[3] = np.full(n, vd, np.float32)
    q = _particles(n, ngx, ngy, 0, seed=10)
    ang = rng.uniform(0, 2 * np.pi, n)
    q[2] = (0.3 * np.cos(ang)).astype(np.float32)
    q[3] = (vd + 0.3 * np.sin(ang)).astype(np.float32)
    consts = (cE, cB, 0.01, 0.01)
    period = 2 * np.pi / (2 * math.atan(cB * B0))        # steps per gyration
    steps = int(round(19 * period))                       # close to 19 whole periods
    mean_vx, mean_vy = np.zeros(n), np.zeros(n)
    for _ in range(steps):
        if kind == "fast":
            p = list(orc.push_fast(g, Ex, Ey, Bz, consts, p, uniform_bz=True, bz0=B0))
            q = list(orc.push_fast(g, Ex, Ey, Bz, consts, q, uniform_bz=True, bz0=B0))
        else:
            p = list(orc.push_ref(g, Ex, Ey, Bz, consts, p)[:5])
            q = list(orc.push_ref(g, Ex, Ey, Bz, consts, q)[:5])
        mean_vx += q[2]
        mean_vy += q[3]
    np.testing.assert_allclose(p[2], 0.0, atol=1e-4)
    np.testing.assert_allclose(p[3], vd, atol=1e-4)
    assert abs(mean_vx.mean() / steps) < 5e-3
    assert abs(mean_vy.mean() / steps - vd) < 5e-3
