"""Pins of the oracle's deposit (src/Particle.mm:281-389) and moments
(src/Moment.mm:33-135): single-particle outer products, total charge,
agreement of the int64 fixed-point scheme with the reference's sequential
FP32 sums and with FP64 accumulation, and the exact float->int64 rounding."""
from fractions import Fraction

import numpy as np
import pytest


def _parts(n, ngx, ngy, ngb, seed=0):
    rng = np.random.default_rng(seed)
    x = (ngb + rng.random(n) * ngx).astype(np.float32)
    y = (ngb + rng.random(n) * ngy).astype(np.float32)
    v = rng.standard_normal((3, n)).astype(np.float32)
    return [x, y, v[0], v[1], v[2]]


@pytest.mark.parametrize("order", [1, 3, 5])
def test_single_particle_outer_product(orc, order):
    ngx = ngy = 8
    ngb = order // 2
    g = orc.grid(ngx, ngy, order)
    x, y = np.float32(4.3), np.float32(5.7)
    p = [np.array([x]), np.array([y]), np.zeros(1, np.float32), np.zeros(1, np.float32), np.zeros(1, np.float32)]
    rho, _, _ = orc.deposit_fx(g, p)
    wx = orc.weights(order, float(x - np.floor(x)), "fast")
    wy = orc.weights(order, float(y - np.floor(y)), "fast")
    # order 3: 16-bit weights, products scaled by 2^14; orders 1, 5: 23 bits
    bits, sh = (16, 14) if order == 3 else (23, 0)
    qx = np.rint(wx * np.float32(2 ** bits)).astype(np.int64)
    qy = np.rint(wy * np.float32(2 ** bits)).astype(np.int64)
    i2, j1 = int(np.floor(x)), int(np.floor(y))
    exp = np.zeros_like(rho)
    exp[j1 - ngb:j1 - ngb + order + 1, i2 - ngb:i2 - ngb + order + 1] = np.outer(qy, qx) << sh
    np.testing.assert_array_equal(rho, exp)
    # FP64 form of the reference weights places the same stencil; a product
    # of two weights each within 2^-(bits+1) is within 2^-bits
    r64, _, _ = orc.deposit_f64(g, p, 1.0)
    np.testing.assert_allclose(rho * 2.0 ** -46, r64, atol=2.0 ** -bits + 3e-7)


@pytest.mark.parametrize("order", [1, 3, 5])
def test_fx_matches_reference_sums(orc, order):
    """int64 scheme vs reference FP32 sequential sums vs FP64 sums: normwise
    relative difference <= 1e-6 for rho and J of one species."""
    ngx, ngy = 24, 20
    ngb = order // 2
    g = orc.grid(ngx, ngy, order)
    p = _parts(60000, ngx, ngy, ngb, seed=order)
    scale = np.float32(0.37)
    rho_fx, jx_fx, jy_fx = orc.deposit_fx(g, p, inv_vnorm=1.0 / 8.0, current=True)
    rho_ref, jx_ref, jy_ref = orc.deposit_ref(g, p, scale, current=True)
    rho64, jx64, jy64 = orc.deposit_f64(g, p, float(scale), current=True)
    rho_b = rho_fx * (2.0 ** -46 * float(scale))
    jx_b = jx_fx * (2.0 ** -40 * 8.0 * float(scale))
    jy_b = jy_fx * (2.0 ** -40 * 8.0 * float(scale))
    for b, r, e in ((rho_b, rho_ref, rho64), (jx_b, jx_ref, jx64), (jy_b, jy_ref, jy64)):
        nrm = np.linalg.norm(e)
        assert np.linalg.norm(b - e) / nrm < 1e-6
        assert np.linalg.norm(r - e) / nrm < 1e-5
        assert np.linalg.norm(b - r) / nrm < 1e-5
    # total charge: sum over nodes = N * scale (partition of unity)
    assert abs(rho_b.sum() / (len(p[0]) * float(scale)) - 1.0) < 1e-6


def test_fx_deterministic_across_threads(orc):
    g = orc.grid(16, 16, 3)
    p = _parts(30000, 16, 16, 1, seed=9)
    a = orc.deposit_fx(g, p, 0.5, True, threads=1)
    b = orc.deposit_fx(g, p, 0.5, True, threads=7)
    for u, v in zip(a, b):
        np.testing.assert_array_equal(u, v)


def test_fx_from_float_exact(orc):
    rng = np.random.default_rng(4)
    vals = np.concatenate([rng.standard_normal(2000) * 10.0 ** rng.integers(-9, 5, 2000),
                           [0.0, 0.5, -0.5, 2.0 ** -41, 3 * 2.0 ** -41, -(2.0 ** -41)]]).astype(np.float32)
    for v in vals:
        exact = Fraction(float(v)) * 2 ** 40
        fl = exact.numerator // exact.denominator
        rem = exact - fl
        r = fl + 1 if (rem > Fraction(1, 2) or (rem == Fraction(1, 2) and fl % 2 == 1)) else fl
        assert orc.lib().gpo_fx_from_float(float(v), 40) == r


def test_moments_density(orc):
    """n moment integrates to N*w; n*v moments integrate to sum(v)."""
    g = orc.grid(16, 12, 3)
    p = _parts(20000, 16, 12, 1, seed=11)
    m = orc.moments_f64(g, p, 1.0)
    assert abs(m[0].sum() - 20000) < 1e-6
    np.testing.assert_allclose(m[1].sum(), p[2].astype(np.float64).sum(), rtol=1e-9, atol=1e-7)
    np.testing.assert_allclose(m[9].sum(), (p[4].astype(np.float64) ** 2).sum(), rtol=1e-9)
