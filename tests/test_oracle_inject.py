"""The oracle's half-Maxwellian wall injector pinned by its distributions:
vx - u0 is Rayleigh with scale vth (CDF 1 - exp(-v^2 / 2 vth^2)), vy and vz
are Gaussian (mean u, variance vth^2) and uncorrelated, y is uniform on the
wall, and each particle sits within the distance it flew during a random
fraction of the step."""
import numpy as np

from oracle import pyoracle as orc


def test_injector_distributions():
    n, vth, u = 400000, 1.5, (0.25, -0.5, 0.75)
    dtdx, lo, ngx, ngy = 0.2, 2.0, 64.0, 48.0
    x, y, vx, vy, vz = orc.inject_wall(7, 1000, n, vth, u, dtdx, lo, ngx, ngy)
    v = np.sort((vx - u[0]).astype(np.float64))
    cdf = 1.0 - np.exp(-v * v / (2 * vth * vth))
    ks = np.max(np.abs(cdf - (np.arange(1, n + 1) / n)))
    assert ks < 1.63 / np.sqrt(n) * 1.5          # well inside the 1% KS bound
    for w, m in ((vy, u[1]), (vz, u[2])):
        assert abs(w.mean() - m) < 5 * vth / np.sqrt(n)
        assert abs(w.var() / vth ** 2 - 1) < 5 * np.sqrt(2.0 / n)
    assert abs(np.corrcoef(vy, vz)[0, 1]) < 5 / np.sqrt(n)
    assert y.min() >= lo and y.max() < lo + ngy
    assert abs((y - lo).mean() / ngy - 0.5) < 5 * np.sqrt(1 / 12 / n)
    assert np.all(x >= lo) and np.all(x - lo <= vx * dtdx * (1 + 1e-6))
    # x - lo = vx dt/dx r4 with r4 uniform: E[(x - lo) / (vx dt/dx)] = 1/2
    assert abs(((x - lo) / (vx * dtdx)).mean() - 0.5) < 5 * np.sqrt(1 / 12 / n)


def test_injector_streams_are_disjoint_and_reproducible():
    a = orc.inject_wall(3, 0, 1000, 1.0, (0, 0, 0), 0.1, 1.0, 16.0, 16.0)
    b = orc.inject_wall(3, 500, 500, 1.0, (0, 0, 0), 0.1, 1.0, 16.0, 16.0)
    for p, q in zip(a, b):
        np.testing.assert_array_equal(p[500:], q)      # id = stream + k
    c = orc.inject_wall(4, 0, 1000, 1.0, (0, 0, 0), 0.1, 1.0, 16.0, 16.0)
    assert not np.array_equal(a[1], c[1])


def test_injected_position_below_the_upper_face():
    """Stream 24,807,051 of seed 1 draws r5 = 1 - 2^-24 (its fifth PCG32 output
    is 2^32 - 210, whose float is 2^32 - 256).  On a 512-cell wall starting at
    grid position 1, lo + r5 * ngy = 513 - 2^-15 rounds to 513, the upper face;
    the injector keeps the particle in the last cell, as the loader does."""
    lo, ngy = 1.0, 512.0
    x, y, vx, vy, vz = orc.inject_wall(1, 24807051, 1, 1.0, (0, 0, 0), 0.1, lo, 512.0, ngy)
    assert np.float32(lo) + np.float32(1 - 2.0 ** -24) * np.float32(ngy) == np.float32(lo + ngy)
    assert y[0] == np.nextafter(np.float32(lo + ngy), np.float32(0))
    # a neighbouring stream is unaffected by the guard
    _, y2, _, _, _ = orc.inject_wall(1, 24807050, 2, 1.0, (0, 0, 0), 0.1, lo, 512.0, ngy)
    assert y2[1] == y[0] and y2[0] < lo + ngy
