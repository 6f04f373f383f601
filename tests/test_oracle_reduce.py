"""Pins of the oracle's absorption compaction and kinetic energy
(src/Particle.mm:893-1002).

The device compaction fills the holes below n_new in ascending order with
the live particles of the tail [n_new, n) taken from the end.  That closed
form is checked here against the literal restatement of the reference loop.
The loop over-counts by one when the array ends in a run of absorbed
particles that the tail scan has already passed (flags [.., 1, 1]); that
case is pinned separately as a reference quirk."""
import numpy as np
import pytest


def hole_donor(parts, flags):
    n = len(flags)
    dead = np.nonzero(flags > 0)[0]
    n_new = n - len(dead)
    holes = dead[dead < n_new]
    tail = np.arange(n_new, n)
    donors = tail[flags[n_new:] == 0]
    out = [a.copy() for a in parts]
    for k, h in enumerate(holes):
        d = donors[len(holes) - 1 - k]
        for a, src in zip(out, parts):
            a[h] = src[d]
    counts = [int(np.sum(flags == w)) for w in (1, 2, 3, 4)]
    return [a[:n_new] for a in out], counts, n_new


def _quirk(flags):
    """True when the reference loop ends inside a dead tail run (see module doc)."""
    n = len(flags)
    kmax, k1 = n, 0
    while k1 < n:
        if k1 >= kmax:
            return flags[k1] > 0
        if flags[k1] > 0:
            k2 = kmax - 1
            while k2 >= k1:
                kmax -= 1
                if flags[k2] == 0:
                    break
                k2 -= 1
        k1 += 1
    return False


@pytest.mark.parametrize("seed", range(40))
def test_literal_loop_equals_hole_donor(orc, seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(1, 400))
    frac = rng.choice([0.0, 0.01, 0.1, 0.5, 0.9])
    flags = np.where(rng.random(n) < frac, rng.integers(1, 5, n), 0).astype(np.int32)
    parts = [rng.standard_normal(n).astype(np.float32) for _ in range(5)]
    lit, lcnt, ln = orc.reduce_literal(parts, flags)
    mod, mcnt, mn = hole_donor(parts, flags)
    if _quirk(flags):
        assert ln == mn - 1
        return
    assert ln == mn and lcnt == mcnt
    for a, b in zip(lit, mod):
        np.testing.assert_array_equal(a, b)


def test_reference_quirk_case(orc):
    flags = np.array([0, 0, 1, 1], np.int32)
    parts = [np.arange(4, dtype=np.float32) for _ in range(5)]
    _, cnt, n = orc.reduce_literal(parts, flags)
    assert n == 1          # the reference loop drops a live particle here
    assert hole_donor(parts, flags)[2] == 2


def test_kinetic_energy(orc):
    rng = np.random.default_rng(0)
    n = 10000
    parts = [np.abs(rng.random(n)).astype(np.float32) + 1] + [rng.standard_normal(n).astype(np.float32) * 1e8
                                                               for _ in range(4)]
    parts[0][::7] = -1.0     # absorbed markers are excluded
    m, w = np.float32(9.10938188e-28), np.float32(1234.5)
    vv = (parts[2] * parts[2] + parts[3] * parts[3] + parts[4] * parts[4]).astype(np.float32)
    exp = np.sum((0.5 * np.float64(m) * vv.astype(np.float64) * np.float64(w))[parts[0] >= 0])
    np.testing.assert_allclose(orc.kinetic(parts, m, w), exp, rtol=1e-12)
