"""Whole runs of the bounded-plasma and Darwin configurations
(gample_amd/scenarios.py) at small sizes.

Bounded plasma: every cycle the particle count moves by exactly the
injected minus the absorbed particles (Particle.mm:893-1002 reduce, the
wall source appended after it); the first cycle's absorption through each
wall matches the one-sided Maxwellian flux n vbar / 4 Ly dt / w within
five standard deviations of its Poisson count; every solve reaches the
tolerance.  Darwin: every cycle's four solves (Poisson, chi, Ax, Ay)
converge and the energies stay finite."""
import math
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_bounded_plasma_particle_balance(gp):
    from gample_amd import scenarios
    say("[scenario] bounded plasma 64x64, 32 ppc, 40 cycles")
    ngx = 64
    ppc = 32
    out = scenarios.bounded_plasma(ngx=ngx, ngy=ngx, ppc=ppc, steps=40, order=3, capacity_factor=1.5)
    h = out["history"]
    n0 = ppc * ngx * ngx
    for sp in ("e", "i"):
        prev = n0
        for c, a, j in zip(h["count_" + sp], h["absorbed_" + sp], h["injected_" + sp]):
            assert c == prev - a + j
            prev = c
    flux = out["config"]["flux_per_cycle"]["electrons"]
    assert flux == pytest.approx(ppc * ngx * math.sqrt(8 / math.pi) * 0.2 / 4, rel=1e-6)
    # both walls absorb at cycle 1: expected 2 * flux
    assert abs(h["absorbed_e"][0] - 2 * flux) < 5 * math.sqrt(2 * flux)
    assert sum(h["injected_e"]) == pytest.approx(40 * flux, abs=1.0)
    assert max(h["relres"]) <= 1e-8
    assert all(np.isfinite(h["kinetic"])) and all(np.isfinite(h["field"]))
    # the electrons leave faster than the ions: the slab charges positive
    assert h["count_e"][-1] < h["count_i"][-1]


def test_darwin_run(gp):
    from gample_amd import scenarios
    say("[scenario] darwin 64x64, 16 ppc, 20 cycles")
    out = scenarios.darwin(ngx=64, ngy=64, ppc=16, steps=20, order=3)
    h = out["history"]
    assert max(h["relres"]) <= 1e-8 and max(h["darwin_relres"]) <= 1e-8
    # the loader places each electron on its ion: cycle 1 solves rho = 0
    assert h["iterations"][0] == 0 and all(it > 0 for it in h["iterations"][1:])
    assert all(0 < it < 100 for d in h["darwin_iterations"] for it in d)
    k = [v for v in h["kinetic"] if not math.isnan(v)]
    assert len(k) == 2 and all(np.isfinite(k))
    assert all(np.isfinite([v for v in h["field"] if not math.isnan(v)]))
