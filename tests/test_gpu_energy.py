"""Energy-drift history over 1000 steps: the device step loop against the
CPU oracle loop on identical seeded inputs (the plain electrostatic case:
64x64 periodic, dx = 0.5 Debye length, dt = 0.1/omega_pe, CIC, electrons and
Xe+ at 100 particles per cell each, Te/Ti = 10).

Both loops record the FP64 total energy per step, kinetic energy after the
push plus the field energy E^2/(8 pi) of the solve.  The trajectories
separate once the two Poisson solutions (MG-CG and the oracle's CG, both to
a 1e-12 relative residual) first differ in an FP32 field bit, so the check
is on the histories: equal to FP64 summation-order level for the first steps
(the device sums the kinetic energy per workgroup, then in a fixed order; the
oracle sums sequentially), and the same drift (relative to the initial
total) over the run.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ME, EC, CL, KB = 9.10938188e-28, 4.803204197e-10, 2.99792458e10, 1.3806503e-16
MXE = 2.1801714e-22


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def plasma(n0=5e10, te_ev=10.0, ti_ev=1.0, ppc=100):
    te, ti = te_ev * 11604.0, ti_ev * 11604.0
    wpe = np.sqrt(4 * np.pi * n0 * EC * EC / ME)
    lde = np.sqrt(KB * te / (4 * np.pi * n0 * EC * EC))
    dx = 0.5 * lde
    return dict(dx=dx, dt=0.1 / wpe, w=n0 * dx * dx / ppc, vth_e=np.sqrt(2 * KB * te / ME),
                vth_i=np.sqrt(2 * KB * ti / MXE))


def consts(q, m, dt, dx):
    cE = np.float32(0.5 * float(np.float32(q)) * dt / float(np.float32(m)))
    cB = np.float32(0.5 * float(np.float32(q)) * dt / float(np.float32(m)) / CL)
    cX = np.float32(dt / float(np.float32(dx)))
    return float(cE), float(cB), float(cX), float(cX)


@pytest.mark.slow
def test_energy_drift_history_matches_oracle(gp, orc):
    nsteps, ngx, ppc = 1000, 64, 100
    P = plasma(ppc=ppc)
    dx, dt, w = P["dx"], P["dt"], P["w"]
    n = ppc * ngx * ngx
    say(f"[energy] {nsteps} steps, 64x64, 2 x {n} particles, dt {dt:.3e} s")
    cfg = gp.make_config(ngx, ngx, dx, dx, order=1)
    ctx = gp.Context(cfg)
    se = ctx.add_species(-EC, ME, w, n, vth_hint=P["vth_e"] / np.sqrt(2))
    si = ctx.add_species(EC, MXE, w, n, vth_hint=P["vth_i"] / np.sqrt(2))
    G = gp.Gen()
    G.gen_type = 0
    L = ngx * float(np.float32(dx))
    G.xmin, G.xmax, G.ymin, G.ymax = 0.0, L, 0.0, L
    G.ele_vth, G.ion_vth = P["vth_e"], P["vth_i"]
    ctx.load_pair(se, si, G, 0, 0, n)
    ctx.sort(se); ctx.sort(si)
    pe, pi = ctx.download(se), ctx.download(si)
    ctx.reset_grids(); ctx.deposit(se); ctx.deposit(si)
    eg = np.zeros(nsteps)
    for cyc in range(1, nsteps + 1):
        st, dg = ctx.step(dt, tol=1e-12, maxiter=500, sort_interval=20, cycle=cyc, diag=True)
        eg[cyc - 1] = dg.kinetic[0] + dg.kinetic[1] + dg.we_cgs
    ctx.close()
    say(f"   device loop done: E0 {eg[0]:.9e} E_end {eg[-1]:.9e}")
    # oracle loop
    g = orc.grid(ngx, ngx, 1)
    f = orc.field(ngx, ngx, 0, float(np.float32(dx)), float(np.float32(dx)))
    ptr, col, val, rbc = orc.poisson_csr(f)
    Bz = np.zeros((ngx + 2, ngx + 2), np.float32)
    ce, ci = consts(-EC, ME, dt, dx), consts(EC, MXE, dt, dx)
    se_, si_ = [float(np.float32(float(np.float32(q)) * float(np.float32(w))) / (float(np.float32(dx)) ** 2))
                for q in (-EC, EC)]
    se_, si_ = float(np.float32(se_)), float(np.float32(si_))
    ec = np.zeros(nsteps)
    sol = None
    for k in range(nsteps):
        rho = np.zeros((ngx + 1, ngx + 1))
        for parts, sc in ((pe, se_), (pi, si_)):
            acc, _, _ = orc.deposit_fx(g, parts, threads=16)
            rho += acc * (2.0 ** -46 * sc)
        rho = orc.fold(f, rho)
        rhs = orc.poisson_rhs(f, rho, rbc)
        rhs -= rhs.mean()
        sol, _, _ = orc.cg(ptr, col, val, rhs, x0=sol, tol=1e-12, project=True)
        phi = orc.reconstruct(f, sol)
        Ex, Ey = orc.gradient(f, phi)
        _, wcgs, _ = orc.field_energy(f, Ex, Ey)
        pe = orc.push_fast(g, Ex, Ey, Bz, ce, pe, uniform_bz=True, threads=16)
        pi = orc.push_fast(g, Ex, Ey, Bz, ci, pi, uniform_bz=True, threads=16)
        ec[k] = orc.kinetic(pe, float(np.float32(ME)), float(np.float32(w))) + \
            orc.kinetic(pi, float(np.float32(MXE)), float(np.float32(w))) + wcgs
    dg_, dc_ = (eg - eg[0]) / eg[0], (ec - ec[0]) / ec[0]
    rel = np.abs(eg - ec) / np.abs(ec)
    say(f"   relative difference of the totals: first 10 steps max {rel[:10].max():.2e}, run max {rel.max():.2e}")
    for k in (9, 99, 249, 499, 749, 999):
        say(f"   step {k + 1:4d}: drift device {dg_[k]:+.4e} oracle {dc_[k]:+.4e}")
    np.save("gpurun_out/energy_history.npy", np.stack([eg, ec])) if __import__("os").path.isdir("gpurun_out") else None
    rel_e0 = abs(eg[0] - ec[0]) / abs(ec[0])
    say(f"   initial total energy relative difference {rel_e0:.2e}")
    assert rel_e0 < 1e-12
    # the first steps: same particles to the bit, FP64 sums in another order
    assert rel[:10].max() < 1e-10
    # same drift over the run: the difference of the drift curves stays small
    # against the drift itself (or the FP32 level where there is none)
    tolr = 0.1 * np.max(np.abs(dc_)) + 1e-5
    say(f"   max |drift difference| {np.max(np.abs(dg_ - dc_)):.3e} (tolerance {tolr:.3e})")
    assert np.max(np.abs(dg_ - dc_)) <= tolr
