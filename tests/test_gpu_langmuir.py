"""End-to-end physics: a Langmuir oscillation through the device step.

Electrons loaded with a cosine density perturbation (mode 2 of a 128-cell
periodic box, dn/n = 0.05) over uniform, effectively immobile Xe+ ions
oscillate at the Bohm-Gross frequency w^2 = wpe^2 (1 + 3 k^2 lambda_D^2).
The field energy oscillates at twice that frequency; its spectral peak,
measured from 400 steps of gp_step with the FP64 diagnostics, has to match
2 w within 2 %.  The check pins the whole chain of units at once: the
deposit scale q w / (dx dy), the 4 pi of the Poisson right-hand side, the
gradient and the push constants q dt / (2 m) (a missing factor of two or a
wrong mass shows up as a factor of sqrt 2 or more).  The reference measures
the same frequency from its probe series (tools/fft.py)."""
import math
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ME, EC, KB, MXE = 9.10938188e-28, 4.803204197e-10, 1.3806503e-16, 2.1801714e-22


def test_langmuir_frequency(gp):
    print("[langmuir] 128x16, order 3, 64 ppc, 400 steps", flush=True)
    ngx, ngy, ppc, mode, steps = 128, 16, 64, 2, 400
    n0, te = 5e10, 1.0 * 11604.0
    wpe = math.sqrt(4 * math.pi * n0 * EC * EC / ME)
    lde = math.sqrt(KB * te / (4 * math.pi * n0 * EC * EC))
    dx, dt = lde, 0.1 / wpe
    w = n0 * dx * dx / ppc
    cfg = gp.make_config(ngx, ngy, dx, dx, order=3)
    ctx = gp.Context(cfg)
    n = ngx * ngy * ppc
    se = ctx.add_species(-EC, ME, w, n, vth_hint=math.sqrt(KB * te / ME))
    si = ctx.add_species(EC, MXE, w, n, vth_hint=1e4)
    L, H = ngx * float(np.float32(dx)), ngy * float(np.float32(dx))
    G = gp.Gen()
    G.gen_type = 7
    G.xmin, G.xmax, G.ymin, G.ymax = 0.0, L, 0.0, H
    G.ele_vth = math.sqrt(2 * KB * te / ME)
    G.pert_delta, G.pert_mode = 0.05, mode
    ctx.load_pair(se, -1, G, 1, 0, n)                 # electrons only, perturbed
    Gi = gp.Gen()
    Gi.gen_type = 0
    Gi.xmin, Gi.xmax, Gi.ymin, Gi.ymax = 0.0, L, 0.0, H
    Gi.ele_vth = math.sqrt(2 * KB * 0.01 * 11604.0 / MXE)
    ctx.load_pair(si, -1, Gi, 2, 0, n)                # the ions, uniform
    ctx.sort(se)
    ctx.sort(si)
    ctx.reset_grids()
    ctx.deposit(se)
    ctx.deposit(si)
    we = []
    for cyc in range(1, steps + 1):
        st, dg = ctx.step(dt, tol=1e-10, maxiter=400, cycle=cyc, diag=True)
        assert st.relres <= 1e-10
        we.append(dg.we_cgs)
    ctx.close()
    we = np.asarray(we)
    sig = (we - we.mean()) * np.hanning(steps)
    spec = np.abs(np.fft.rfft(sig, 64 * steps))
    f = np.fft.rfftfreq(64 * steps, d=dt)
    lo = np.searchsorted(f, 0.5 * wpe / math.pi)      # ignore the slow drift below 0.5 (2 wpe)
    k = lo + int(np.argmax(spec[lo:]))
    w_meas = math.pi * f[k]                           # field energy at 2 w: w = 2 pi f / 2
    kk = 2 * math.pi * mode / (ngx * dx)
    w_bg = wpe * math.sqrt(1 + 3 * (kk * lde) ** 2)
    print(f"[langmuir] w/wpe measured {w_meas / wpe:.4f}, Bohm-Gross {w_bg / wpe:.4f}", flush=True)
    sys.stdout.flush()
    assert abs(w_meas / w_bg - 1) < 0.02
