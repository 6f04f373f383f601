"""Runs of the bounded-plasma and Darwin configurations through the binding.

bounded_plasma: a 2D slab between absorbing walls in x (Dirichlet phi = 0,
particles deleted, Particle.mm:215-269 code 1), periodic in y, uniform Bz.
Every cycle each species loses the particles that crossed a wall (gp_step's
reduce, Particle.mm:893-1002) and gains a half-Maxwellian wall flux through
x = 0: Gamma = n0 * vbar / 4 per unit wall length and time, vbar the mean
speed sqrt(8 kT / (pi m)), drawn by gp_inject_wall; the charge of the new
particles is deposited at once, as the reference's next-cycle deposit would
(main.mm:154-233).

darwin: the Darwin-model cycle (EMField = 2) on a wall-bounded, periodic-y
grid: Poisson, then the transverse solves for chi, Ax, Ay
(EMField.mm:1138-1361), with the current deposit; the three transverse
solves' iterations and residuals are recorded per cycle.

Both return per-cycle histories (counts, absorbed and injected particles,
energies, solver iterations, wall time) so that a caller can check the
particle balance and report the run.
"""
from __future__ import annotations

import math
import time

import numpy as np

from . import Context, Gen, make_config

ME = 9.10938188e-28      # g
EC = 4.803204197e-10     # statC
CL = 2.99792458e10       # cm/s
KB = 1.3806503e-16       # erg/K
EVTOK = 11604.0
MXE = 2.1801714e-22      # Xe+ (data/condition_large.json)


def plasma(n0=5e10, te_ev=10.0, ti_ev=1.0, ppc=128, wce_over_wpe=0.5):
    """dx = 0.5 Debye length, dt = 0.1 / omega_pe, weight n0 dx^2 / ppc,
    vth = sqrt(2 kT / m) (the loader's width, Collision.mm:264)"""
    te, ti = te_ev * EVTOK, ti_ev * EVTOK
    wpe = math.sqrt(4 * math.pi * n0 * EC * EC / ME)
    lde = math.sqrt(KB * te / (4 * math.pi * n0 * EC * EC))
    dx = 0.5 * lde
    return dict(n0=n0, te=te, ti=ti, dx=dx, dt=0.1 / wpe, w=n0 * dx * dx / ppc, wpe=wpe,
                bz=wce_over_wpe * wpe * ME * CL / EC,
                vth_e=math.sqrt(2 * KB * te / ME), vth_i=math.sqrt(2 * KB * ti / MXE))


def _load(ctx, se, si, P, ngx, ngy, n, seed=0):
    G = Gen()
    G.gen_type = 0
    G.xmin, G.xmax = 0.0, ngx * float(np.float32(P["dx"]))
    G.ymin, G.ymax = 0.0, ngy * float(np.float32(P["dx"]))
    G.ele_vth, G.ion_vth = P["vth_e"], P["vth_i"]
    ctx.load_pair(se, si, G, seed, 0, n)
    ctx.sort(se)
    ctx.sort(si)


def bounded_plasma(ngx=512, ngy=512, ppc=256, steps=500, order=3, tol=1e-8, sort_drift=3e-3, seed=0,
                   capacity_factor=1.25, log=None):
    P = plasma(ppc=ppc)
    n = ppc * ngx * ngy
    cfg = make_config(ngx, ngy, P["dx"], P["dx"], order=order, pbc=("Delete", "Delete", "periodic", "periodic"),
                      fbc=("Dirichlet", "Dirichlet", "periodic", "periodic"), b0=(0.0, 0.0, P["bz"]))
    ctx = Context(cfg)
    cap = int(n * capacity_factor)
    se = ctx.add_species(-EC, ME, P["w"], cap, vth_hint=P["vth_e"] / math.sqrt(2))
    si = ctx.add_species(EC, MXE, P["w"], cap, vth_hint=P["vth_i"] / math.sqrt(2))
    _load(ctx, se, si, P, ngx, ngy, n, seed)
    ctx.set_sort_drift(sort_drift)
    ctx.reset_grids()
    ctx.deposit(se)
    ctx.deposit(si)
    dt, ly = P["dt"], ngy * float(np.float32(P["dx"]))
    # wall flux per cycle in particles: n0 vbar / 4 * Ly * dt / w
    flux = {}
    for sid, (m, kt) in ((se, (ME, P["te"])), (si, (MXE, P["ti"]))):
        vbar = math.sqrt(8 * KB * kt / (math.pi * m))
        flux[sid] = P["n0"] * vbar / 4 * ly * dt / P["w"]
    vth1 = {se: math.sqrt(KB * P["te"] / ME), si: math.sqrt(KB * P["ti"] / MXE)}   # per-component width
    carry = {se: 0.0, si: 0.0}
    # particle k of a call draws from PCG stream (stream + k): advancing the
    # stream by each call's count keeps every injected particle's stream distinct
    next_stream = {se: 0, si: 0}
    hist = {k: [] for k in ("cycle", "count_e", "count_i", "absorbed_e", "absorbed_i", "injected_e", "injected_i",
                            "kinetic", "field", "iterations", "relres", "ms")}
    ctx.sync()
    t_all = time.perf_counter()
    for cyc in range(1, steps + 1):
        t0 = time.perf_counter()
        st, dg = ctx.step(dt, tol=tol, maxiter=400, sort_interval=0, cycle=cyc, diag=True)
        inj = {}
        for sid in (se, si):
            want = flux[sid] + carry[sid]
            k = int(want)
            carry[sid] = want - k
            if k > 0:
                ctx.inject_wall(sid, k, float(np.float32(vth1[sid])), (0.0, 0.0, 0.0), dt, seed + 1 + sid,
                                next_stream[sid], deposit=1)
                next_stream[sid] += k
            inj[sid] = k
        ctx.sync()
        hist["cycle"].append(cyc)
        hist["count_e"].append(ctx.count(se))
        hist["count_i"].append(ctx.count(si))
        hist["absorbed_e"].append(int(sum(dg.absorbed[se])))
        hist["absorbed_i"].append(int(sum(dg.absorbed[si])))
        hist["injected_e"].append(inj[se])
        hist["injected_i"].append(inj[si])
        hist["kinetic"].append(dg.kinetic[se] + dg.kinetic[si])
        hist["field"].append(dg.we_cgs)
        hist["iterations"].append(st.iterations)
        hist["relres"].append(st.relres)
        hist["ms"].append((time.perf_counter() - t0) * 1e3)
        if log and (cyc % 50 == 0 or cyc == steps):
            log(f"[bounded] cycle {cyc}: electrons {hist['count_e'][-1]}, ions {hist['count_i'][-1]}, "
                f"{hist['ms'][-1]:.2f} ms")
    seconds = time.perf_counter() - t_all
    out = dict(config=dict(ngx=ngx, ngy=ngy, ppc=ppc, order=order, steps=steps, n0=n, tol=tol,
                           flux_per_cycle={"electrons": flux[se], "ions": flux[si]}),
               history=hist, seconds=seconds, sorts=[ctx.sort_stats(se)[0], ctx.sort_stats(si)[0]])
    ctx.close()
    return out


def darwin(ngx=512, ngy=512, ppc=128, steps=500, order=3, tol=1e-8, sort_drift=3e-3, seed=0, log=None):
    P = plasma(ppc=ppc)
    n = ppc * ngx * ngy
    cfg = make_config(ngx, ngy, P["dx"], P["dx"], order=order, pbc=("Reflect", "Reflect", "periodic", "periodic"),
                      fbc=("Dirichlet", "Dirichlet", "periodic", "periodic"), b0=(0.0, 0.0, P["bz"]),
                      cur_dens=1, em_model=2)
    ctx = Context(cfg)
    se = ctx.add_species(-EC, ME, P["w"], n, vth_hint=P["vth_e"] / math.sqrt(2))
    si = ctx.add_species(EC, MXE, P["w"], n, vth_hint=P["vth_i"] / math.sqrt(2))
    _load(ctx, se, si, P, ngx, ngy, n, seed)
    ctx.set_sort_drift(sort_drift)
    ctx.reset_grids()
    ctx.deposit(se)
    ctx.deposit(si)
    dt = P["dt"]
    hist = {k: [] for k in ("cycle", "iterations", "relres", "darwin_iterations", "darwin_relres", "kinetic", "field",
                            "ms")}
    ctx.sync()
    t_all = time.perf_counter()
    for cyc in range(1, steps + 1):
        t0 = time.perf_counter()
        st, dg = ctx.step(dt, tol=tol, maxiter=400, sort_interval=0, cycle=cyc, diag=(cyc % 10 == 0))
        ctx.sync()
        hist["cycle"].append(cyc)
        hist["iterations"].append(st.iterations)
        hist["relres"].append(st.relres)
        d3 = ctx.darwin_stats()
        hist["darwin_iterations"].append([d.iterations for d in d3])
        hist["darwin_relres"].append(max(d.relres for d in d3))
        hist["kinetic"].append(dg.kinetic[se] + dg.kinetic[si] if dg is not None else float("nan"))
        hist["field"].append(dg.we_cgs if dg is not None else float("nan"))
        hist["ms"].append((time.perf_counter() - t0) * 1e3)
        if log and (cyc % 50 == 0 or cyc == steps):
            log(f"[darwin] cycle {cyc}: {hist['ms'][-1]:.2f} ms")
    seconds = time.perf_counter() - t_all
    out = dict(config=dict(ngx=ngx, ngy=ngy, ppc=ppc, order=order, steps=steps, n0=n, tol=tol),
               history=hist, seconds=seconds)
    ctx.close()
    return out


def summary(out):
    """the run's figures for a report: sizes, wall time, per-cycle time,
    solver iterations, and for the bounded plasma the particle balance"""
    h = out["history"]
    ms = np.array(h["ms"])
    s = dict(config=out["config"], seconds=out["seconds"], ms_per_cycle_median=float(np.median(ms)),
             ms_per_cycle_mean=float(ms.mean()), iterations_mean=float(np.mean(h["iterations"])),
             iterations_max=int(np.max(h["iterations"])), relres_max=float(np.max(h["relres"])))
    if "darwin_iterations" in h:
        d = np.array(h["darwin_iterations"])
        s["darwin_iterations_mean"] = d.mean(axis=0).tolist()      # chi, Ax, Ay
        s["darwin_iterations_max"] = d.max(axis=0).tolist()
        s["darwin_relres_max"] = float(np.max(h["darwin_relres"]))
    if "count_e" in h:
        n0 = out["config"]["n0"]
        ok = True
        for sp in ("e", "i"):
            prev = n0
            for c, a, j in zip(h["count_" + sp], h["absorbed_" + sp], h["injected_" + sp]):
                ok &= c == prev - a + j
                prev = c
            s["final_count_" + sp] = h["count_" + sp][-1]
            s["absorbed_total_" + sp] = int(np.sum(h["absorbed_" + sp]))
            s["injected_total_" + sp] = int(np.sum(h["injected_" + sp]))
        s["balance_exact"] = bool(ok)
        s["sorts"] = out["sorts"]
        s["particle_steps_per_s"] = float(np.sum(h["count_e"]) + np.sum(h["count_i"])) / (ms.sum() * 1e-3)
    k = [v for v in h["kinetic"] if not math.isnan(v)]
    f = [v for v in h["field"] if not math.isnan(v)]
    if k:
        s["kinetic_first_last"] = [k[0], k[-1]]
        s["field_first_last"] = [f[0], f[-1]]
    return s


def main(argv=None):
    import argparse
    import json
    ap = argparse.ArgumentParser(description="bounded-plasma or Darwin run on one GPU")
    ap.add_argument("which", choices=["bounded", "darwin"])
    ap.add_argument("--ngx", type=int, default=512)
    ap.add_argument("--ppc", type=int, default=None)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--order", type=int, default=3)
    ap.add_argument("--history", default=None, help="write the per-cycle history here as JSON")
    a = ap.parse_args(argv)
    log = lambda m: print(m, flush=True)
    if a.which == "bounded":
        out = bounded_plasma(ngx=a.ngx, ngy=a.ngx, ppc=a.ppc or 256, steps=a.steps, order=a.order, log=log)
    else:
        out = darwin(ngx=a.ngx, ngy=a.ngx, ppc=a.ppc or 128, steps=a.steps, order=a.order, log=log)
    if a.history:
        with open(a.history, "w") as f:
            json.dump(out, f)
    print(json.dumps(summary(out)), flush=True)


if __name__ == "__main__":
    main()
