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

learned: the learned Poisson preconditioner trained at the workload's scale
on thousands of the run's own time-loop increments, kept on the device, and
compared with MG-CG on held-out steps of the same run.  darwin_learned: the
network trained on a Darwin run's Poisson increments, fine-tuned on its
Darwin ones and used in all four solves, against the same run with MG-CG.
"""
from __future__ import annotations

import math
import time

import numpy as np

from . import F_RHS, F_SOL, Context, Gen, make_config

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


class WallSource:
    """Half-Maxwellian flux through the x = 0 wall, n0 vbar / 4 per unit wall
    length and time (vbar = sqrt(8 kT / (pi m))), in whole particles per cycle
    with the fraction carried over; particle k of a call draws from PCG stream
    (stream + k), and each call advances the stream by its count, so every
    injected particle's stream is distinct."""

    def __init__(self, ctx, se, si, P, ngy, seed):
        self.ctx, self.dt, self.seed = ctx, P["dt"], seed
        ly = ngy * float(np.float32(P["dx"]))
        self.flux = {}
        for sid, (m, kt) in ((se, (ME, P["te"])), (si, (MXE, P["ti"]))):
            vbar = math.sqrt(8 * KB * kt / (math.pi * m))
            self.flux[sid] = P["n0"] * vbar / 4 * ly * P["dt"] / P["w"]
        self.vth1 = {se: math.sqrt(KB * P["te"] / ME), si: math.sqrt(KB * P["ti"] / MXE)}   # per-component width
        self.carry = {se: 0.0, si: 0.0}
        self.next_stream = {se: 0, si: 0}

    def inject(self):
        """one cycle's injection of every species (charge deposited at once);
        the particles injected per species"""
        inj = {}
        for sid in self.flux:
            want = self.flux[sid] + self.carry[sid]
            k = int(want)
            self.carry[sid] = want - k
            if k > 0:
                self.ctx.inject_wall(sid, k, float(np.float32(self.vth1[sid])), (0.0, 0.0, 0.0), self.dt,
                                     self.seed + 1 + sid, self.next_stream[sid], deposit=1)
                self.next_stream[sid] += k
            inj[sid] = k
        return inj


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
    dt = P["dt"]
    src = WallSource(ctx, se, si, P, ngy, seed)
    flux = src.flux
    hist = {k: [] for k in ("cycle", "count_e", "count_i", "absorbed_e", "absorbed_i", "injected_e", "injected_i",
                            "kinetic", "field", "iterations", "relres", "ms")}
    ctx.sync()
    t_all = time.perf_counter()
    for cyc in range(1, steps + 1):
        t0 = time.perf_counter()
        st, dg = ctx.step(dt, tol=tol, maxiter=400, sort_interval=0, cycle=cyc, diag=True)
        inj = src.inject()
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


def learned(kind="periodic", ngx=1024, ppc=None, pairs=10000, eval_fields=8, warmup=20, train_seconds=120.0,
            max_epochs=100, batch=4, lr=1e-3, tol=1e-8, order=3, seed=0, save=None, log=None):
    """The learned Poisson preconditioner at the workload's scale: the network
    is trained on `pairs` time-loop increments (b_n - b_{n-1}, x_n - x_{n-1})
    that the run itself produces, kept on the device (gp_nn_add_increment),
    then compared with MG-CG on held-out steps of the same run.

    kind "periodic": the metric workload (ngx^2 periodic, uniform Bz, 128 ppc);
    "bounded": the bounded plasma (walls in x with the wall source, 256 ppc).
    Held-out step n is solved from x_{n-1} by MG-CG (mg_warm), by MG-CG after
    the network's S-optimally scaled correction (nn_warm_mg), by the network
    flexible CG (nn_warm_fcg), and from zero by MG-CG (mg_cold).  On the
    held-out increments: the network's relative energy-norm error and the
    residual it leaves, against the warm start's."""
    periodic = kind == "periodic"
    ppc = ppc or (128 if periodic else 256)
    P = plasma(ppc=ppc)
    n = ppc * ngx * ngx
    if periodic:
        cfg = make_config(ngx, ngx, P["dx"], P["dx"], order=order, b0=(0.0, 0.0, P["bz"]))
    else:
        cfg = make_config(ngx, ngx, P["dx"], P["dx"], order=order, pbc=("Delete", "Delete", "periodic", "periodic"),
                          fbc=("Dirichlet", "Dirichlet", "periodic", "periodic"), b0=(0.0, 0.0, P["bz"]))
    ctx = Context(cfg)
    cap = n if periodic else int(n * 1.25)
    se = ctx.add_species(-EC, ME, P["w"], cap, vth_hint=P["vth_e"] / math.sqrt(2))
    si = ctx.add_species(EC, MXE, P["w"], cap, vth_hint=P["vth_i"] / math.sqrt(2))
    _load(ctx, se, si, P, ngx, ngx, n, seed)
    ctx.set_sort_drift(3e-3)
    ctx.reset_grids(); ctx.deposit(se); ctx.deposit(si)
    src = None if periodic else WallSource(ctx, se, si, P, ngx, seed)
    dt = P["dt"]
    cyc = 0

    def step():
        nonlocal cyc
        cyc += 1
        ctx.step(dt, tol=tol, maxiter=400, sort_interval=0, cycle=cyc)
        if src is not None:
            src.inject()

    for _ in range(warmup):
        step()
    ctx.nn_init(32, seed=1)
    ctx.nn_config(1, 2, 0)
    t0 = time.perf_counter()
    got = 0
    while got < pairs:
        step()
        got += 1 if ctx.nn_add_increment() else 0
        if log and got and got % 2000 == 0:
            log(f"[learned] {got} increments ({time.perf_counter() - t0:.0f} s)")
    my, mx = ctx.field_shape(F_SOL)[0]
    bs, xs = [], []
    for _ in range(eval_fields + 1):                 # held out: not in the training set
        step()
        bs.append(ctx.field(F_RHS).ravel().copy())
        xs.append(ctx.field(F_SOL).ravel().copy())
    t1 = time.perf_counter()
    loss0 = ctx.nn_train(epochs=1, batch=batch, lr=0.0)
    losses, epochs = [loss0], 0
    while epochs < max_epochs and time.perf_counter() - t1 < train_seconds:
        losses.append(ctx.nn_train(epochs=1, batch=batch, lr=lr))
        epochs += 1
        if log:
            log(f"[learned] epoch {epochs}: loss {losses[-1]:.4f}")
    t2 = time.perf_counter()
    h = float(np.float32(P["dx"]))

    def S(v):      # the solver's operator on the unknown grid (periodic y; periodic or Dirichlet x)
        u = v.reshape(my, mx)
        if periodic:
            ex = np.roll(u, 1, 1) + np.roll(u, -1, 1)
        else:
            ex = np.zeros_like(u)
            ex[:, 1:] += u[:, :-1]
            ex[:, :-1] += u[:, 1:]
        return ((4 * u - np.roll(u, 1, 0) - np.roll(u, -1, 0) - ex) / h ** 2).ravel()

    rows = {k: [] for k in ("mg_cold", "mg_warm", "nn_warm_mg", "nn_warm_fcg")}
    err, res_warm, res_nn = [], [], []
    for k in range(1, eval_fields + 1):
        b, xp = bs[k], xs[k - 1]
        db, dxk = bs[k] - bs[k - 1], xs[k] - xs[k - 1]
        for key, solver, x0 in (("mg_cold", 0, None), ("mg_warm", 0, xp), ("nn_warm_mg", 2, xp),
                                ("nn_warm_fcg", 1, xp)):
            _, st, _ = ctx.solve_rhs(b, x0=x0, tol=tol, maxiter=400, solver=solver)
            rows[key].append((st.iterations, 1e3 * st.seconds, st.relres, st.fell_back))
        z = ctx.nn_infer(db)
        sz = S(z)
        a = (z @ db) / (z @ sz)                      # S dx = db: the S-optimal multiple of z
        e = dxk - a * z
        err.append(float(np.sqrt((e @ S(e)) / (dxk @ S(dxk)))))
        ar = (db @ sz) / (sz @ sz)
        nb = np.linalg.norm(b)
        res_warm.append(float(np.linalg.norm(db) / nb))
        res_nn.append(float(np.linalg.norm(db - ar * sz) / nb))
    out = dict(config=dict(kind=kind, ngx=ngx, ppc=ppc, particles=2 * n, order=order, tol=tol, warmup=warmup),
               train_pairs=pairs, pairs="time-loop increments (b_n - b_{n-1}, x_n - x_{n-1}), device resident",
               harvest_s=t1 - t0, train_s=t2 - t1, epochs=epochs, batch=batch, lr=lr,
               loss_untrained=loss0, loss_trained=losses[-1], loss_curve=losses,
               heldout_fields=eval_fields, heldout_energy_error=float(np.mean(err)),
               heldout_relres_warm_start=float(np.mean(res_warm)),
               heldout_relres_after_network=float(np.mean(res_nn)))
    for key, v in rows.items():
        a = np.array(v, dtype=float)
        out[key] = {"iterations": float(a[:, 0].mean()), "ms": float(a[:, 1].mean()),
                    "max_relres": float(a[:, 2].max()), "fell_back": int(a[:, 3].sum())}
    out["nn_warm_mg_over_mg_warm_time"] = out["nn_warm_mg"]["ms"] / out["mg_warm"]["ms"]
    out["nn_warm_fcg_over_mg_cold_time"] = out["nn_warm_fcg"]["ms"] / out["mg_cold"]["ms"]
    if save:
        ctx.nn_save(save)
    ctx.close()
    return out


def darwin_learned(ngx=512, ppc=128, warmup=20, poisson_pairs=1000, darwin_steps=1000, eval_steps=100,
                   train_seconds=60.0, solver=2, tol=1e-8, order=3, seed=0, log=None):
    """The Darwin model (the darwin run's set-up) with the learned
    preconditioner: the network is trained on the run's own Poisson increments,
    then fine-tuned on its Darwin increments (the chi, Ax and Ay systems,
    EMField.mm:1147-1275, each b_n - b_{n-1} -> x_n - x_{n-1}), and serves as
    the warm start (solver 2) or flexible-CG preconditioner (solver 1) of all
    four solves for eval_steps cycles.  The same seeded run with plain MG-CG
    gives the reference figures for the same cycles."""

    def make():
        P = plasma(ppc=ppc)
        n = ppc * ngx * ngx
        cfg = make_config(ngx, ngx, P["dx"], P["dx"], order=order, pbc=("Reflect", "Reflect", "periodic", "periodic"),
                          fbc=("Dirichlet", "Dirichlet", "periodic", "periodic"), b0=(0.0, 0.0, P["bz"]),
                          cur_dens=1, em_model=2)
        ctx = Context(cfg)
        se = ctx.add_species(-EC, ME, P["w"], n, vth_hint=P["vth_e"] / math.sqrt(2))
        si = ctx.add_species(EC, MXE, P["w"], n, vth_hint=P["vth_i"] / math.sqrt(2))
        _load(ctx, se, si, P, ngx, ngx, n, seed)
        ctx.set_sort_drift(3e-3)
        ctx.reset_grids(); ctx.deposit(se); ctx.deposit(si)
        return ctx, P["dt"]

    total = warmup + (poisson_pairs + 1) + (darwin_steps + 1) + eval_steps

    def record(ctx, st, rows, t):
        d3 = ctx.darwin_stats()
        rows.append([st.iterations, 1e3 * st.seconds] + [d.iterations for d in d3] +
                    [1e3 * d.seconds for d in d3] + [max(st.relres, *[d.relres for d in d3]),
                                                     st.fell_back + sum(d.fell_back for d in d3), t])

    # plain MG-CG
    ctx, dt = make()
    ref = []
    for cyc in range(1, total + 1):
        t0 = time.perf_counter()
        st, _ = ctx.step(dt, tol=tol, maxiter=400, cycle=cyc)
        ctx.sync()
        if cyc > total - eval_steps:
            record(ctx, st, ref, 1e3 * (time.perf_counter() - t0))
    ctx.close()
    if log:
        log("[darwin-learned] MG-CG run done")
    # learned
    ctx, dt = make()
    cyc = 0

    def step(solver_=0):
        nonlocal cyc
        cyc += 1
        t0 = time.perf_counter()
        st, _ = ctx.step(dt, tol=tol, maxiter=400, solver=solver_, cycle=cyc)
        ctx.sync()
        return st, 1e3 * (time.perf_counter() - t0)

    for _ in range(warmup):
        step()
    ctx.nn_init(32, seed=1)
    ctx.nn_config(1, 2, 0)
    for _ in range(poisson_pairs + 1):
        step()
        ctx.nn_add_increment()

    def train(budget):
        t = time.perf_counter()
        losses = [ctx.nn_train(epochs=1, batch=4, lr=0.0)]
        while time.perf_counter() - t < budget and len(losses) < 200:
            losses.append(ctx.nn_train(epochs=1, batch=4, lr=1e-3))
        return losses

    lp = train(train_seconds / 2)
    if log:
        log(f"[darwin-learned] Poisson training: loss {lp[0]:.3f} -> {lp[-1]:.4f} in {len(lp) - 1} epochs")
    ctx.nn_set_system(1)
    added = 0
    for _ in range(darwin_steps + 1):
        step()
        added += ctx.nn_add_darwin_increments()
    ld = train(train_seconds / 2)
    if log:
        log(f"[darwin-learned] Darwin fine-tuning on {added} pairs: loss {ld[0]:.3f} -> {ld[-1]:.4f} "
            f"in {len(ld) - 1} epochs")
    got = []
    for _ in range(eval_steps):
        st, t = step(solver)
        record(ctx, st, got, t)
    ctx.close()
    names = ["poisson_iterations", "poisson_ms", "chi_iterations", "ax_iterations", "ay_iterations", "chi_ms",
             "ax_ms", "ay_ms"]
    A, B = np.array(ref, dtype=float), np.array(got, dtype=float)
    res = {"mg_cg": {k: float(A[:, i].mean()) for i, k in enumerate(names)},
           "learned": {k: float(B[:, i].mean()) for i, k in enumerate(names)}}
    for key, M in (("mg_cg", A), ("learned", B)):
        res[key]["max_relres"] = float(M[:, 8].max())
        res[key]["fell_back"] = int(M[:, 9].sum())
        res[key]["step_ms"] = float(M[:, 10].mean())
        res[key]["darwin_ms"] = float(M[:, 5:8].sum(axis=1).mean())
    res["darwin_time_ratio"] = res["learned"]["darwin_ms"] / res["mg_cg"]["darwin_ms"]
    res["poisson_time_ratio"] = res["learned"]["poisson_ms"] / res["mg_cg"]["poisson_ms"]
    return dict(config=dict(ngx=ngx, ppc=ppc, order=order, tol=tol, warmup=warmup, poisson_pairs=poisson_pairs,
                            darwin_steps=darwin_steps, darwin_pairs=added, eval_steps=eval_steps, solver=solver),
                poisson_loss=[lp[0], lp[-1], len(lp) - 1], darwin_loss=[ld[0], ld[-1], len(ld) - 1], **res)


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
    ap = argparse.ArgumentParser(description="bounded-plasma, Darwin or learned-preconditioner run on one GPU")
    ap.add_argument("which", choices=["bounded", "darwin", "learned", "darwin_learned"])
    ap.add_argument("--kind", choices=["periodic", "bounded"], default="periodic",
                    help="learned: the workload the pairs come from")
    ap.add_argument("--pairs", type=int, default=10000, help="learned: training increments")
    ap.add_argument("--eval-fields", type=int, default=8)
    ap.add_argument("--train-seconds", type=float, default=120.0)
    ap.add_argument("--save", default=None, help="learned: write the trained network here (gp_nn_save)")
    ap.add_argument("--solver", type=int, default=2, help="darwin_learned: 2 network warm start, 1 flexible CG")
    ap.add_argument("--ngx", type=int, default=512)
    ap.add_argument("--ppc", type=int, default=None)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--order", type=int, default=3)
    ap.add_argument("--history", default=None, help="write the per-cycle history here as JSON")
    a = ap.parse_args(argv)
    log = lambda m: print(m, flush=True)
    if a.which == "darwin_learned":
        out = darwin_learned(ngx=a.ngx, ppc=a.ppc or 128, train_seconds=a.train_seconds, order=a.order,
                             poisson_pairs=a.pairs, darwin_steps=a.pairs, eval_steps=a.steps, solver=a.solver, log=log)
        print(json.dumps(out), flush=True)
        return
    if a.which == "learned":
        out = learned(kind=a.kind, ngx=a.ngx, ppc=a.ppc, pairs=a.pairs, eval_fields=a.eval_fields,
                      train_seconds=a.train_seconds, order=a.order, save=a.save, log=log)
        print(json.dumps(out), flush=True)
        return
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
