"""Benchmark of the MI355X PIC step (one JSON line on stdout).

Workload (the metric's configuration): 1024 x 1024 periodic grid, 3rd-order
B-spline weighting, two species (electrons, Xe+ at mass ratio 239,332) at
128 particles per cell = 268,435,456 particles, uniform Bz with
omega_ce/omega_pe = 0.5, electrostatic, dx = 0.5 Debye length,
dt = 0.1/omega_pe, FP32 particles.  A step is one cycle of the reference's
loop (main.mm:154-254): Poisson solve (GPU MG-CG to a 1e-8 relative
residual, FP64 refinement) -> fused gather + Boris push + int64 deposit of
both species -> tile sort of a species when its last push gathered more
than a fraction (default 3e-3) of its particles outside their tile windows
(--sort-every N adds a fixed interval).  Particles are generated on
the device (seeded PCG32 streams), so inputs are resident in HBM when the
timed region starts.

The roofline object is for the fused push+deposit kernel.  Its algorithmic
bytes are the per-particle-step figure of the hot-path scope table, 40 B
(x, y, vx, vy, vz read and written in FP32, the reference's unit of work),
times the particles of a launch, over the kernel's mean duration from HIP
events recorded on the context stream around each of its launches.  The
timed steps themselves carry no event marks (each is a gap between kernels):
the same number of steps follows with the events, and the kernel durations,
species_push_ms (each species' push with the window reduction and energy
sum after it) and breakdown_ms_per_step come from that window.  The kernel itself moves 32 B of them:
vz never changes (Ez = Bx = By = 0), so the particles-periodic push neither
reads nor writes it; kernel_bytes_per_particle and frac_at_kernel_bytes give
that view.  stream_ceiling is the same 32 B per particle (read x, y, vx, vy,
write them) moved by a kernel that does nothing else (gp_stream_probe): the
HBM rate this access pattern reaches on the box, the practical ceiling of
the push's own bandwidth.  sort_cycle times the steps from one drift-
triggered sort of the electrons to the next after the timed steps, so its
rate carries the sort's amortised cost whatever the timed window held.
cpu_baseline times the C/OpenMP oracle's push+deposit on a bounded sample of
the same workload on the host cores.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

ME = 9.10938188e-28          # g (Constant.h)
EC = 4.803204197e-10         # statC
CL = 2.99792458e10           # cm/s
KB = 1.3806503e-16
EVTOK = 11604.0
MXE = 2.1801714e-22          # Xe+ mass (data/condition_large.json)
HBM_PEAK = 8.0e12            # MI355X HBM3E, B/s
METRIC = "particle-steps/s + HBM%, 1024² 3rd-order; Poisson time to 1e-8: NN-PCG vs MG-CG"


def plasma(n0=5e10, te_ev=10.0, ti_ev=1.0, ngx=1024, ppc=128):
    te, ti = te_ev * EVTOK, ti_ev * EVTOK
    wpe = np.sqrt(4 * np.pi * n0 * EC * EC / ME)
    lde = np.sqrt(KB * te / (4 * np.pi * n0 * EC * EC))
    dx = 0.5 * lde
    dt = 0.1 / wpe
    bz = 0.5 * wpe * ME * CL / EC
    w = n0 * dx * dx / ppc                                  # Init.mm:192-200 InitialNumDens
    vth_e = np.sqrt(2 * KB * te / ME)                       # Collision.mm:264 (sqrt(2kT/m))
    vth_i = np.sqrt(2 * KB * ti / MXE)
    return dict(dx=dx, dt=dt, bz=bz, w=w, vth_e=vth_e, vth_i=vth_i, wpe=wpe)


def gpu_run(args):
    import gample_amd as gp
    P = plasma(ngx=args.ngx, ppc=args.ppc)
    ngx = ngy = args.ngx
    n = args.ppc * ngx * ngy
    cfg = gp.make_config(ngx, ngy, P["dx"], P["dx"], order=args.order, boris=2, b0=(0.0, 0.0, P["bz"]))
    ctx = gp.Context(cfg)
    se = ctx.add_species(-EC, ME, P["w"], n, vth_hint=P["vth_e"] / np.sqrt(2))
    si = ctx.add_species(EC, MXE, P["w"], n, vth_hint=P["vth_i"] / np.sqrt(2))
    G = gp.Gen()
    G.gen_type = 0
    G.xmin, G.xmax, G.ymin, G.ymax = 0.0, ngx * float(np.float32(P["dx"])), 0.0, ngy * float(np.float32(P["dx"]))
    G.ele_vth, G.ion_vth = P["vth_e"], P["vth_i"]
    ctx.load_pair(se, si, G, 0, 0, n)
    ctx.sort(se)
    ctx.sort(si)
    ctx.set_sort_drift(args.sort_drift)
    ctx.reset_grids()
    ctx.deposit(se)
    ctx.deposit(si)
    dt, tol, sort_every = P["dt"], 1e-8, args.sort_every
    cyc = 0
    for _ in range(args.warmup):
        cyc += 1
        ctx.step(dt, tol=tol, maxiter=400, sort_interval=sort_every, cycle=cyc)
    ctx.sync()
    # the timed steps run without event marks (each mark is a gap between
    # kernels); the same number of steps follows with HIP events around the
    # push kernel, each species' push, the sorts and the solves
    sorts0 = [ctx.sort_stats(se)[0], ctx.sort_stats(si)[0]]
    iters = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        cyc += 1
        st, _ = ctx.step(dt, tol=tol, maxiter=400, sort_interval=sort_every, cycle=cyc)
        iters.append(st.iterations)
    ctx.sync()
    t1 = time.perf_counter()
    sorts = [ctx.sort_stats(se)[0] - sorts0[0], ctx.sort_stats(si)[0] - sorts0[1]]
    ctx.timing(True)
    sorts1 = [ctx.sort_stats(se)[0], ctx.sort_stats(si)[0]]
    for _ in range(args.steps):
        cyc += 1
        ctx.step(dt, tol=tol, maxiter=400, sort_interval=sort_every, cycle=cyc)
    ctx.sync()
    sorts_ev = [ctx.sort_stats(se)[0] - sorts1[0], ctx.sort_stats(si)[0] - sorts1[1]]
    push_ms, push_n = ctx.kernel_time(0)
    pushk_ms, pushk_n = ctx.kernel_time(28)
    e_ms, e_n = ctx.kernel_time(4 + 3 * se)
    i_ms, i_n = ctx.kernel_time(4 + 3 * si)
    sort_ms, sort_n = ctx.kernel_time(1)
    solve_ms, solve_n = ctx.kernel_time(2)
    ctx.timing(False)
    probe_ms = ctx.stream_probe(se, 10)
    cycle_rep = sort_cycle(ctx, se, dt, tol, sort_every, cyc, args.cycle_max_steps)
    cyc += cycle_rep.pop("steps_run")
    _, dg, _ = ctx.solve_fields(dt, tol=tol, diag=True)
    nn = ctx.count(se) + ctx.count(si)
    nnrep = None if args.no_nn else nn_poisson(ctx, args, dt, tol, cyc, float(np.float32(P["dx"])))
    ctx.close()
    return dict(n_total=nn, seconds=t1 - t0, push_ms=push_ms, push_n=push_n, pushk_ms=pushk_ms, pushk_n=pushk_n,
                species_push_ms={"electrons": e_ms / max(1, e_n), "ions": i_ms / max(1, i_n)},
                probe_ms=probe_ms,
                sort_ms=sort_ms, sort_n=sort_n,
                solve_ms=solve_ms, solve_n=solve_n, iters=iters, P=P, n_species=n, nn=nnrep, sorts=sorts,
                sorts_ev=sorts_ev,
                cycle=cycle_rep)


def sort_cycle(ctx, se, dt, tol, sort_every, cyc, max_steps):
    """Steps from one drift-triggered sort of the electrons (the species
    sorted most often) to the next, timed as a whole (sorts, solves and
    pushes of both species): the throughput with the sort amortised."""
    n_all = ctx.count(se) + ctx.count(1 - se)
    run = 0

    def step():
        nonlocal cyc, run
        cyc += 1
        run += 1
        ctx.step(dt, tol=tol, maxiter=400, sort_interval=sort_every, cycle=cyc)

    s0 = ctx.sort_stats(se)[0]
    while ctx.sort_stats(se)[0] == s0 and run < max_steps:
        step()
    if ctx.sort_stats(se)[0] == s0:
        return {"steps_run": run, "cycle_steps": None, "note": f"no electron sort within {max_steps} steps"}
    ctx.sync()
    s1, k0 = ctx.sort_stats(se)[0], run
    t0 = time.perf_counter()
    while ctx.sort_stats(se)[0] == s1 and run < 2 * max_steps:
        step()
    if ctx.sort_stats(se)[0] == s1:
        return {"steps_run": run, "cycle_steps": None, "note": f"no second electron sort within {max_steps} steps"}
    # the last step began with the next sort: the k timed steps hold exactly
    # one electron sort
    ctx.sync()
    t1 = time.perf_counter()
    k = run - k0
    return {"steps_run": run, "cycle_steps": k, "ms_per_step": 1e3 * (t1 - t0) / k,
            "value": n_all * k / (t1 - t0),
            "note": "the steps after one electron sort up to and including the step of the next: one sort "
                    "cycle, with its sort, solves and pushes"}


def nn_poisson(ctx, args, dt, tol, cyc, h):
    """Learned warm start on the time loop's own increments.  After the timed
    steps the run continues; step n's right-hand side b_n and solution x_n
    give the pair (b_n - b_{n-1}, x_n - x_{n-1}): the first residual of a solve
    warm-started from x_{n-1}, and the correction it needs.  The network is
    trained on these pairs with its relative energy-norm loss; then held-out
    steps are solved from x_{n-1} with MG-CG and with each network mode, and
    from zero with MG-CG.  Also reported, on the held-out increments: the
    network's relative energy-norm error (its output at the S-optimal scale)
    and the residual it leaves at the residual-optimal scale, against the
    warm start's."""
    import gample_amd as gp
    t0 = time.perf_counter()
    npair, nev = args.nn_pairs, args.nn_eval_fields
    my, mx = ctx.field_shape(gp.F_SOL)[0]
    db = np.empty((npair + nev, my * mx)); dx = np.empty_like(db)
    bprev = xprev = None
    for k in range(-1, npair + nev):
        cyc += 1
        ctx.step(dt, tol=tol, maxiter=400, sort_interval=args.sort_every, cycle=cyc)
        b, x = ctx.field(gp.F_RHS).ravel().copy(), ctx.field(gp.F_SOL).ravel().copy()
        if k >= 0:
            db[k] = b - bprev
            dx[k] = x - xprev
        bprev, xprev = b, x
    ctx.nn_init(32, seed=1)
    ctx.nn_config(1, 2, 0)
    ctx.nn_train(epochs=0, rho=db[:npair], phi=dx[:npair])
    t1 = time.perf_counter()
    loss0 = ctx.nn_train(epochs=1, batch=4, lr=0.0)
    epochs, loss = 0, loss0
    while epochs < args.nn_max_epochs and time.perf_counter() - t1 < args.nn_train_seconds:
        loss = ctx.nn_train(epochs=1, batch=4, lr=1e-3)
        epochs += 1
    t2 = time.perf_counter()

    def S(v):      # the solver's operator on the periodic unknown grid
        u = v.reshape(my, mx)
        return ((4 * u - np.roll(u, 1, 0) - np.roll(u, -1, 0) - np.roll(u, 1, 1) - np.roll(u, -1, 1)) / h ** 2).ravel()

    rows = {k: [] for k in ("mg_cold", "mg_warm", "nn_warm_mg", "nn_warm_fcg")}
    err, res_warm, res_nn = [], [], []
    # held-out steps: b_n = b_{n-1} + db, x_{n-1} = x_n - dx, counted back from the last step
    b, x = bprev, xprev
    for k in range(npair + nev - 1, npair - 1, -1):
        bp, xp = b - db[k], x - dx[k]
        for key, solver, x0 in (("mg_cold", 0, None), ("mg_warm", 0, xp), ("nn_warm_mg", 2, xp),
                                ("nn_warm_fcg", 1, xp)):
            _, st, _ = ctx.solve_rhs(b, x0=x0, tol=tol, maxiter=400, solver=solver)
            rows[key].append((st.iterations, 1e3 * st.seconds, st.relres, st.fell_back))
        z = ctx.nn_infer(db[k])
        sz = S(z)
        a = (z @ db[k]) / (z @ sz)            # S dx = db: the S-optimal multiple of z
        e = dx[k] - a * z
        err.append(np.sqrt((e @ S(e)) / (dx[k] @ S(dx[k]))))
        ar = (db[k] @ sz) / (sz @ sz)
        nb = np.linalg.norm(b)
        res_warm.append(np.linalg.norm(db[k]) / nb)
        res_nn.append(np.linalg.norm(db[k] - ar * sz) / nb)
        b, x = bp, xp
    out = {"train_pairs": npair, "pairs": "time-loop increments (b_n - b_{n-1}, x_n - x_{n-1})",
           "harvest_s": t1 - t0, "train_s": t2 - t1, "epochs": epochs, "loss_untrained": loss0, "loss_trained": loss,
           "network": "4-level conv encoder-decoder 32-64-128-128, 3x3, BF16 MFMA / FP32 accumulate",
           "fields": nev, "tol": tol,
           "heldout_energy_error": float(np.mean(err)), "heldout_relres_warm_start": float(np.mean(res_warm)),
           "heldout_relres_after_network": float(np.mean(res_nn))}
    for k, v in rows.items():
        a = np.array(v, dtype=float)
        out[k] = {"iterations": float(a[:, 0].mean()), "ms": float(a[:, 1].mean()),
                  "max_relres": float(a[:, 2].max()), "fell_back": int(a[:, 3].sum())}
    out["nn_warm_mg_over_mg_warm_time"] = out["nn_warm_mg"]["ms"] / out["mg_warm"]["ms"]
    out["nn_warm_fcg_over_mg_cold_time"] = out["nn_warm_fcg"]["ms"] / out["mg_cold"]["ms"]
    return out


def cpu_baseline(args, P):
    """C/OpenMP oracle push + deposit on a bounded sample of the workload."""
    from oracle import pyoracle as orc
    ngx = args.ngx
    n = args.cpu_sample
    threads = int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count() or 1)
    rng = np.random.default_rng(0)
    g = orc.grid(ngx, ngx, args.order, boris=2)
    ngb = args.order // 2
    nx = ngx + 2 * ngb
    Ex = (rng.standard_normal((nx + 1, nx + 2)) * 1e-3).astype(np.float32)
    Ey = (rng.standard_normal((nx + 2, nx + 1)) * 1e-3).astype(np.float32)
    Bz = np.full((nx + 2, nx + 2), P["bz"], np.float32)
    sp = []
    for vth, q, m in ((P["vth_e"], -EC, ME), (P["vth_i"], EC, MXE)):
        x = (ngb + rng.random(n) * ngx).astype(np.float32)
        y = (ngb + rng.random(n) * ngx).astype(np.float32)
        v = (vth / np.sqrt(2) * rng.standard_normal((3, n))).astype(np.float32)
        cE = float(np.float32(0.5 * float(np.float32(q)) * P["dt"] / float(np.float32(m))))
        cB = float(np.float32(0.5 * float(np.float32(q)) * P["dt"] / float(np.float32(m)) / CL))
        cX = float(np.float32(P["dt"] / float(np.float32(P["dx"]))))
        sp.append(([x, y, v[0], v[1], v[2]], (cE, cB, cX, cX)))

    def one_step(state):
        out = []
        for parts, cs in state:
            p2 = orc.push_fast(g, Ex, Ey, Bz, cs, parts, uniform_bz=True, bz0=float(np.float32(P["bz"])),
                               threads=threads)
            orc.deposit_fx(g, p2, threads=threads)
            out.append((p2, cs))
        return out

    sp = one_step(sp)                     # warm-up
    steps = 0
    t0 = time.perf_counter()
    while steps < args.cpu_steps or time.perf_counter() - t0 < args.cpu_seconds:
        sp = one_step(sp)
        steps += 1
    t1 = time.perf_counter()
    rate = 2 * n * steps / (t1 - t0)
    return dict(value=rate, unit="particle-steps/s", cores=threads, kind="port", cpu_model=cpu_model(),
                sample=f"oracle push_fast + deposit_fx (C/OpenMP, -O2 x86-64-v3) of 2 x {n:,} particles on the "
                       f"{ngx}x{ngx} 3rd-order grid, {steps} timed steps ({t1 - t0:.1f} s) after 1 warm-up; "
                       f"Poisson not included")


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def copy_calibration(mib):
    """A plain device-to-device copy of known size: reads and writes mib MiB."""
    import torch
    n = mib * (1 << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda:0")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(f"copy calibration: {mib} MiB read + {mib} MiB written per launch", file=sys.stderr, flush=True)


def traffic_from_profiles():
    """HBM bytes per launch of the fused kernel from the committed rocprof
    counter summary (FETCH_SIZE and WRITE_SIZE, each scaled by a plain copy)."""
    f = os.path.join(ROOT, "profiles", "push_deposit_counters.md")
    try:
        for line in open(f):
            if line.startswith("traffic_bytes_per_launch:"):
                return {"bytes_per_launch": float(line.split(":")[1])}
    except OSError:
        pass
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--ngx", type=int, default=1024)
    ap.add_argument("--ppc", type=int, default=128)
    ap.add_argument("--order", type=int, default=3)
    ap.add_argument("--sort-every", type=int, default=0, help="fixed tile-sort interval (0: none)")
    ap.add_argument("--sort-drift", type=float, default=3e-3,
                    help="sort a species when this fraction of it drifted out of its tile windows (0: off)")
    ap.add_argument("--cpu-sample", type=int, default=4_000_000)
    ap.add_argument("--cpu-steps", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-nn", action="store_true", help="skip the learned Poisson preconditioner report")
    ap.add_argument("--nn-pairs", type=int, default=150, help="time-loop increments to train the network on")
    ap.add_argument("--nn-train-seconds", type=float, default=30.0)
    ap.add_argument("--nn-max-epochs", type=int, default=15)
    ap.add_argument("--nn-eval-fields", type=int, default=4)
    ap.add_argument("--cycle-max-steps", type=int, default=60,
                    help="steps allowed to reach each of the two electron sorts that bound the sort cycle")
    ap.add_argument("--copy-calib", type=int, default=0, metavar="MIB",
                    help="first run a plain device copy of MIB MiB (3 times), the reference kernel for scaling "
                         "the FETCH_SIZE / WRITE_SIZE counters of a --pmc run")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; the multi-GPU (particle-sharded, int64 "
                                   "all-reduce of the charge grid) design is described in DESIGN.md",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    if args.copy_calib:
        copy_calibration(args.copy_calib)
    r = gpu_run(args)
    ms_step = 1e3 * r["seconds"] / args.steps
    value = r["n_total"] * args.steps / r["seconds"]
    push_avg_ms = r["pushk_ms"] / max(1, r["pushk_n"])
    bytes_per_launch = 40.0 * r["n_species"]        # per-unit figure of the scope table x particles
    kernel_bytes = 32.0 * r["n_species"]            # what the kernel moves (vz neither read nor written)
    achieved = bytes_per_launch / (push_avg_ms * 1e-3)
    tr = traffic_from_profiles()
    roof = {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
            "frac": achieved / HBM_PEAK, "traffic": (tr or {}).get("bytes_per_launch"),
            "kernel": "k_push<3,uniform B,rho deposit,push> (fused gather+Boris+int64 deposit)",
            "algorithmic_bytes_per_launch": bytes_per_launch, "algorithmic_bytes_per_particle": 40,
            "kernel_bytes_per_particle": 32,
            "frac_at_kernel_bytes": kernel_bytes / (push_avg_ms * 1e-3) / HBM_PEAK, "avg_launch_ms": push_avg_ms,
            "launches": r["pushk_n"],
            "timing": "HIP events around each k_push launch on the context stream, in a window of as many steps "
                      "after the timed ones (the timed steps carry no event marks)",
            "species_push_ms": r["species_push_ms"],
            "stream_ceiling": {"kernel": "k_stream_probe: read x,y,vx,vy, write them (32 B per particle)",
                               "ms": r["probe_ms"], "achieved": kernel_bytes / (r["probe_ms"] * 1e-3) / 1e9,
                               "frac": kernel_bytes / (r["probe_ms"] * 1e-3) / HBM_PEAK,
                               "push_over_stream": push_avg_ms / r["probe_ms"]}}
    line = {
        "metric": METRIC, "value": value, "unit": "particle-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "fp32 (fp64 Poisson refinement and reductions)", "data": "synthetic",
        "config": {"workload": f"PIC step {args.ngx}x{args.ngx} periodic, order {args.order} B-spline, "
                               f"{args.ppc} ppc x 2 species ({r['n_total']:,} particles), uniform Bz "
                               f"(wce/wpe=0.5), electrostatic, Poisson MG-CG to 1e-8 (FP64 refined), "
                               + (f"tile sort every {args.sort_every} steps, " if args.sort_every else "")
                               + f"tile sort of a species at drift fraction > {args.sort_drift:g}",
                   "global_batch": r["n_total"], "parallelism": "single GPU"},
        "roofline": roof,
        "breakdown_ms_per_step": {"push_deposit": r["push_ms"] / args.steps, "sort": r["sort_ms"] / args.steps,
                                  "field_solve": r["solve_ms"] / args.steps},
        "sorts_timed": {"electrons": r["sorts"][0], "ions": r["sorts"][1]},
        "sorts_in_event_window": {"electrons": r["sorts_ev"][0], "ions": r["sorts_ev"][1]},
        "sort_cycle": r["cycle"],
        "poisson": {"mg_cg_iterations_mean": float(np.mean(r["iters"])), "ms_per_solve":
                    r["solve_ms"] / max(1, r["solve_n"]), "nn_pcg": r["nn"]},
    }
    if not args.no_cpu:
        cb = cpu_baseline(args, r["P"])
        line["cpu_baseline"] = cb
        line["gpu_over_cpu"] = value / cb["value"]
    print(json.dumps(line))


if __name__ == "__main__":
    main()
