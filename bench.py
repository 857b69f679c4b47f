"""Benchmark of the staggered CG hot path on one MI355X.

A step is one CG solve of (D D^dag) x = b to a true FP64 relative residual of
1e-10 on a 32^4 lattice (1,048,576 sites; links U = exp(i 0.3 H), m = 0.01,
one Gaussian right-hand side), run in mixed precision: FP16 inner iterations
(per-site max-norm, FP32 accumulation), FP64 reliable update whenever the
iterated residual drops 10x, FP64 reductions.  The plain FP64 CG on the same
inputs is timed beside it, also stopped on the FP64 true residual.  Inputs
are resident in HBM before the timed region starts.

Prints one JSON line: value = mixed-precision solves/s, the FP64 solves/s,
p50/p99 solve latency, the HBM roofline of the dominant kernel of the mixed
solve (the second hop of each iteration with its dot products) and of the
Dslash (hop) alone in both precisions, measured with HIP events on the
library stream, and the CPU baseline (the FP64 oracle, timed on a bounded
sample on the host's cores and on one core).

``--workload 16^4`` runs the 16^4, m = 0.05 solve with FP32 inner iterations
instead; ``batch`` the 64 uneven batched systems, ``serve`` the solve service
under Poisson arrivals, ``precision`` the errors of the mixed-precision
results against the FP64 oracle at 16^4, ``traj32`` a 32^4 Symanzik HMC
trajectory in FP64 and FP16-mixed, ``heatbath`` Cabibbo-Marinari heatbath
sweeps of the Symanzik action at 32^4 and 16^4, ``hex`` HEX smearing steps
at 32^4 and 16^4, ``force`` the gauge and fermion force kernels at 32^4 and
16^4.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "CG solves/s + Dslash HBM GB/s, mixed vs FP64 at 1e-10 resid; p99 solve latency"
PEAK_HBM_GBS = 8000.0

WORKLOADS = {
    "16^4": dict(dims=(16, 16, 16, 16), eps=0.3, mass=0.05, mixed="f32",
                 desc="16^4 lattice, links exp(i 0.3 H), m=0.05, 1 Gaussian RHS, CG on D D^dag to true rel. "
                      "residual 1e-10: FP32 inner + FP64 reliable updates vs plain FP64"),
    "32^4": dict(dims=(32, 32, 32, 32), eps=0.3, mass=0.01, mixed="f16",
                 desc="32^4 lattice, links exp(i 0.3 H), m=0.01, 1 Gaussian RHS, CG on D D^dag to true rel. "
                      "residual 1e-10: FP16 inner (per-site norms, FP32 accumulation) + FP64 reliable updates "
                      "vs plain FP64"),
}

# Algorithmic (compulsory) bytes per lattice site of one full-lattice hop:
# 4 links x 18 reals, 6 reals in, 6 reals out, + per-site norms for FP16.
HOP_BYTES_PER_SITE = {"f64": 72 * 8 + 12 * 8, "f32": 72 * 4 + 12 * 4, "f16": 72 * 2 + 12 * 2 + 2 * 4}
# The hop that closes each CG iteration also reads p for <p, v>.
DOT_EXTRA_PER_SITE = {"f64": 48, "f32": 24, "f16": 16}
# The two steps of one mixed-precision iteration, compulsory bytes per site:
# A reads links, r, v, p, x and writes r, p, x, t; B reads links, t, p, r and
# writes v (FP16 vectors carry one FP32 norm per site; x is FP32).
MIX_A_BYTES_PER_SITE = {"f32": 288 + 3 * 24 + 2 * 24 + 2 * 24 + 24, "f16": 144 + 3 * 16 + 2 * 16 + 2 * 24 + 16}
MIX_B_BYTES_PER_SITE = {"f32": 288 + 24 + 24 + 2 * 24, "f16": 144 + 16 + 16 + 2 * 16}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def host_cpus():
    """CPU model, nproc (the affinity mask), the cgroup CPU quota if one is
    set, and the threads the CPU leg uses: nproc, capped by the quota."""
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    nproc = len(os.sched_getaffinity(0))
    quota = None
    try:
        q, per = open("/sys/fs/cgroup/cpu.max").read().split()[:2]
        if q != "max":
            quota = max(1, int(int(q) / int(per)))
    except (OSError, ValueError):
        pass
    threads = min(nproc, quota) if quota else nproc
    return {"model": model, "nproc": nproc, "cgroup_cpu_quota": quota, "threads": threads}


def cpu_baseline(dims, mass, U, phi, iters_per_solve, budget_s=12.0, budget_1t_s=8.0):
    """FP64 oracle (the reference's CG recurrence, AoS layout and OpenMP
    parallel-for, g++ -O3 -fopenmp) on the host: times a bounded number of CG
    iterations with OMP_NUM_THREADS = nproc (capped by the cgroup CPU quota)
    and with one thread, OMP_PROC_BIND=spread, OMP_PLACES=threads, and scales
    each to the iteration count of one FP64 solve (extrapolated)."""
    # the OpenMP runtime reads these when the oracle library is first loaded
    os.environ.setdefault("OMP_PROC_BIND", "spread")
    os.environ.setdefault("OMP_PLACES", "threads")
    from oracle import oracle as oc
    info = host_cpus()

    def per_iteration(threads, budget):
        oc.set_threads(threads)
        t1 = oc.time_cg_iterations(dims, mass, U, phi, 1)  # warm + estimate
        n = int(max(2, min(iters_per_solve, budget / max(t1, 1e-6))))
        return oc.time_cg_iterations(dims, mass, U, phi, n), n

    t_n, n_n = per_iteration(info["threads"], budget_s)
    t_1, n_1 = per_iteration(1, budget_1t_s)
    oc.set_threads(info["threads"])
    extrap = n_n < iters_per_solve
    return {
        "value": 1.0 / (t_n * iters_per_solve),
        "unit": "solves/s",
        "cores": info["threads"],
        "kind": "port",
        "sample": f"{n_n} FP64 CG iterations (D D^dag + 2 dots + 2 axpy) of the oracle at "
                  f"{'x'.join(map(str, dims))} on {info['threads']} threads, {t_n * 1e3:.1f} ms/iteration, "
                  f"scaled to the {iters_per_solve} iterations of one FP64 solve"
                  + (" (extrapolated)" if extrap else ""),
        "one_thread": {"value": 1.0 / (t_1 * iters_per_solve), "unit": "solves/s", "cores": 1,
                       "sample": f"{n_1} iterations, {t_1 * 1e3:.1f} ms/iteration, scaled to "
                                 f"{iters_per_solve} (extrapolated)"},
        "host": {"cpu_model": info["model"], "nproc": info["nproc"], "cgroup_cpu_quota": info["cgroup_cpu_quota"],
                 "omp": "OMP_PROC_BIND=spread OMP_PLACES=threads, g++ -O3 -fopenmp"},
    }


def traffic_from_profiles(kernel, workload):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 counter
    summary (FETCH_SIZE + WRITE_SIZE, calibrated on a kernel of known size)."""
    path = os.path.join(ROOT, "profiles", "kernel_traffic.json")
    if not os.path.exists(path):
        return None
    try:
        data = json.load(open(path))
        return data.get(workload, {}).get(kernel)
    except Exception:
        return None


BATCH_SIZES = [(8, 8, 8, 8), (12, 12, 12, 12), (16, 16, 16, 16), (24, 24, 24, 24)]


def make_batch_systems(dev, ctx, n_per_size=16, seed=42):
    """64 systems: 16 each of 8^4, 12^4, 16^4, 24^4; m ~ U[0.01, 0.1];
    eps in {0.2, 0.3, 0.4}; 1-8 Gaussian right-hand sides each."""
    rng = np.random.default_rng(seed)
    systems = []
    k = 0
    for dims in BATCH_SIZES:
        for _ in range(n_per_size):
            eps = float(rng.choice([0.2, 0.3, 0.4]))
            m = float(rng.uniform(0.01, 0.1))
            nrhs = int(rng.integers(1, 9))
            g = dev.Gauge(ctx, dims, dev.random_links(dims, eps, seed=10000 + k))
            rf = [dev.Field(g, dev.random_spinor(dims, seed=20000 + 16 * k + j)) for j in range(nrhs)]
            xf = [dev.Field(g) for _ in range(nrhs)]
            systems.append(dict(dims=dims, eps=eps, mass=m, g=g, rf=rf, xf=xf))
            k += 1
    return systems


def bench_batch(args, dev):
    """64 uneven systems with 1-8 RHS each, solved together (FP32 inner +
    FP64 reliable updates, graph-replayed chunks) vs one after another."""
    ctx = dev.Context(0)
    log("[bench] batch: building 64 systems")
    systems = make_batch_systems(dev, ctx)
    nrhs = sum(len(s["rf"]) for s in systems)
    spec = [(s["g"], s["mass"], s["rf"], s["xf"]) for s in systems]
    if args.kernels_only:
        # one batched solve of the 64 systems (default launch shape), then one
        # FP32 link conversion of a 24^4 lattice (reads 576 B, writes 288 B per
        # site) that calibrates the traffic counters
        dev.cg_batch(ctx, spec, tol=1e-10, precision="f32")
        g = dev.Gauge(ctx, (24, 24, 24, 24), dev.random_links((24, 24, 24, 24), 0.3, seed=5))
        g.time_hop("f32", False, 1)
        ctx.synchronize()
        log("[bench] batch kernels done")
        return
    dev.cg_batch(ctx, spec[:2], tol=1e-10, precision="f32")  # warm-up
    # the batched hop t = D^dag p with every right-hand side active, at the
    # engine's default shape (256 sites, 8 RHS per block: the shape the
    # traffic counters were read at), timed with HIP events (5 launches)
    _, _, hop_default_us = dev.tune_batch(ctx, spec, tol=1e-10, precision="f32", shape=(256, 8))
    # autotuner: launch shape of the batched hop for this batch (untimed)
    shape_threads, shape_tile, shape_us = dev.tune_batch(ctx, spec, tol=1e-10, precision="f32")
    log(f"[bench] batch: tuned hop shape {shape_threads} sites/block, {shape_tile} RHS/block ({shape_us:.1f} us)")
    ctx.synchronize()
    ctx.set_profiling(1)
    t0 = time.perf_counter()
    stats = dev.cg_batch(ctx, spec, tol=1e-10, precision="f32")
    ctx.synchronize()
    makespan = time.perf_counter() - t0
    chunk_ms, nchunks = ctx.kernel_time("batch_chunk_f32")
    ctx.set_profiling(0)
    worst = max(s["true_rel_resid"] for s in stats)
    log(f"[bench] batch: {nrhs} solves in {makespan:.3f}s, worst true residual {worst:.2e}")
    # algorithmic bytes: FP32 links read twice per iteration while any RHS of
    # the system is active, plus 336 B/site of FP32 vectors per RHS iteration
    k = 0
    link_bytes = vec_bytes = 0.0
    for s in systems:
        V = int(np.prod(s["dims"]))
        its = [stats[k + j]["iterations"] for j in range(len(s["rf"]))]
        k += len(s["rf"])
        link_bytes += 2 * 288.0 * V * max(its)
        vec_bytes += 336.0 * V * sum(its)
    # sequential baselines over the same systems: plain FP64 and mixed single solves
    seq = {}
    for prec in ("f64", "f32"):
        t0 = time.perf_counter()
        for s in systems:
            for rf, xf in zip(s["rf"], s["xf"]):
                s["g"].cg_field(rf, xf, s["mass"], tol=1e-10, max_iter=100000, precision=prec)
        ctx.synchronize()
        seq[prec] = time.perf_counter() - t0
        log(f"[bench] batch: sequential {prec} {seq[prec]:.3f}s")
    value = nrhs / makespan
    achieved = (link_bytes + vec_bytes) / (chunk_ms * 1e-3) / 1e9 if chunk_ms else 0.0
    # algorithmic bytes of one batched hop with all right-hand sides active:
    # every FP32 link once (288 B/site) and each right-hand side's input and
    # output spinor once (48 B per site and RHS)
    hop_bytes = sum(288.0 * np.prod(s["dims"]) + 48.0 * np.prod(s["dims"]) * len(s["rf"]) for s in systems)
    hop_gbs = hop_bytes / (hop_default_us * 1e-6) / 1e9
    out = {
        "metric": METRIC, "value": round(value, 3), "unit": "solves/s", "n_gpus": 1, "steps": 1, "warmup": 1,
        "ms_per_step": round(makespan * 1e3, 2), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 inner / f64 reliable updates and reductions", "data": "synthetic (seeded Philox)",
        "config": {"workload": "64 systems (16 each of 8^4, 12^4, 16^4, 24^4), m~U[0.01,0.1], eps in "
                               "{0.2,0.3,0.4}, 1-8 RHS each, batched CG to true residual 1e-10, hipGraph chunks",
                   "rhs_total": nrhs, "parallelism": "single GPU",
                   "tuned_hop_shape": {"sites_per_block": shape_threads, "rhs_per_block": shape_tile,
                                       "hop_us": round(shape_us, 2)}},
        "sequential_fp64_solves_per_s": round(nrhs / seq["f64"], 3),
        "sequential_mixed_solves_per_s": round(nrhs / seq["f32"], 3),
        "speedup_vs_sequential_fp64": round(seq["f64"] / makespan, 3),
        "worst_true_rel_residual": worst,
        "roofline": {"bound": "hbm",
                     "kernel": "kb_hop (t = D^dag p for every system and right-hand side, all active; "
                               "256 sites x 8 RHS per block)",
                     "achieved": round(hop_gbs, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": round(hop_gbs / PEAK_HBM_GBS, 4),
                     "traffic": traffic_from_profiles("kb_hop", "batch"),
                     "algorithmic_bytes_per_launch": int(hop_bytes), "avg_launch_us": round(hop_default_us, 2),
                     "chunk": {"what": "whole chunks (two hops, the vector updates and the reliable updates), "
                                       "algorithmic bytes of the iterations each solve ran",
                               "achieved": round(achieved, 1), "frac": round(achieved / PEAK_HBM_GBS, 4)}},
        "cpu_baseline": None,
    }
    print(json.dumps(out))


def bench_serve(args, dev):
    """32 chains on {8,12,16,24}^4 with resident links; Poisson arrivals at
    rates swept around the sequential FP64 capacity; p50/p99 latency and
    solves/s of the service vs a sequential FP64 CG server on the same
    arrivals (FIFO replay of measured FP64 solve times)."""
    import threading
    ctx = dev.Context(0)
    rng = np.random.default_rng(7)
    chains = []
    for k in range(32):
        dims = BATCH_SIZES[k % 4]
        m = float(rng.uniform(0.01, 0.1))
        g = dev.Gauge(ctx, dims, dev.random_links(dims, 0.3, seed=30000 + k))
        rhs = [dev.Field(g, dev.random_spinor(dims, seed=40000 + 8 * k + j)) for j in range(4)]
        chains.append(dict(dims=dims, mass=m, g=g, rhs=rhs, free=[dev.Field(g) for _ in range(8)]))
    log("[bench] serve: measuring sequential FP64 solve times")
    fp64_t = {}
    for k, c in enumerate(chains):
        t0 = time.perf_counter()
        c["g"].cg_field(c["rhs"][0], c["free"][0], c["mass"], tol=1e-10, max_iter=100000, precision="f64")
        fp64_t[k] = time.perf_counter() - t0
    mean_fp64 = float(np.mean(list(fp64_t.values())))
    capacity = 1.0 / mean_fp64
    svc = dev.Service(0, tol=1e-10)
    ids = [svc.add_chain(c["g"], c["mass"], max_inflight=8) for c in chains]
    results = {}
    for mult in [float(r) for r in args.rates.split(",")]:
        lam = mult * capacity
        n = args.requests
        gaps = rng.exponential(1.0 / lam, size=n)
        which = rng.integers(0, len(chains), size=n)
        done = {}
        inflight = {}
        lock = threading.Lock()

        def collect():
            while True:
                with lock:
                    if len(done) >= n:
                        return
                for t, st in svc.poll(timeout=0.5):
                    with lock:
                        k, xf, ta = inflight.pop(t)
                        chains[k]["free"].append(xf)
                        done[t] = (ta, st)

        th = threading.Thread(target=collect)
        th.start()
        t_start = time.perf_counter()
        t_next = t_start
        arrivals = []
        for r in range(n):
            t_next += gaps[r]
            now = time.perf_counter()
            if t_next > now:
                time.sleep(t_next - now)
            k = int(which[r])
            c = chains[k]
            while True:
                with lock:
                    xf = c["free"].pop() if c["free"] else None
                if xf is not None:
                    break
                time.sleep(0.0005)
            ta = time.perf_counter() - t_start
            with lock:
                t = svc.submit(ids[k], c["rhs"][r % 4], xf)
                inflight[t] = (k, xf, ta)
            arrivals.append((ta, k))
        th.join()
        makespan = time.perf_counter() - t_start
        lat = np.array([st["seconds"] for _, st in done.values()]) * 1e3
        worst = max(st["true_rel_resid"] for _, st in done.values())
        # sequential FP64 FIFO server on the same arrivals
        fin, seq_lat = 0.0, []
        for ta, k in arrivals:
            fin = max(fin, ta) + fp64_t[k]
            seq_lat.append((fin - ta) * 1e3)
        seq_lat = np.array(seq_lat)
        results[f"{mult:g}x"] = {
            "rate_per_s": round(lam, 3), "requests": n, "solves_per_s": round(n / makespan, 3),
            "p50_ms": round(float(np.percentile(lat, 50)), 2), "p99_ms": round(float(np.percentile(lat, 99)), 2),
            "worst_true_rel_residual": worst,
            "sequential_fp64": {"solves_per_s": round(n / fin, 3), "p50_ms": round(float(np.percentile(seq_lat, 50)), 2),
                                "p99_ms": round(float(np.percentile(seq_lat, 99)), 2)},
        }
        log(f"[bench] serve {mult}x: {results[f'{mult:g}x']}")
    svc.close()
    top = results[max(results, key=lambda k: float(k[:-1]))]
    out = {
        "metric": METRIC, "value": top["solves_per_s"], "unit": "solves/s", "n_gpus": 1, "steps": args.requests,
        "warmup": 0, "ms_per_step": round(1e3 / max(top["solves_per_s"], 1e-9), 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32 inner / f64 reliable updates and reductions",
        "data": "synthetic (seeded Philox)",
        "config": {"workload": "serving: 32 chains on {8,12,16,24}^4 with resident links (eps=0.3), "
                               "m~U[0.01,0.1], Poisson arrivals, continuous batching, true residual 1e-10",
                   "parallelism": "single GPU"},
        "p50_solve_ms": top["p50_ms"], "p99_solve_ms": top["p99_ms"],
        "sequential_fp64_capacity_per_s": round(capacity, 3),
        "rates": results,
        "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(out))


def _base_line(value, unit, steps, ms, dtype, workload, extra):
    out = {"metric": METRIC, "value": value, "unit": unit, "n_gpus": 1, "steps": steps, "warmup": 1,
           "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": dtype,
           "data": "synthetic (seeded Philox links, momenta and Gaussian fields)",
           "config": {"workload": workload, "parallelism": "single GPU"}, "roofline": None, "cpu_baseline": None}
    out.update(extra)
    return out


def bench_precision(args, dev):
    """Errors of the mixed-precision results against the FP64 oracle on the
    16^4, m = 0.05 lattice (links exp(i 0.3 H)): solution and true residual
    of one solve, the force L-inf built from it, and one HMC trajectory
    (Wilson beta = 5.6, tau = 1, 10 steps, one pseudofermion): Delta H and
    the plaquette at its end, for the device FP64, FP32-mixed and
    FP16-mixed paths."""
    from oracle import oracle as oc
    oc.set_threads(min(16, os.cpu_count() or 1))
    dims, m, tol = (16, 16, 16, 16), 0.05, 1e-10
    U = dev.random_links(dims, 0.3, seed=1234)
    chi = dev.random_spinor(dims, seed=99)
    P0 = dev.random_momenta(dims, seed=77)
    ctx = dev.Context(0)
    g = dev.Gauge(ctx, dims, U)
    phi = g.apply_d(m, chi)
    log("[bench] precision: oracle solve and force")
    xo, *_ = oc.cg(dims, m, U, phi, tol=1e-13)
    Fo = oc.force(dims, U, xo, oc.apply_d(dims, m, U, xo, dagger=True), 5.6, 1.0, 0.0)
    solve, force = {}, {}
    for prec in ("f64", "f32", "f16"):
        x, st = g.cg(phi, m, tol=tol, precision=prec, max_iter=100000)
        r = phi - oc.apply_ddag(dims, m, U, x)
        F = g.force(5.6, 1.0, 0.0, x, g.apply_d(m, x, dagger=True))
        solve[prec] = {"iterations": st["iterations"], "true_rel_residual": float(np.sqrt(r @ r / (phi @ phi))),
                       "rel_error_vs_oracle": float(np.linalg.norm(x - xo) / np.linalg.norm(xo))}
        force[prec] = {"linf_error": float(np.max(np.abs(F - Fo))), "linf_force": float(np.max(np.abs(Fo))),
                       "rel_linf_error": float(np.max(np.abs(F - Fo)) / np.max(np.abs(Fo)))}
    md = dict(mass=m, beta=5.6, c11=1.0, c12=0.0, eps=0.1, nsteps=10)
    log("[bench] precision: oracle trajectory")
    t0 = time.perf_counter()
    Uo, Ho, _ = oc.hmc_trajectory(dims, m, U, chi, P0, 0.0, beta=5.6, c11=1.0, c12=0.0, eps=0.1, nsteps=10,
                                  tol=1e-13)
    t_oracle = time.perf_counter() - t0
    plaq_o = oc.avg_plaq(dims, Uo)
    traj = {"oracle": {"delta_H": float(-Ho[6]), "plaquette": plaq_o, "seconds": t_oracle}}
    for prec in ("f64", "f32", "f16"):
        g.upload(U)
        t0 = time.perf_counter()
        H, _ = g.hmc_trajectory(chi, P0, 0.0, tol=tol, max_iter=100000, precision=prec, **md)
        ctx.synchronize()
        dt = time.perf_counter() - t0
        plaq = g.avg_plaq()
        traj[prec] = {"delta_H": float(-H[6]), "delta_H_error": float(abs(H[6] - Ho[6])), "plaquette": plaq,
                      "plaquette_error": float(abs(plaq - plaq_o)), "seconds": round(dt, 4)}
        log(f"[bench] precision: {prec} trajectory dH {-H[6]:.6f} (oracle {-Ho[6]:.6f}) in {dt:.3f}s")
    print(json.dumps(_base_line(
        round(1.0 / traj["f32"]["seconds"], 3), "trajectories/s", 1, round(traj["f32"]["seconds"] * 1e3, 2),
        "f32 inner / f64 reliable updates, reductions, forces and Metropolis",
        "16^4, m=0.05, links exp(i 0.3 H): errors against the FP64 CPU oracle of one solve to true residual 1e-10, "
        "of the force built from it, and of one Wilson HMC trajectory (beta 5.6, tau 1, 10 steps)",
        {"solve": solve, "force": force, "trajectory": traj})))


def bench_traj32(args, dev):
    """One HMC trajectory at 32^4 (Symanzik tree level c11 = 5/3, c12 = -1/12,
    beta = 4.0, tau = 1, 20 steps, m = 0.01, one pseudofermion): device FP64
    CG against FP16-mixed CG (FP64 reliable updates), both to a true
    relative residual of 1e-10, from the same start; time per trajectory,
    Delta H, plaquette, and the force L-inf difference at the start."""
    dims, m, tol = (32, 32, 32, 32), 0.01, 1e-10
    U = dev.random_links(dims, 0.3, seed=1234)
    chi = dev.random_spinor(dims, seed=99)
    P0 = dev.random_momenta(dims, seed=77)
    ctx = dev.Context(0)
    g = dev.Gauge(ctx, dims, U)
    phi = g.apply_d(m, chi)
    md = dict(mass=m, beta=4.0, c11=5.0 / 3.0, c12=-1.0 / 12.0, eps=0.05, nsteps=20)
    F = {}
    for prec in ("f64", "f16"):
        x, _ = g.cg(phi, m, tol=tol, precision=prec, max_iter=100000)
        F[prec] = g.force(4.0, 5.0 / 3.0, -1.0 / 12.0, x, g.apply_d(m, x, dagger=True))
    traj = {}
    for prec in ("f64", "f16"):
        g.upload(U)
        t0 = time.perf_counter()
        H, _ = g.hmc_trajectory(chi, P0, 0.0, tol=tol, max_iter=100000, precision=prec, **md)
        ctx.synchronize()
        dt = time.perf_counter() - t0
        traj[prec] = {"seconds": round(dt, 4), "delta_H": float(-H[6]), "plaquette": g.avg_plaq()}
        log(f"[bench] traj32: {prec} trajectory {dt:.3f}s dH {-H[6]:.6f}")
    dF = float(np.max(np.abs(F["f16"] - F["f64"])))
    # the Symanzik gauge force kernel alone (HIP events on the library stream)
    ctx.set_profiling(True)
    for _ in range(3):
        g.force(4.0, 5.0 / 3.0, -1.0 / 12.0, which=1)
    fms, fn = ctx.kernel_time("force")
    ctx.set_profiling(False)
    force_ms = fms / fn
    force_tfs = int(np.prod(dims)) * 4 * 85 * 198 / (force_ms * 1e-3) / 1e12
    print(json.dumps(_base_line(
        round(1.0 / traj["f16"]["seconds"], 4), "trajectories/s", 1, round(traj["f16"]["seconds"] * 1e3, 1),
        "f16 inner / f64 reliable updates, reductions, forces and Metropolis",
        "32^4 HMC trajectory, Symanzik (5/3, -1/12), beta 4.0, tau 1, 20 steps, m=0.01, CG to true residual "
        "1e-10: FP16-mixed vs FP64 device CG",
        {"trajectory": traj, "speedup_mixed_vs_fp64": round(traj["f64"]["seconds"] / traj["f16"]["seconds"], 3),
         "delta_H_difference": abs(traj["f16"]["delta_H"] - traj["f64"]["delta_H"]),
         "plaquette_difference": abs(traj["f16"]["plaquette"] - traj["f64"]["plaquette"]),
         "force_linf_difference": dF, "force_linf": float(np.max(np.abs(F["f64"]))),
         "force_rel_linf_difference": dF / float(np.max(np.abs(F["f64"]))),
         "gauge_force_kernel_ms": round(force_ms, 3), "gauge_force_fp64_tflops": round(force_tfs, 2)})))


# Symanzik gauge force per site: 4 links x (12 plaquette + 72 rectangle
# staple products + U A) = 85 SU(3) products x 198 flop each.
FORCE_FLOP_PER_SITE = 4 * 85 * 198
# Fermion force + momentum kick per site (compulsory bytes): U read (72
# reals), X and Y read (6 + 6), P read and written (72 + 72), FP64.
FERMION_FORCE_BYTES_PER_SITE = (72 + 6 + 6 + 72 + 72) * 8


def bench_latency(args, dev):
    """Kernel times at small lattices (8^4 to 24^4): the plain hop with and
    without its grid reduction, FP32 and FP64, and the two kernels of one
    FP32-inner mixed iteration in both forms (fused update + hop, split), each
    the average of 100 back-to-back launches between HIP events.  Where the
    times barely change with the lattice, a launch is fixed cost, not data."""
    ctx = dev.Context(0)
    sizes = {}
    for L in (8, 12, 16, 24):
        dims = (L,) * 4
        g = dev.Gauge(ctx, dims, dev.random_links(dims, 0.3, seed=1))
        r = {"sites": L ** 4, "waves": L ** 4 // 64}
        for prec in ("f32", "f64"):
            r[f"hop_{prec}_us"] = round(g.time_hop(prec, False, 100), 2)
            r[f"hop_dot_{prec}_us"] = round(g.time_hop(prec, True, 100), 2)
        # the four-lanes-per-site hop form (four times the waves)
        for prec in ("f32", "f16", "f64"):
            dev.set_hop_variant(prec, 1)
            r[f"hop4_{prec}_us"] = round(g.time_hop(prec, False, 100), 2)
            dev.set_hop_variant(prec, 0)
            r[f"hop1_{prec}_us"] = round(g.time_hop(prec, False, 100), 2)
            dev.set_hop_variant(prec, -1)
        for form, name in ((0, "fused"), (1, "split")):
            dev.set_cg_form("f32", form)
            a, b = g.time_cg_kernels("f32", 100)
            r[f"mix_f32_{name}_us"] = [round(a, 2), round(b, 2)]
        dev.set_cg_form("f32", -1)
        sizes[f"{L}^4"] = r
        log(f"[bench] latency {L}^4: {r}")
    out = {"metric": "kernel time at small lattices (us per launch)", "value": sizes["16^4"]["hop_f32_us"],
           "unit": "us (FP32 hop, 16^4)", "n_gpus": 1, "steps": 100, "warmup": 1,
           "ms_per_step": round(sizes["16^4"]["hop_f32_us"] / 1e3, 5), "higher_is_better": False,
           "scaling": "weak", "vs_baseline": None, "dtype": "f32 / f64", "data": "synthetic (seeded Philox)",
           "config": {"workload": "hop and mixed-iteration kernels at 8^4, 12^4, 16^4, 24^4, links exp(i 0.3 H)",
                      "parallelism": "single GPU"},
           "roofline": None, "cpu_baseline": None, "sizes": sizes}
    print(json.dumps(out))


def bench_force(args, dev):
    """The force kernel as the leapfrog runs it (momentum kick fused) on 32^4
    links exp(i 0.3 H), 16^4 beside it: the Symanzik gauge force (c11 = 5/3,
    c12 = -1/12, beta = 4.0) against the FP64 VALU peak, the fermion force
    against HBM, and both together (the leapfrog's kick), each the average of
    `steps` back-to-back launches between HIP events on the library stream."""
    ctx = dev.Context(0)
    n = max(args.steps, 3)
    res = {}
    for dims in [(32, 32, 32, 32), (16, 16, 16, 16)]:
        V = int(np.prod(dims))
        g = dev.Gauge(ctx, dims, dev.random_links(dims, 0.3, seed=1234))
        if args.kernels_only:
            for which in (1, 2, 3):
                g.time_force(which=which, n=2)
            ctx.synchronize()
            log(f"[bench] force kernels done {dims}")
            continue
        us = {w: g.time_force(which=w, n=n) for w in (1, 2, 3)}
        res["x".join(map(str, dims))] = {
            "gauge_ms": round(us[1] * 1e-3, 4),
            "gauge_fp64_tflops": round(V * FORCE_FLOP_PER_SITE / (us[1] * 1e-6) / 1e12, 2),
            "fermion_ms": round(us[2] * 1e-3, 4),
            "fermion_tb_s": round(V * FERMION_FORCE_BYTES_PER_SITE / (us[2] * 1e-6) / 1e12, 3),
            "both_ms": round(us[3] * 1e-3, 4)}
        log(f"[bench] force {dims}: {res['x'.join(map(str, dims))]}")
    if args.kernels_only:
        return
    top = res["32x32x32x32"]
    out = _base_line(
        round(1e3 / top["both_ms"], 3), "force evaluations/s", n, top["both_ms"], "f64",
        "32^4 force evaluation as the leapfrog runs it: Symanzik gauge force (5/3, -1/12), beta 4.0, plus the "
        "staggered fermion force, momentum kick fused, links exp(i 0.3 H)",
        {"metric": "force evaluations/s (32^4, Symanzik + fermion force)", "sizes": res,
         "roofline": {"bound": "mfma", "achieved": top["gauge_fp64_tflops"], "peak": PEAK_FP64_TFS,
                      "unit": "TFLOP/s", "frac": round(top["gauge_fp64_tflops"] / PEAK_FP64_TFS, 4),
                      "traffic": traffic_from_profiles("force_gauge", "32^4"),
                      "kernel": "k_force, gauge part (85 x 198 flop per link); bound: FP64 vector ALU (no FP64 "
                                "matrix-core advantage on this part)"},
         "fermion_force_roofline": {"bound": "hbm", "achieved": round(top["fermion_tb_s"] * 1e3, 1),
                                    "peak": PEAK_HBM_GBS, "unit": "GB/s",
                                    "frac": round(top["fermion_tb_s"] * 1e3 / PEAK_HBM_GBS, 4),
                                    "traffic": traffic_from_profiles("force_fermion", "32^4")}})
    print(json.dumps(out))


# One link update of the heatbath: 12 plaquette and 72 rectangle staple
# products, about 3 more for the three SU(2) subgroup rotations (the 2 x 2
# block of U A and two rotated rows each): 87 SU(3) products x 198 flop.
HEATBATH_FLOP_PER_LINK = 87 * 198
PEAK_FP64_TFS = 78.6


def bench_heatbath(args, dev):
    """Cabibbo-Marinari heatbath sweeps of the Symanzik action (c11 = 5/3,
    c12 = -1/12, beta = 4.1) on 32^4 device-resident links from a hot start
    (links exp(i 0.6 H)), after `warmup` untimed sweeps; 16^4 beside it.  The
    roofline is the FP64 rate of the sweep's 16 class launches, timed with
    HIP events on the library stream."""
    beta, c11, c12 = 4.1, 5.0 / 3.0, -1.0 / 12.0
    steps = max(args.steps, 4)
    ctx = dev.Context(0)
    if args.kernels_only:
        # two sweeps at 32^4, then one FP32 link conversion (reads 576 B and
        # writes 288 B per site) that calibrates the traffic counters
        dims = (32, 32, 32, 32)
        g = dev.Gauge(ctx, dims, dev.random_links(dims, 0.6, seed=8))
        g.heatbath(beta, c11, c12, seed=1, sweep0=0, nsweeps=2)
        g.time_hop("f32", False, 1)
        ctx.synchronize()
        log("[bench] heatbath kernels done")
        return
    res = {}
    for dims in [(32, 32, 32, 32), (16, 16, 16, 16)]:
        V = int(np.prod(dims))
        g = dev.Gauge(ctx, dims, dev.random_links(dims, 0.6, seed=8))
        g.heatbath(beta, c11, c12, seed=1, sweep0=0, nsweeps=max(args.warmup, 1))
        ctx.synchronize()
        t0 = time.perf_counter()
        pl, ac = g.heatbath(beta, c11, c12, seed=1, sweep0=100, nsweeps=steps)
        dt = (time.perf_counter() - t0) / steps
        ctx.set_profiling(True)
        g.heatbath(beta, c11, c12, seed=1, sweep0=200, nsweeps=steps)
        ms, n = ctx.kernel_time("heatbath")
        ctx.set_profiling(False)
        per_launch_us = ms * 1e3 / (16 * n)
        tfs = (V / 4) * HEATBATH_FLOP_PER_LINK / (per_launch_us * 1e-6) / 1e12
        res["x".join(map(str, dims))] = {"ms_per_sweep": round(dt * 1e3, 3), "sweeps_per_s": round(1 / dt, 2),
                                         "class_launch_us": round(per_launch_us, 2), "fp64_tflops": round(tfs, 2),
                                         "plaquette": float(pl[-1]), "action_density": float(ac[-1])}
        log(f"[bench] heatbath {dims}: {dt * 1e3:.3f} ms/sweep, {per_launch_us:.1f} us per class launch, "
            f"{tfs:.2f} TF/s")
    top = res["32x32x32x32"]
    cpu = None
    if not args.no_cpu:
        from oracle import oracle as oc
        threads = min(16, os.cpu_count() or 1)
        oc.set_threads(threads)
        dims = (32, 32, 32, 32)
        U = dev.random_links(dims, 0.6, seed=8)
        t0 = time.perf_counter()
        oc.heatbath(dims, U, beta, c11, c12, seed=1, sweep0=0, nsweeps=1, order=1)
        tc = time.perf_counter() - t0
        cpu = {"value": round(1.0 / tc, 4), "unit": "sweeps/s", "cores": threads, "kind": "port",
               "sample": f"one colour-ordered oracle sweep at 32^4 ({tc:.1f} s, OpenMP over each colour class)"}
    traffic = traffic_from_profiles("heatbath", "32^4")
    out = _base_line(
        top["sweeps_per_s"], "sweeps/s", steps, top["ms_per_sweep"], "f64",
        "32^4 Cabibbo-Marinari heatbath sweep, Symanzik (5/3, -1/12), beta 4.1, hot start",
        {"metric": "heatbath sweeps/s (Symanzik action, 32^4)", "sizes": res,
         "roofline": {"bound": "mfma", "achieved": top["fp64_tflops"], "peak": PEAK_FP64_TFS, "unit": "TFLOP/s",
                      "frac": round(top["fp64_tflops"] / PEAK_FP64_TFS, 4), "traffic": traffic,
                      "kernel": "k_heatbath (one colour class, 262144 links, 87 x 198 flop each)"},
         "cpu_baseline": cpu})
    print(json.dumps(out))


# One HEX step per site: level 1 makes 12 links (4 staple products, the
# stout map G U^dagger, Q^2, Q^3 and exp(iQ) U: 8 SU(3) products each),
# level 2 makes 12 (8 staple products + 4), level 3 makes 4 (12 + 4).
HEX_FLOP_PER_SITE = (12 * 8 + 12 * 12 + 4 * 16) * 198
# Compulsory bytes per site: level 1 reads U and writes 12 level-1 links,
# level 2 reads those and U and writes 12 level-2 links, level 3 reads those
# and U and writes U.
HEX_BYTES_PER_SITE = 576 + 1728 + 1728 + 576 + 1728 + 1728 + 576 + 576


def bench_hex(args, dev):
    """HEX smearing steps (Smearing.cpp:304-311, alpha 0.95, 0.76, 0.38) of
    32^4 device-resident links (links exp(i 0.6 H)), after `warmup` untimed
    steps; 16^4 beside it.  The roofline is the FP64 rate of the timed steps
    (three level launches each, back to back on the library stream)."""
    alpha = (0.95, 0.76, 0.38)
    steps = max(args.steps, 4)
    ctx = dev.Context(0)
    if args.kernels_only:
        dims = (32, 32, 32, 32)
        g = dev.Gauge(ctx, dims, dev.random_links(dims, 0.6, seed=8))
        g.hex_smear(alpha, 2)
        g.time_hop("f32", False, 1)
        ctx.synchronize()
        log("[bench] hex kernels done")
        return
    res = {}
    for dims in [(32, 32, 32, 32), (16, 16, 16, 16)]:
        V = int(np.prod(dims))
        U0 = dev.random_links(dims, 0.6, seed=8)
        g = dev.Gauge(ctx, dims, U0)
        g.hex_smear(alpha, max(args.warmup, 1))
        ctx.synchronize()
        t0 = time.perf_counter()
        g.hex_smear(alpha, steps)
        dt = (time.perf_counter() - t0) / steps
        # per-level device times from the same steps on the same starting
        # links (each call checks the links: after many steps the rounding
        # leaves them further from SU(3) than the two-row threshold, and the
        # full-row kernels run)
        gp = dev.Gauge(ctx, dims, U0)
        gp.hex_smear(alpha, max(args.warmup, 1))
        ctx.set_profiling(True)
        gp.hex_smear(alpha, steps)
        lv = {k: ctx.kernel_time(k) for k in ("hex1", "hex2", "hex3")}
        ctx.set_profiling(False)
        us = {k: ms * 1e3 / n for k, (ms, n) in lv.items()}
        # the rate over the timed steps (the three level launches back to back)
        step_us = dt * 1e6
        tfs = V * HEX_FLOP_PER_SITE / (step_us * 1e-6) / 1e12
        res["x".join(map(str, dims))] = {
            "ms_per_step": round(dt * 1e3, 3), "steps_per_s": round(1 / dt, 2),
            "level_us": {k: round(v, 1) for k, v in us.items()}, "fp64_tflops": round(tfs, 2),
            "compulsory_tb_s": round(V * HEX_BYTES_PER_SITE / (step_us * 1e-6) / 1e12, 2),
            "avg_plaquette": g.avg_plaq()}
        log(f"[bench] hex {dims}: {dt * 1e3:.3f} ms/step, levels {us}, {tfs:.2f} TF/s")
    top = res["32x32x32x32"]
    cpu = None
    if not args.no_cpu:
        from oracle import oracle as oc
        threads = min(16, os.cpu_count() or 1)
        oc.set_threads(threads)
        dims = (32, 32, 32, 32)
        U = dev.random_links(dims, 0.6, seed=8)
        t0 = time.perf_counter()
        oc.hex_smear(dims, U, alpha=alpha, nsteps=1)
        tc = time.perf_counter() - t0
        cpu = {"value": round(1.0 / tc, 4), "unit": "steps/s", "cores": threads, "kind": "port",
               "sample": f"one oracle HEX step at 32^4 ({tc:.1f} s, OpenMP over sites)"}
    out = _base_line(
        top["steps_per_s"], "HEX steps/s", steps, top["ms_per_step"], "f64",
        "32^4 HEX smearing step (alpha 0.95, 0.76, 0.38), links exp(i 0.6 H)",
        {"metric": "HEX smearing steps/s (32^4)", "sizes": res,
         "roofline": {"bound": "mfma", "achieved": top["fp64_tflops"], "peak": PEAK_FP64_TFS, "unit": "TFLOP/s",
                      "frac": round(top["fp64_tflops"] / PEAK_FP64_TFS, 4),
                      "traffic": traffic_from_profiles("hex_step", "32^4"),
                      "kernel": "k_hex1 + k_hex2 + k_hex3 (one step, 304 x 198 flop per site)"},
         "cpu_baseline": cpu})
    print(json.dumps(out))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="32^4",
                    choices=sorted(WORKLOADS) + ["batch", "serve", "precision", "traj32", "heatbath", "hex", "force",
                                                          "latency"])
    ap.add_argument("--rates", default="0.5,1,2", help="serve: arrival rates as multiples of the measured "
                                                     "sequential FP64 capacity")
    ap.add_argument("--requests", type=int, default=200, help="serve: requests per arrival rate")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-graphs", action="store_true",
                    help="solve workloads: launch each CG chunk's kernels directly instead of replaying HIP graphs "
                         "(the same kernels; for kernel tracers)")
    ap.add_argument("--legs", default="both", choices=["both", "mixed", "f64"],
                    help="solve workloads: time the mixed-precision leg, the FP64 leg or both (one leg prints its "
                         "timings to stderr instead of the JSON line, e.g. for a kernel trace of one leg)")
    ap.add_argument("--kernels-only", action="store_true",
                    help="only the kernel timing loops of the workload (the hop in FP64 and in the inner precision, "
                         "the two steps of the mixed iteration), e.g. under a hardware-counter profiler; the first "
                         "use of the low-precision link copy also runs its conversion kernel once (reads 576 B, "
                         "writes 288 B (FP32) or 144 B (FP16) per site); for the heatbath workload, two 32^4 sweeps "
                         "and one FP32 link conversion; for hex, two 32^4 steps and one FP32 link conversion")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "runs on one GPU only; --gpus must be 1 (the lattice would shard by "
                                   "4-D domain decomposition with halo exchange, which is not implemented)"}))
        sys.exit(2)

    from staggered_fermions_hmc_amd import device as dev

    if args.workload == "batch":
        return bench_batch(args, dev)
    if args.workload == "serve":
        return bench_serve(args, dev)
    if args.workload == "precision":
        return bench_precision(args, dev)
    if args.workload == "traj32":
        return bench_traj32(args, dev)
    if args.workload == "heatbath":
        return bench_heatbath(args, dev)
    if args.workload == "hex":
        return bench_hex(args, dev)
    if args.workload == "force":
        return bench_force(args, dev)
    if args.workload == "latency":
        return bench_latency(args, dev)

    wl = WORKLOADS[args.workload]
    dims, mass = wl["dims"], wl["mass"]
    V = int(np.prod(dims))
    log(f"[bench] workload {args.workload}: generating inputs")
    U = dev.random_links(dims, wl["eps"], seed=1234)
    b = dev.random_spinor(dims, seed=5678)

    ctx = dev.Context(0)
    if args.no_graphs:
        ctx.set_graph_replay(False)
    g = dev.Gauge(ctx, dims, U)
    if args.kernels_only:
        out = {"workload": args.workload, "lattice": "x".join(map(str, dims))}
        for prec in (wl["mixed"], "f64"):
            for dot in (False, True):
                us = g.time_hop(prec, with_dot=dot, n=20)
                out[f"hop{'_dot' if dot else ''}_{prec}_us"] = round(us, 2)
        ua, ub = g.time_cg_kernels(wl["mixed"], n=20)
        out[f"mix_a_{wl['mixed']}_us"], out[f"mix_b_{wl['mixed']}_us"] = round(ua, 2), round(ub, 2)
        print(json.dumps(out))
        return
    phi = dev.Field(g, b)
    x = dev.Field(g, None)
    tol = 1e-10

    def solve(prec):
        t0 = time.perf_counter()
        # both legs stop on the FP64 true residual: mixed precision by its
        # reliable updates, FP64 by a true-residual check and restart once the
        # iterated residual meets the target
        st = g.cg_field(phi, x, mass, tol=tol, max_iter=100000, precision=prec,
                        true_residual="stop" if prec == "f64" else False)
        return time.perf_counter() - t0, st

    res = {}
    legs = {"both": (wl["mixed"], "f64"), "mixed": (wl["mixed"],), "f64": ("f64",)}[args.legs]
    for prec in legs:
        for _ in range(args.warmup):
            solve(prec)
        ctx.synchronize()
        lat, stats = [], None
        t0 = time.perf_counter()
        for _ in range(args.steps):
            dt, stats = solve(prec)
            lat.append(dt)
        ctx.synchronize()
        total = time.perf_counter() - t0
        # verify the last solution against an independent FP64 residual
        r = b - g.apply_ddag(mass, x.download())
        true_res = float(np.sqrt(r @ r / (b @ b)))
        # per-kernel device time inside one more solve, every launch bracketed
        # by HIP events on the library stream (direct launches instead of the
        # chunk graphs, so each kernel can be timed)
        ctx.set_profiling(1)
        solve(prec)
        names = ([f"hop_{prec}", f"hop_dot_{prec}", "cg_blas"] if prec == "f64" else
                 [f"mix_a_{prec}", "mix_upd", f"hop_{prec}", f"mix_b_{prec}", "reliable_update"])
        kt = {k: ctx.kernel_time(k) for k in names}
        ctx.set_profiling(0)
        # the same solve through the host-buffer entry point (sfh_cg: phi
        # uploaded and x downloaded over PCIe inside the call), outside the
        # timed region; reported beside the device-resident rate
        t1 = time.perf_counter()
        for _ in range(3):
            g.cg(b, mass, tol=tol, max_iter=100000, precision=prec,
                 true_residual="stop" if prec == "f64" else False)
        pcie_sps = 3 / (time.perf_counter() - t1)
        # the hop kernel alone: back-to-back launches between two HIP events
        hop_us = g.time_hop(prec, with_dot=False, n=50)
        hop_dot_us = g.time_hop(prec, with_dot=True, n=50)
        # the two steps of the mixed iteration as the solver runs them
        mix_us = g.time_cg_kernels(prec, n=50) if prec != "f64" else None
        res[prec] = dict(total=total, lat=lat, stats=stats, kt=kt, true_res=true_res, hop_us=hop_us,
                         hop_dot_us=hop_dot_us, mix_us=mix_us, pcie_sps=pcie_sps)
        log(f"[bench] {prec}: {args.steps} solves in {total:.3f}s, iters {stats['iterations']}, "
            f"reliable updates {stats['reliable_updates']}, true residual {true_res:.2e}, hop {hop_us:.2f} us")

    if args.legs != "both":
        r = res[legs[0]]
        log(json.dumps({"leg": legs[0], "solves_per_s": args.steps / r["total"], "iterations": r["stats"]["iterations"],
                        "true_rel_residual": r["true_res"], "hop_us": r["hop_us"], "hop_dot_us": r["hop_dot_us"],
                        "mix_us": r["mix_us"]}))
        return
    mixed = res[wl["mixed"]]
    f64 = res["f64"]
    value = args.steps / mixed["total"]

    def hop_gbs(prec, r):
        out = {"hop_avg_us": r["hop_us"], "hop_gbs": HOP_BYTES_PER_SITE[prec] * V / (r["hop_us"] * 1e-6) / 1e9,
               "hop_dot_avg_us": r["hop_dot_us"],
               "hop_dot_gbs": (HOP_BYTES_PER_SITE[prec] + DOT_EXTRA_PER_SITE[prec]) * V / (r["hop_dot_us"] * 1e-6)
               / 1e9}
        its = max(1, r["stats"]["iterations"])
        for k, (ms, n) in r["kt"].items():
            if n:
                out[f"in_solve_{k}_us_per_iter"] = ms * 1e3 / its
        return out

    hm = hop_gbs(wl["mixed"], mixed)
    h64 = hop_gbs("f64", f64)
    mp = wl["mixed"]

    def gbs(bytes_per_site, us):
        return bytes_per_site * V / (us * 1e-6) / 1e9

    # Dominant kernel of the timed (mixed) solve, by its time inside the
    # solve: step A of the iteration (first hop with the vector updates) when
    # it is one fused kernel and takes longer, else step B (the second hop
    # with its four dot products).  Both timed as the solver launches them.
    us_a, us_b = mixed["mix_us"]
    in_a = mixed["kt"].get(f"mix_a_{mp}", (0.0, 0))[0]
    in_b = mixed["kt"].get(f"mix_b_{mp}", (0.0, 0))[0]
    step = "a" if in_a > in_b else "b"
    kern = f"mix_{step}_{mp}"
    bps = (MIX_A_BYTES_PER_SITE if step == "a" else MIX_B_BYTES_PER_SITE)[mp]
    us_k = us_a if step == "a" else us_b
    achieved = gbs(bps, us_k)
    traffic = traffic_from_profiles(kern, args.workload)
    other = ("b", MIX_B_BYTES_PER_SITE[mp], us_b) if step == "a" else ("a", MIX_A_BYTES_PER_SITE[mp], us_a)
    roofline = {
        "bound": "hbm",
        "kernel": kern,
        "achieved": round(achieved, 1),
        "peak": PEAK_HBM_GBS,
        "unit": "GB/s",
        "frac": round(achieved / PEAK_HBM_GBS, 4),
        "traffic": traffic,
        "algorithmic_bytes_per_launch": bps * V,
        "avg_launch_us": round(us_k, 2),
        "in_solve_ms": {"mix_a": round(in_a, 3), "mix_b": round(in_b, 3)},
        f"step_{other[0]}": {"achieved": round(gbs(other[1], other[2]), 1),
                             "frac": round(gbs(other[1], other[2]) / PEAK_HBM_GBS, 4), "avg_us": round(other[2], 2),
                             "algorithmic_bytes": other[1] * V},
        "dslash": {"precision": mp, "achieved": round(hm["hop_gbs"], 1),
                   "frac": round(hm["hop_gbs"] / PEAK_HBM_GBS, 4), "avg_launch_us": round(hm["hop_avg_us"], 2),
                   "algorithmic_bytes_per_launch": HOP_BYTES_PER_SITE[mp] * V},
        "dslash_fp64": {"achieved": round(h64["hop_gbs"], 1), "frac": round(h64["hop_gbs"] / PEAK_HBM_GBS, 4),
                        "avg_launch_us": round(h64["hop_avg_us"], 2),
                        "algorithmic_bytes_per_launch": HOP_BYTES_PER_SITE["f64"] * V},
        "timing": "HIP events on the library stream around 50 back-to-back launches of each kernel as the solver "
                  "runs it at this lattice",
        "bytes_note": "algorithmic bytes count the full 18-real links (72 link reals per site); the links here are "
                      "SU(3), so every hop (FP64 included) reads two rows per link and rebuilds the third, and the "
                      "measured traffic (traffic) is below the algorithmic bytes",
    }

    cpu = None
    if not args.no_cpu:
        log("[bench] CPU baseline (oracle, bounded sample)")
        cpu = cpu_baseline(dims, mass, U, b, f64["stats"]["iterations"])

    lat_ms = np.array(mixed["lat"]) * 1e3
    out = {
        "metric": METRIC,
        "value": round(value, 4),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(mixed["total"] / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": f"{wl['mixed']} inner / f64 reliable updates and reductions",
        "data": "synthetic (seeded Philox links and Gaussian RHS)",
        "config": {"workload": wl["desc"], "lattice": "x".join(map(str, dims)), "mass": mass, "link_eps": wl["eps"],
                   "tol_true_residual": tol, "parallelism": "single GPU"},
        "p50_solve_ms": round(float(np.percentile(lat_ms, 50)), 3),
        "p99_solve_ms": round(float(np.percentile(lat_ms, 99)), 3),
        "mixed": {"precision": wl["mixed"], "solves_per_s": round(value, 4),
                  "iterations": mixed["stats"]["iterations"],
                  "reliable_updates": mixed["stats"]["reliable_updates"],
                  "true_rel_residual": mixed["true_res"],
                  "pcie_inclusive_solves_per_s": round(mixed["pcie_sps"], 3),
                  **{k: round(v, 2) for k, v in hm.items()}},
        "fp64": {"solves_per_s": round(args.steps / f64["total"], 4), "iterations": f64["stats"]["iterations"],
                 "true_rel_residual": f64["true_res"], "pcie_inclusive_solves_per_s": round(f64["pcie_sps"], 3),
                 **{k: round(v, 2) for k, v in h64.items()}},
        "speedup_mixed_vs_fp64": round(f64["total"] / mixed["total"], 3),
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
