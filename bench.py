"""Benchmark of the staggered CG hot path on one MI355X.

A step is one CG solve of (D D^dag) x = b to a true FP64 relative residual of
1e-10 on a 16^4 lattice (links U = exp(i 0.3 H), m = 0.05, one Gaussian
right-hand side), run in mixed precision (FP32 inner iterations, FP64
reliable update whenever the iterated residual drops 10x).  The plain FP64 CG
on the same inputs is timed beside it.  Inputs are resident in HBM before the
timed region starts.

Prints one JSON line: value = mixed-precision solves/s, the FP64 solves/s,
p50/p99 solve latency, the Dslash (hop) HBM roofline of the dominant kernel
measured with HIP events on the library stream over the timed region, and
the CPU baseline (the FP64 oracle, timed on a bounded sample on the host).

``--workload 32^4`` runs the 32^4, m = 0.01 solve with FP16 inner iterations
instead.
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


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(dims, mass, U, phi, iters_per_solve, budget_s=15.0):
    """FP64 oracle (same recurrence, AoS layout, OpenMP) on the host: times a
    bounded number of CG iterations and scales to one solve."""
    from oracle import oracle as oc
    threads = min(16, os.cpu_count() or 1)
    oc.set_threads(threads)
    t1 = oc.time_cg_iterations(dims, mass, U, phi, 2)  # warm + estimate
    n = int(max(3, min(iters_per_solve, budget_s / max(t1, 1e-6))))
    per_iter = oc.time_cg_iterations(dims, mass, U, phi, n)
    return {
        "value": 1.0 / (per_iter * iters_per_solve),
        "unit": "solves/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{n} FP64 CG iterations (D D^dag + 2 dots + 2 axpy) of the oracle at "
                  f"{'x'.join(map(str, dims))}, {per_iter * 1e3:.2f} ms/iteration, scaled to the "
                  f"{iters_per_solve} iterations of one FP64 solve",
    }


def traffic_from_profiles(kernel, workload):
    path = os.path.join(ROOT, "profiles", "hop_traffic.json")
    if not os.path.exists(path):
        return None
    try:
        data = json.load(open(path))
        return data.get(workload, {}).get(kernel)
    except Exception:
        return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="16^4", choices=sorted(WORKLOADS))
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 (multi-GPU sharding is a "
                                   "design only, see DESIGN.md)"}))
        sys.exit(2)

    from staggered_fermions_hmc_amd import device as dev

    wl = WORKLOADS[args.workload]
    dims, mass = wl["dims"], wl["mass"]
    V = int(np.prod(dims))
    log(f"[bench] workload {args.workload}: generating inputs")
    U = dev.random_links(dims, wl["eps"], seed=1234)
    b = dev.random_spinor(dims, seed=5678)

    ctx = dev.Context(0)
    g = dev.Gauge(ctx, dims, U)
    phi = dev.Field(g, b)
    x = dev.Field(g, None)
    tol = 1e-10

    def solve(prec):
        t0 = time.perf_counter()
        st = g.cg_field(phi, x, mass, tol=tol, max_iter=100000, precision=prec, true_residual=(prec == "f64"))
        return time.perf_counter() - t0, st

    res = {}
    for prec in (wl["mixed"], "f64"):
        for _ in range(args.warmup):
            solve(prec)
        ctx.synchronize()
        ctx.set_profiling(4)  # time one launch in 4 of each kernel family, no syncs added
        lat, stats = [], None
        ctx.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            dt, stats = solve(prec)
            lat.append(dt)
        ctx.synchronize()
        total = time.perf_counter() - t0
        kt = {k: ctx.kernel_time(f"{k}_{prec}") for k in ("hop", "hop_dot")}
        blas = ctx.kernel_time("cg_blas")
        ctx.set_profiling(0)
        # verify the last solution against an independent FP64 residual
        r = b - g.apply_ddag(mass, x.download())
        true_res = float(np.sqrt(r @ r / (b @ b)))
        res[prec] = dict(total=total, lat=lat, stats=stats, kt=kt, blas=blas, true_res=true_res)
        log(f"[bench] {prec}: {args.steps} solves in {total:.3f}s, iters {stats['iterations']}, "
            f"reliable updates {stats['reliable_updates']}, true residual {true_res:.2e}")

    mixed = res[wl["mixed"]]
    f64 = res["f64"]
    value = args.steps / mixed["total"]

    def hop_gbs(prec, r):
        ms, n = r["kt"]["hop"]
        msd, nd = r["kt"]["hop_dot"]
        out = {}
        if n:
            avg = ms / n * 1e-3
            out["hop_avg_us"] = avg * 1e6
            out["hop_gbs"] = HOP_BYTES_PER_SITE[prec] * V / avg / 1e9
        if nd:
            avgd = msd / nd * 1e-3
            out["hop_dot_avg_us"] = avgd * 1e6
            out["hop_dot_gbs"] = (HOP_BYTES_PER_SITE[prec] + DOT_EXTRA_PER_SITE[prec]) * V / avgd / 1e9
        return out

    hm = hop_gbs(wl["mixed"], mixed)
    h64 = hop_gbs("f64", f64)
    kern = f"hop_{wl['mixed']}"
    traffic = traffic_from_profiles(kern, args.workload)
    roofline = {
        "bound": "hbm",
        "kernel": kern,
        "achieved": round(hm.get("hop_gbs", 0.0), 1),
        "peak": PEAK_HBM_GBS,
        "unit": "GB/s",
        "frac": round(hm.get("hop_gbs", 0.0) / PEAK_HBM_GBS, 4),
        "traffic": traffic,
        "algorithmic_bytes_per_launch": HOP_BYTES_PER_SITE[wl["mixed"]] * V,
        "avg_launch_us": round(hm.get("hop_avg_us", 0.0), 2),
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
                  "true_rel_residual": mixed["true_res"], **{k: round(v, 2) for k, v in hm.items()}},
        "fp64": {"solves_per_s": round(args.steps / f64["total"], 4), "iterations": f64["stats"]["iterations"],
                 "true_rel_residual": f64["true_res"], **{k: round(v, 2) for k, v in h64.items()}},
        "speedup_mixed_vs_fp64": round(f64["total"] / mixed["total"], 3),
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
