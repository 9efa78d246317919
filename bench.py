"""Throughput of the batched explicit RVE solve on one MI355X.

A step is one batched solve of the whole workload: 4096 RVEs of 500 nodes
(k=4 nearest-neighbour fibers, linear truss reaction r=3.5e-8, E=4.32e7,
rho=1000), each under its own F = I + sym(U), U_ij ~ U(-0.05, 0.05), run for
5000 explicit steps with the boundary ramped linearly.  The headline is the
FP32 + compensated solver; the plain FP64 solver runs the same workload for
the speedup and as the error reference.  The CPU baseline is the FP64 oracle
on a bounded sample of the same workload.

Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "RVE solves/s at ≤1e-4 rel stress err vs FP64; fraction of BW roofline"
WORKLOAD = dict(n_rve=4096, n_nodes=500, k=4, explicit_steps=5000, F_amp=0.05)
# MI355X LDS read bandwidth from the data sheet: 256 CUs x 128 B/clk x 2.4 GHz
LDS_SPEC = 256 * 128 * 2.4e9
# HBM bytes per launch from rocprofv3 FETCH_SIZE / WRITE_SIZE runs of this
# bench (scaled by a plain copy kernel of known size), see profiles/
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic.json")


def measured_traffic(solver):
    try:
        with open(TRAFFIC_FILE) as f:
            t = json.load(f).get(solver)
        return None if t is None else int(t["bytes_per_launch"])
    except (OSError, ValueError, KeyError):
        return None


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def strain_F(seed, n, amp):
    import mumfim_amd as mf
    u = mf.uniform(seed, 77, 9 * n).reshape(n, 3, 3)
    U = (2 * u - 1) * amp
    return np.eye(3)[None] + 0.5 * (U + U.transpose(0, 2, 1))


def algorithmic_bytes(g, prec):
    """Compulsory state bytes per RVE-step summed over the batch:
    FP64 108*Nf + 24*Nb + 24*M, FP32c 80*Nf + 24*Nb + 20*M."""
    st = g.stats()
    Nf, Nb, M = st["Nf"], st["Nb"], st["M"]
    if prec == "fp64":
        return float((108 * Nf + 24 * Nb + 24 * M).sum())
    return float((80 * Nf + 24 * Nb + 20 * M).sum())


def time_solver(g, F_dev, sig_dev, steps, warmup):
    import mumfim_amd as mf
    for _ in range(warmup):
        g.run_device(F_dev, sig_dev)
        g.synchronize()
    kern_ms = []
    mf.device_synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        g.run_device(F_dev, sig_dev)
        g.synchronize()
        kern_ms.append(g.last_solve_ms()[0])
    mf.device_synchronize()
    t1 = time.perf_counter()
    return (t1 - t0), kern_ms


def cpu_baseline(nets_fn, F, n_steps, target_s=12.0):
    """FP64 oracle on a bounded sample of the workload (OpenMP over RVEs)."""
    import mumfim_amd as mf
    from oracle.oracle import Oracle
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or len(os.sched_getaffinity(0))
    threads = max(1, min(threads, 16))
    os.environ["OMP_NUM_THREADS"] = str(threads)
    p = mf.default_params(mode=mf.MF_MODE_RAMP, n_steps=n_steps)
    # calibrate with one RVE on one thread-slot
    o1 = Oracle(nets_fn(1), p)
    t = time.perf_counter()
    o1.run(F[:1], update_coords=False)
    t1 = time.perf_counter() - t
    n = int(max(threads, min(4096, round(target_s / max(t1, 1e-3) * threads))))
    n = max(threads, (n // threads) * threads)
    o = Oracle(nets_fn(n), p)
    t = time.perf_counter()
    # trial runs (update_coords=False) like the timed GPU batches: sigma / V0
    s = o.run(F[:n], update_coords=False)
    dt = time.perf_counter() - t
    return n / dt, n, threads, dt, s


def rel_err(a, b):
    return np.abs(a - b).max(1) / np.abs(b).max(1)


def bench_cfg3(args):
    """16384 RVEs, N log-uniform in [200, 4000], half the fibers on the
    nonlinear toe reaction, F = I + sym(U) with U ~ U(-0.1, 0.1), 10000 ramp
    steps, FP32c in one batch (all kernel-variant groups launched together);
    FP64 on a 256-RVE subset for the error."""
    import mumfim_amd as mf
    from oracle.oracle import Oracle
    R = 16384 if args.n_rve == WORKLOAD["n_rve"] else args.n_rve
    S = 10000 if args.explicit_steps == WORKLOAD["explicit_steps"] else args.explicit_steps
    u = mf.uniform(3, 11, R)
    sizes = np.rint(np.exp(np.log(200) + u * (np.log(4000) - np.log(200)))).astype(np.int32)
    F = strain_F(3, R, 0.1)
    gen = mf.default_gen_params(seed=100001, nonlinear_fraction=0.5)
    base = dict(mode=mf.MF_MODE_RAMP, n_steps=S, reaction_policy=mf.MF_REACT_PER_FIBER)
    t = time.perf_counter()
    g = mf.BatchedFiberRVEAnalysisExplicit(
        params=mf.default_params(precision=mf.MF_PREC_FP32C, **base),
        generated=(gen, R), n_nodes=sizes)
    log(f"[cfg3] batch built in {time.perf_counter() - t:.1f}s, plan groups {g.plan()[0]}")
    F_dev = mf.DeviceArray.from_host(F.reshape(R, 9))
    sig = mf.DeviceArray((R, 6))
    wall, kms = time_solver(g, F_dev, sig, args.steps, args.warmup)
    s32 = sig.numpy()
    st = g.stats()
    nsub = 256
    sub = np.arange(0, R, R // nsub)[:nsub]
    nets = [mf.generate_network(seed=100001 + int(r), n_nodes=int(sizes[r]),
                                nonlinear_fraction=0.5) for r in sub]
    g64 = mf.BatchedFiberRVEAnalysisExplicit(
        nets, mf.default_params(precision=mf.MF_PREC_FP64, **base))
    s64 = g64.run(F[sub], update_coords=False)
    e = rel_err(s32[sub], s64)
    res = {"config": "cfg3", "workload": f"{R} RVEs, nodes log-uniform [200,4000] "
           f"(total {int(st['N'].sum())} nodes, {int(st['M'].sum())} fibers), 50% nonlinear "
           f"fibers, F=I+sym(U) U~U(-0.1,0.1), {S} ramp steps",
           "value": round(R * args.steps / wall, 3), "unit": "RVE solves/s", "dtype": "f32",
           "ms_per_batch": round(1e3 * wall / args.steps, 3),
           "kernel_ms": round(float(np.mean(kms)), 3), "launch_groups": g.plan()[0],
           "rel_stress_err_vs_fp64_subset": {"n": nsub, "max": float(e.max()),
                                              "median": float(np.median(e))}}
    if not args.no_cpu:
        o = Oracle(nets[:32], mf.default_params(precision=mf.MF_PREC_FP64, **base))
        so = o.run(F[sub[:32]], update_coords=False)
        res["fp64_gpu_vs_oracle_max"] = float(rel_err(s64[:32], so).max())
    print(json.dumps(res), flush=True)


def bench_cfg4(args):
    """4096 RVEs of 1000-2000 nodes, mixed reactions; per RVE the base solve
    plus 12 central perturbations (h = 1e-6) in one batched FP64 launch
    (53248 solves), FD tangent in FP64; tangent error vs the oracle on a
    sample."""
    import mumfim_amd as mf
    from oracle.oracle import Oracle
    R = args.n_rve
    S = args.explicit_steps
    u = mf.uniform(4, 11, R)
    sizes = (1000 + np.floor(u * 1001)).astype(np.int32)
    F = strain_F(4, R, 0.05)
    gen = mf.default_gen_params(seed=200001, nonlinear_fraction=0.5)
    p = mf.default_params(mode=mf.MF_MODE_RAMP, n_steps=S,
                          reaction_policy=mf.MF_REACT_PER_FIBER,
                          tangent_scheme=mf.MF_TANGENT_CENTRAL, fd_step=1e-6)
    g = mf.BatchedFiberRVEAnalysisExplicit(params=p, generated=(gen, R), n_nodes=sizes)
    times = []
    first = None
    for k in range(max(1, args.warmup) + args.steps):
        t = time.perf_counter()
        sigma, Cm = g.run_with_tangent(F)  # commits: later batches start deformed
        dt = time.perf_counter() - t
        if first is None:
            first = (sigma, Cm)
        elif len(times) < args.steps:
            times.append(dt)
    sigma, Cm = first
    wall = float(np.mean(times))
    res = {"config": "cfg4", "workload": f"{R} RVEs, nodes U[1000,2000], 50% nonlinear fibers, "
           f"base + 12 central perturbations (h=1e-6) per RVE = {13 * R} solves of {S} ramp "
           f"steps, FP64 solves + FP64 differencing",
           "value": round(13 * R / wall, 3), "unit": "RVE solves/s (incl. perturbed)",
           "tangents_per_s": round(R / wall, 3), "ms_per_batch": round(1e3 * wall, 3),
           "dtype": "f64"}
    if not args.no_cpu:
        n = 16
        nets = [mf.generate_network(seed=200001 + r, n_nodes=int(sizes[r]),
                                    nonlinear_fraction=0.5) for r in range(n)]
        o = Oracle(nets, p)
        so = o.run(F[:n])
        Co = o.computeMaterialStiffness()
        res["sigma_err_vs_oracle_max"] = float(rel_err(sigma[:n], so).max())
        res["tangent_err_vs_oracle_max"] = float(
            rel_err(Cm[:n].reshape(n, 36), Co.reshape(n, 36)).max())
    print(json.dumps(res), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n-rve", type=int, default=WORKLOAD["n_rve"])
    ap.add_argument("--explicit-steps", type=int, default=WORKLOAD["explicit_steps"])
    ap.add_argument("--fp64-steps", type=int, default=None,
                    help="timed batches of the FP64 solver (default: --steps)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--only", choices=["fp32c", "fp64"], default=None,
                    help="time one solver only (profiling runs)")
    ap.add_argument("--config", choices=["cfg2", "cfg3", "cfg4"], default="cfg2",
                    help="cfg2 (default): 4096 x 500 nodes; cfg3: 16384 RVEs of 200-4000 "
                         "nodes, mixed reactions, 10000 steps; cfg4: 4096 RVEs of "
                         "1000-2000 nodes, base + 12 perturbed solves per RVE")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not "
              "supported (RVEs shard by RVE across GPUs, see DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    import mumfim_amd as mf

    if args.config == "cfg3":
        return bench_cfg3(args)
    if args.config == "cfg4":
        return bench_cfg4(args)
    R, S = args.n_rve, args.explicit_steps
    gen = mf.default_gen_params(seed=1, n_nodes=WORKLOAD["n_nodes"])
    F = strain_F(2, R, WORKLOAD["F_amp"])
    # inputs resident in HBM, allocated by the library's own HIP runtime
    F_dev = mf.DeviceArray.from_host(F.reshape(R, 9))
    out = {}
    solvers = {}
    for name, prec in (("fp32c", mf.MF_PREC_FP32C), ("fp64", mf.MF_PREC_FP64)):
        if args.only and args.only != name:
            continue
        t = time.perf_counter()
        p = mf.default_params(mode=mf.MF_MODE_RAMP, n_steps=S, precision=prec)
        g = mf.BatchedFiberRVEAnalysisExplicit(params=p, generated=(gen, R))
        log(f"[bench] {name}: batch of {R} RVEs built in {time.perf_counter() - t:.1f}s")
        sig = mf.DeviceArray((R, 6), np.float64)
        steps = args.steps if name == "fp32c" else (args.fp64_steps or args.steps)
        wall, kms = time_solver(g, F_dev, sig, steps, args.warmup)
        log(f"[bench] {name}: {steps} batches in {wall:.3f}s, kernel ms {kms}")
        solvers[name] = g
        out[name] = dict(wall=wall, steps=steps, kernel_ms=kms,
                         sigma=sig.numpy(),
                         bytes=algorithmic_bytes(g, name))
    head = "fp32c" if "fp32c" in out else "fp64"
    h = out[head]
    value = R * h["steps"] / h["wall"]
    ms_per_step = 1e3 * h["wall"] / h["steps"]
    kern_ms = float(np.mean(h["kernel_ms"]))
    lds_meas = mf.measure_lds_bandwidth()
    hbm_meas = mf.measure_hbm_bandwidth(1 << 30)
    achieved = h["bytes"] * S / (kern_ms * 1e-3)
    res = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "RVE solves/s",
        "n_gpus": 1,
        "steps": h["steps"],
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if head == "fp32c" else "f64",
        "data": "synthetic",
        "config": {
            "workload": f"{R} RVEs x {WORKLOAD['n_nodes']} nodes (k=4 NN fibers, linear truss), "
                        f"F=I+sym(U) U~U(-0.05,0.05), {S} ramp steps per solve",
            "n_rve": R, "n_nodes": WORKLOAD["n_nodes"], "explicit_steps": S,
            "solver": "fp32 forces/updates + compensated (hi,lo) fp32 positions + fp64 stress"
                      if head == "fp32c" else "fp64 reference arithmetic",
            "parallelism": "1 GPU, one workgroup per RVE solve",
        },
        "roofline": {
            "bound": "lds",
            "achieved": round(achieved / 1e9, 1),
            "peak": round(lds_meas / 1e9, 1),
            "unit": "GB/s",
            "frac": round(achieved / lds_meas, 4),
            "traffic": measured_traffic(head),
            "peak_spec": round(LDS_SPEC / 1e9, 1),
            "frac_of_spec": round(achieved / LDS_SPEC, 4),
            "kernel_ms": round(kern_ms, 3),
            "algorithmic_bytes_per_launch": h["bytes"] * S,
            "hbm_copy_measured_GBps": round(hbm_meas / 1e9, 1),
        },
    }
    if "fp64" in out and "fp32c" in out:
        s64, s32 = out["fp64"]["sigma"], out["fp32c"]["sigma"]
        err = np.abs(s32 - s64).max(1) / np.abs(s64).max(1)
        v64 = R * out["fp64"]["steps"] / out["fp64"]["wall"]
        res["fp64"] = {"value": round(v64, 3), "ms_per_step":
                       round(1e3 * out["fp64"]["wall"] / out["fp64"]["steps"], 3),
                       "kernel_ms": round(float(np.mean(out["fp64"]["kernel_ms"])), 3)}
        res["speedup_vs_fp64"] = round(value / v64, 3)
        res["rel_stress_err_vs_fp64"] = {"max": float(err.max()), "median": float(np.median(err))}
    if not args.no_cpu:
        t = time.perf_counter()
        nets_fn = lambda n: [mf.generate_network(seed=1 + r, n_nodes=WORKLOAD["n_nodes"])
                             for r in range(n)]
        cv, n, threads, dt, s_cpu = cpu_baseline(nets_fn, F, S)
        ref = out.get("fp64", out[head])["sigma"][:n]
        perr = float((np.abs(ref - s_cpu).max(1) / np.abs(s_cpu).max(1)).max())
        res["cpu_baseline"] = {
            "value": round(cv, 3), "unit": "RVE solves/s", "cores": threads, "kind": "port",
            "sample": f"first {n} RVEs of the workload ({S} steps each), FP64 oracle, "
                      f"OpenMP {threads} threads, {dt:.1f}s",
            "fp64_gpu_vs_cpu_max_rel_err": perr,
        }
        log(f"[bench] cpu baseline {cv:.2f} solves/s on {threads} threads "
            f"({time.perf_counter() - t:.1f}s)")
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
