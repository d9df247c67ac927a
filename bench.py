#!/usr/bin/env python3
"""Throughput of the coupled two-phase LBM step on one MI355X.

Workload (default): 3D rising bubble, D3Q19 for both distributions, 256^3
interior nodes (N = 258 with CLIP's in-N ghost ring), fully periodic, sphere
R = 32, rhoRatio 1000, muRatio 100, Eotvos 10, Morton 1e-3, interface width 4,
synthetic initial state from the case definition (no input files).  One step
= NSAllen::solve on every interior node (ghost refresh + phase sum + fused
collide/stream kernel).

Prints one JSON line: MLUPS (whole job), the roofline of the dominant kernel
(k_fused, HIP events on the solver's stream) and a CPU baseline (the FP64
OpenMP oracle on a bounded sample of the same workload).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
# thread placement of the OpenMP CPU baseline (read when its runtime starts)
os.environ.setdefault("OMP_PROC_BIND", "close")
os.environ.setdefault("OMP_PLACES", "cores")

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E


POP_STORAGE = {"f64": "f64", "f32": "f32", "mixed_f16": "f16", "mixed_bf16": "bf16"}


def pop_bytes(precision):
    return {"f64": 8, "f32": 4, "mixed_f16": 2, "mixed_bf16": 2}[precision]


def mac_bytes(precision):
    return 8 if precision == "f64" else 4


def workload_text(name):
    from tests import cases
    if name == "bubble3d256":
        return cases.bubble3d(n=256, radius=32.0)
    if name == "bubble3d128":
        return cases.bubble3d(n=128, radius=16.0)
    if name == "rti2d":
        return cases.rti2d(256, 1024)
    if name == "drop2d128":
        return cases.drop2d(n=128, radius=25)
    if name == "jet3d192x192x768":
        return cases.jet3d_c5()
    raise SystemExit(f"unknown workload {name}")


def load_traffic(kernel, workload, precision):
    """HBM bytes per launch of `kernel` measured by rocprofv3 PMC runs and
    committed under profiles/ (None when no measurement exists)."""
    path = os.path.join(ROOT, "profiles", f"traffic_{kernel}_{workload}_{precision}.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f).get("hbm_bytes_per_launch")


def cpu_baseline(text, steps_cap, seconds_budget):
    """FP64 OpenMP oracle (CLIP's per-node phases in CLIP's AoS layout) on the
    same configuration: up to `steps_cap` steps after initialisation, each
    timed on its own; the median step gives the rate."""
    from clip_amd import Config
    from tests.oracle_lib import Oracle

    cfg = Config.from_string(text)
    orc = Oracle(cfg).init()
    interior = 1
    for a in range(cfg.dim):
        interior *= cfg.extents()[a] - 2
    times, t_start = [], time.perf_counter()
    while len(times) < steps_cap:
        t0 = time.perf_counter()
        orc.step(1)
        times.append(time.perf_counter() - t0)
        if time.perf_counter() - t_start > seconds_budget:
            break
    dt = sorted(times)[len(times) // 2]
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    del orc
    return {"value": interior / dt / 1e6, "unit": "MLUPS", "cores": cores, "kind": "port",
            "sample": f"median of {len(times)} timed step(s) of the same workload after initialisation, FP64 "
                      f"oracle, OpenMP over {cores} threads (OMP_PROC_BIND=close, OMP_PLACES=cores), "
                      f"{sum(times):.1f} s"}


def measure(workload, precision, steps=50, warmup=5):
    """MLUPS and the k_fused roofline of one more workload/precision."""
    import torch
    from clip_amd import Config, Solver
    cfg = Config.from_string(workload_text(workload))
    sol = Solver(cfg, precision).init_fresh()
    inf = sol.info()
    sol.step(warmup)
    sol.synchronize()
    t0 = time.perf_counter()
    sol.step(steps)
    sol.synchronize()
    dt = time.perf_counter() - t0
    sol.enable_timing(True)
    sol.step(min(steps, 20))
    tm = sol.last_timing()
    sol.close()
    torch.cuda.synchronize()
    n = int(inf.interior)
    k = min(steps, 20)
    fused_ms = tm["fused_ms"] / k
    alg = n * (4 * inf.q * pop_bytes(precision) + (2 * inf.dim + 1) * mac_bytes(precision))
    return {"workload": workload, "precision": precision, "interior_nodes": n, "lattice": f"D{inf.dim}Q{inf.q}",
            "MLUPS": round(n * steps / dt / 1e6, 2), "ms_per_step": round(dt * 1e3 / steps, 4),
            "kernel_ms_per_step": {kk: round(v / k, 4) for kk, v in tm.items()},
            "k_fused_hbm_frac": round(alg / (fused_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}


def copy_bandwidth(mib=2048, reps=10):
    """Device-to-device copy (read + write) GB/s on torch's stream: the
    practical HBM ceiling the step kernels are compared with."""
    import torch
    n = mib * (1 << 20) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    b.copy_(a)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        b.copy_(a)
    e1.record()
    e1.synchronize()
    ms = e0.elapsed_time(e1) / reps
    del a, b
    return 2.0 * n * 4 / (ms * 1e-3) / 1e9


def adjoint_measure(precision, n=256, radius=50, steps=500, every=50):
    """Reverse-mode cost on the 2D drop (D2Q9, n^2, R = radius, density ratio
    1000): loss = 1/2 |phi_T - phi_target|^2 with the target from a run with
    sigma x 1.2 and mobility x 0.8; phi0 = the case's initial phase field plus
    1e-3 N(0,1) (seed 0) smoothed by one Laplacian pass and confined to the
    interface band by 4 phi (1 - phi).  Times the plain forward (steps x
    NSAllen::solve), the forward with states kept every `every` steps, and the
    reverse sweep (which recomputes each segment once), with the interface
    normal's derivative bounded on plateaus (normal floor 0.1)."""
    import numpy as np
    import torch
    from clip_amd import Config, Solver
    from tests import cases

    cfg = Config.from_string(cases.drop2d(n=n, radius=radius))
    base = Solver(cfg, precision).init_fresh()
    phi0 = base.interior("phi")[..., 0, 0]
    rng = np.random.default_rng(0)
    xi = rng.standard_normal(phi0.shape)
    lap = sum(np.roll(xi, s_, a) for a in (0, 1) for s_ in (1, -1)) - 4 * xi
    phi0 = phi0 + 1e-3 * (xi + 0.25 * lap) * 4 * phi0 * (1 - phi0)
    base.step(10)
    base.synchronize()
    t0 = time.perf_counter()
    base.step(steps)
    base.synchronize()
    fwd = time.perf_counter() - t0
    base.close()
    target = Solver(cfg.with_params(sigma=1.2 * cfg.sigma, mobility=0.8 * cfg.mobility),
                    precision).adjoint_forward(phi0, steps, steps)
    s = Solver(cfg, precision)
    s.adjoint_forward(phi0, 20, 10)  # allocation and first-launch costs outside the timed region
    s.adjoint_backward(np.zeros_like(phi0), 0.1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    phiT = s.adjoint_forward(phi0, steps, every)
    t1 = time.perf_counter()
    g0, gs, gm = s.adjoint_backward(phiT - target, 0.1)
    t2 = time.perf_counter()
    s.close()
    return {"workload": f"drop2d{n}_adjoint", "precision": precision, "lattice": "D2Q9", "steps": steps,
            "checkpoint_every": every, "normal_floor": 0.1, "forward_ms": round(fwd * 1e3, 3),
            "forward_with_checkpoints_ms": round((t1 - t0) * 1e3, 3), "backward_ms": round((t2 - t1) * 1e3, 3),
            "adjoint_over_forward": round((t2 - t1) / fwd, 3),
            "loss": float(0.5 * np.sum((phiT - target) ** 2)), "grad_sigma": gs, "grad_mobility": gm,
            "grad_phi0_norm": float(np.linalg.norm(g0))}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--workload", default="bubble3d256")
    ap.add_argument("--precision", default="f64", choices=["f64", "f32", "mixed_f16", "mixed_bf16"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=3)
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--no-adjoint", action="store_true", help="skip the reverse-mode measurement")
    ap.add_argument("--no-extra", action="store_true", help="skip the mixed-precision and jet lines")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="only run an elementwise FP32 copy (read + write) of this many MiB five times; "
                         "its rocprofv3 FETCH_SIZE / WRITE_SIZE calibrate the traffic counters")
    ap.add_argument("--adjoint-only", choices=["f32", "f64"], default=None,
                    help="only run the reverse-mode measurement at this precision (for kernel traces)")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the solver runs on a single MI355X; "
              "multi-GPU slab decomposition is a design note, not a run.", file=sys.stderr)
        sys.exit(2)

    import torch
    from clip_amd import Config, Solver

    torch.cuda.set_device(0)
    if args.copy_probe_mib:
        n = args.copy_probe_mib * (1 << 20) // 4
        a = torch.ones(n, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(5):
            torch.neg(a, out=b)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes_read": n * 4, "copy_probe_bytes_written": n * 4, "launches": 5}),
              flush=True)
        return
    if args.adjoint_only:
        print(json.dumps(adjoint_measure(args.adjoint_only)), flush=True)
        return
    text = workload_text(args.workload)
    cfg = Config.from_string(text)
    sol = Solver(cfg, args.precision).init_fresh()
    inf = sol.info()
    interior = int(inf.interior)
    Q, D = inf.q, inf.dim

    sol.step(args.warmup)
    sol.synchronize()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sol.step(args.steps)
    sol.synchronize()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    mlups = interior * args.steps / (t1 - t0) / 1e6

    # Per-kernel durations with HIP events on the solver's stream.
    tsteps = min(args.steps, 50)
    sol.enable_timing(True)
    sol.step(tsteps)
    tm = sol.last_timing()
    sol.enable_timing(False)
    fused_ms = tm["fused_ms"] / tsteps
    sp, sm = pop_bytes(args.precision), mac_bytes(args.precision)
    fused_bytes = interior * (4 * Q * sp + (2 * D + 1) * sm)  # f,g read+write; phi read; u read+write
    achieved = fused_bytes / (fused_ms * 1e-3) / 1e9
    step_bytes = 4 * Q * sp + (2 * D + 2) * sm
    health = sol.check_health()
    sol.close()
    del sol
    copy_gbs = copy_bandwidth()

    line = {
        "metric": "MLUPS + % HBM peak (D3Q19 coupled step) FP64 vs mixed; adjoint/forward time",
        "value": round(mlups, 2),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "f64" else "f32",
        "data": "synthetic",
        "config": {"workload": args.workload, "interior_nodes": interior, "lattice": f"D{D}Q{Q}",
                   "precision": args.precision, "population_storage": POP_STORAGE[args.precision],
                   "parallelism": "single-gpu"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": load_traffic("k_fused", args.workload, args.precision),
                     "kernel": "k_fused", "kernel_ms": round(fused_ms, 4),
                     "alg_bytes_per_launch": fused_bytes,
                     "copy_gbs": round(copy_gbs, 1), "frac_of_copy": round(achieved / copy_gbs, 4)},
        "step_hbm_frac": round(mlups * 1e6 * step_bytes / 1e9 / HBM_PEAK_GBS, 4),
        "kernel_ms_per_step": {k: round(v / tsteps, 4) for k, v in tm.items()},
        "phi_sum": health,
    }
    if not args.no_extra:
        other = "mixed_f16" if args.precision == "f64" else "f64"
        line["extra"] = [measure(args.workload, other), measure("jet3d192x192x768", "mixed_f16"),
                         measure("jet3d192x192x768", "f64")]
    if not args.no_adjoint:
        line["adjoint"] = [adjoint_measure("f32"), adjoint_measure("f64")]
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(text, args.cpu_steps, args.cpu_seconds)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
