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
    if name == "slab192x192x96":
        # one rank's share of the jet split eight ways along z: a fully
        # periodic box of that shape, whose z ghost planes (refreshed by the
        # periodic wrap) stand in for the halo planes a rank would receive
        return cases.box3d([(s, "periodic") for s in ("x-", "x+", "y-", "y+", "z-", "z+")], n=(192, 192, 96),
                           radius=16.0)
    raise SystemExit(f"unknown workload {name}")


def load_traffic(kernel, workload, precision):
    """HBM bytes per launch of `kernel` measured by rocprofv3 PMC runs and
    committed under profiles/ (None when no measurement exists)."""
    path = os.path.join(ROOT, "profiles", f"traffic_{kernel}_{workload}_{precision}.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f).get("hbm_bytes_per_launch")


# The cores this process may use, read before any OpenMP region runs (the
# oracle's OMP_PROC_BIND=close binds the calling thread to one core).
HOST_CORES = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)


def cpu_model():
    """The host CPU's model name with its family / model numbers (a virtual
    machine may report only a generic name) and the cores this process may use."""
    fields = {}
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if ":" not in line:
                    if fields:
                        break  # first processor only
                    continue
                k, v = (x.strip() for x in line.split(":", 1))
                fields.setdefault(k, v)
    except OSError:
        return "unknown"
    name = fields.get("model name", "unknown")
    if "cpu family" in fields and "model" in fields:
        name += f" (cpu family {fields['cpu family']}, model {fields['model']})"
    return name + f", {HOST_CORES} cores available"


def cpu_baseline(text, steps, runs=3):
    """FP64 OpenMP oracle (CLIP's per-node phases in CLIP's AoS layout) on the
    same configuration: `runs` timed runs of `steps` steady-state steps after
    initialisation and one untimed step; the median run gives the rate."""
    from clip_amd import Config
    from tests.oracle_lib import Oracle

    cfg = Config.from_string(text)
    orc = Oracle(cfg).init()
    orc.step(1)
    interior = 1
    for a in range(cfg.dim):
        interior *= cfg.extents()[a] - 2
    per_step = []
    for _ in range(runs):
        t0 = time.perf_counter()
        orc.step(steps)
        per_step.append((time.perf_counter() - t0) / steps)
    dt = sorted(per_step)[len(per_step) // 2]
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    del orc
    return {"value": interior / dt / 1e6, "unit": "MLUPS", "cores": cores, "kind": "port",
            "cpu_model": cpu_model(),
            "sample": f"median of {runs} runs of {steps} steady-state steps of the same workload (after "
                      f"initialisation and one untimed step), FP64 oracle, OpenMP over {cores} threads "
                      f"(OMP_PROC_BIND=close, OMP_PLACES=cores) on {cpu_model()}, "
                      f"{sum(per_step) * steps:.1f} s timed"}


def progress(msg):
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def rel_l2(a, b):
    import numpy as np
    den = float(np.linalg.norm(b.ravel()))
    return float(np.linalg.norm((a - b).ravel())) / den if den > 0 else float(np.linalg.norm(a.ravel()))


def measure(workload, precision, steps=50, warmup=5, graph=False, against_f64=False):
    """MLUPS and the k_fused roofline of one more workload/precision
    (graph: steps through the captured two-step HIP graph; against_f64: the
    rel-L2 of phi, u, p and the phi-sum drift against an FP64 run of the
    same number of steps)."""
    progress(f"{workload} {precision}{' (HIP graph)' if graph else ''}")
    import torch
    from clip_amd import Config, Solver
    cfg = Config.from_string(workload_text(workload))
    sol = Solver(cfg, precision).init_fresh()
    if graph:
        sol.enable_graph(True)
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
    sol.enable_timing(False)
    total_steps = warmup + steps + min(steps, 20)
    errors = None
    if against_f64 and precision != "f64":
        fields = {f: sol.interior(f) for f in ("phi", "u", "p")}
        s0 = sol.sum("phi")
        sol.close()
        ref = Solver(cfg, "f64").init_fresh()
        r0 = ref.sum("phi")
        ref.step(total_steps)
        errors = {f"rel_l2_{f}": rel_l2(fields[f], ref.interior(f)) for f in fields}
        errors["phi_sum_drift"] = abs(s0 - ref.sum("phi")) / abs(r0)
        errors["steps"] = total_steps
        ref.close()
    else:
        sol.close()
    torch.cuda.synchronize()
    n = int(inf.interior)
    k = min(steps, 20)
    fused_ms = tm["fused_ms"] / k
    alg = n * (4 * inf.q * pop_bytes(precision) + (2 * inf.dim + 1) * mac_bytes(precision))
    out = {"workload": workload, "precision": precision, "interior_nodes": n, "lattice": f"D{inf.dim}Q{inf.q}",
           "hip_graph": graph, "MLUPS": round(n * steps / dt / 1e6, 2), "ms_per_step": round(dt * 1e3 / steps, 4),
           "kernel_ms_per_step": {kk: round(v / k, 4) for kk, v in tm.items()},
           "k_fused_hbm_frac": round(alg / (fused_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
    if errors:
        out["against_f64"] = errors
    return out


def copy_rate_16b(mib=2048, reps=20):
    """GB/s (read + write) of the library's 16-byte-per-lane streaming copy
    (clip_copy_rate): the best pure-copy rate measured on these boxes, and
    the ceiling frac_of_copy uses."""
    import ctypes
    from clip_amd import _abi
    out = ctypes.c_double()
    _abi.check(_abi.load().clip_copy_rate(mib << 20, reps, ctypes.byref(out)))
    return out.value


def copy_bandwidth(mib=2048, reps=10):
    """Device-to-device copy (read + write) GB/s of torch's elementwise copy
    on its stream (reported beside the 16-byte copy)."""
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
    progress(f"drop2d{n} adjoint {precision}")
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
    seed = phiT - target  # dL/dphi_T of the loss, formed outside the timed backward
    tb = time.perf_counter()
    g0, gs, gm = s.adjoint_backward(seed, 0.1)
    t2 = time.perf_counter()
    s.close()
    # the same workload over 50 steps, where central differences of the loss
    # converge: the floored adjoint against them (past ~150 steps the loss has
    # no converged finite-difference derivative at this density ratio)
    T = 50
    tgt50 = Solver(cfg.with_params(sigma=1.2 * cfg.sigma, mobility=0.8 * cfg.mobility),
                   "f64").adjoint_forward(phi0, T, T)

    def loss(sig=None, mob=None):
        q = Solver(cfg.with_params(sigma=sig, mobility=mob), "f64").adjoint_forward(phi0, T, T)
        return 0.5 * float(np.sum((q - tgt50) ** 2))
    hs, hm = 1e-6 * cfg.sigma, 1e-6 * cfg.mobility
    fds = (loss(sig=cfg.sigma + hs) - loss(sig=cfg.sigma - hs)) / (2 * hs)
    fdm = (loss(mob=cfg.mobility + hm) - loss(mob=cfg.mobility - hm)) / (2 * hm)
    s50 = Solver(cfg, precision)
    p50 = s50.adjoint_forward(phi0, T, 10)
    _, gs50, gm50 = s50.adjoint_backward(p50 - tgt50, 0.1)
    s50.close()
    return {"workload": f"drop2d{n}_adjoint", "precision": precision, "lattice": "D2Q9", "steps": steps,
            "checkpoint_every": every, "normal_floor": 0.1, "forward_ms": round(fwd * 1e3, 3),
            "forward_with_checkpoints_ms": round((t1 - t0) * 1e3, 3), "backward_ms": round((t2 - tb) * 1e3, 3),
            "adjoint_over_forward": round((t2 - tb) / fwd, 3),
            "loss": float(0.5 * np.sum((phiT - target) ** 2)), "grad_sigma": gs, "grad_mobility": gm,
            "grad_phi0_norm": float(np.linalg.norm(g0)),
            "fd_check_50_steps": {"normal_floor": 0.1, "rel_err_sigma": abs(gs50 - fds) / abs(fds),
                                  "rel_err_mobility": abs(gm50 - fdm) / abs(fdm),
                                  "central_difference_h": "1e-6 x parameter, FP64 forward"}}


def adjoint_c5_measure(precision="f32", steps=200, plans=((20, 5), (14, 0))):
    """Reverse mode on the 3D jet (D3Q19, 192x192x768, modulated disc inlet,
    outflow): loss = 1/2 |phi_T - phi_target|^2 after `steps` steps from the
    case's initial phase field, target from a run with the 8 inlet mode
    amplitudes x 1.5; gradients w.r.t. the 8 amplitudes, sigma and the
    mobility.  One entry per checkpoint plan (every, inner): states kept
    every `every` steps and, with inner > 0, every `inner` steps inside each
    recomputed segment (two levels) or not (one level: more kept states,
    one recomputation pass fewer).  Peak device memory from
    hipMemGetInfo (through torch) after the checkpointed forward; the
    forward with kept states is timed on a fresh handle (state pool
    allocation and boundary maps included) and again on the same handle."""
    progress(f"jet3d192x192x768 adjoint {precision}")
    import numpy as np
    import torch
    from clip_amd import Config, Solver
    text = workload_text("jet3d192x192x768")
    cfg = Config.from_string(text)
    tcfg = Config.from_string(text)
    for m in range(tcfg.raw.inlet_modes):
        tcfg.raw.inlet_mode_amp[m] *= 1.5
    tgt = Solver(tcfg, precision).init_fresh()
    tgt.step(steps)
    target = tgt.interior("phi")[..., 0]
    tgt.close()
    base = Solver(cfg, precision).init_fresh()
    phi0 = base.interior("phi")[..., 0]
    base.step(5)
    base.synchronize()
    t0 = time.perf_counter()
    base.step(steps)
    base.synchronize()
    fwd = time.perf_counter() - t0
    base.close()
    out, first = [], None
    for every, inner in plans:
        progress(f"jet3d192x192x768 adjoint {precision}, kept every {every}, inner {inner}")
        torch.cuda.synchronize()
        free0, total = torch.cuda.mem_get_info()
        s = Solver(cfg, precision)
        t0 = time.perf_counter()
        phiT = s.adjoint_forward(phi0, steps, every, inner_every=inner)
        t1 = time.perf_counter()
        free1, _ = torch.cuda.mem_get_info()
        t1a = time.perf_counter()
        phiT = s.adjoint_forward(phi0, steps, every, inner_every=inner)
        t1b = time.perf_counter()
        seed = phiT - target
        tb = time.perf_counter()
        g0, gs, gm = s.adjoint_backward(seed, 0.1)
        t2 = time.perf_counter()
        ga = s.adjoint_inlet_gradient()
        held = int(s.info().device_bytes)
        s.close()
        e = {"workload": "jet3d192x192x768_adjoint", "precision": precision, "lattice": "D3Q19", "steps": steps,
             "checkpoint_every": every, "inner_checkpoint_every": inner, "normal_floor": 0.1,
             "forward_ms": round(fwd * 1e3, 3), "forward_with_checkpoints_ms": round((t1 - t0) * 1e3, 3),
             "forward_with_checkpoints_again_ms": round((t1b - t1a) * 1e3, 3),
             "backward_ms": round((t2 - tb) * 1e3, 3), "adjoint_over_forward": round((t2 - tb) / fwd, 3),
             "peak_device_bytes_used": int(total - free1), "device_bytes_before": int(total - free0),
             "handle_device_bytes": held,
             "loss": float(0.5 * np.sum((phiT - target) ** 2)),
             "grad_inlet_amplitudes": [float(x) for x in ga], "grad_sigma": gs, "grad_mobility": gm,
             "grad_phi0_norm": float(np.linalg.norm(g0))}
        if first is None:
            first = (ga, gs, gm, g0)
        else:
            e["gradients_equal_to_first_plan"] = bool(np.array_equal(ga, first[0]) and gs == first[1]
                                                      and gm == first[2] and np.array_equal(g0, first[3]))
        del g0, seed, phiT
        out.append(e)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--workload", default="bubble3d256")
    ap.add_argument("--precision", default="f64", choices=["f64", "f32", "mixed_f16", "mixed_bf16"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=8, help="steady-state steps per CPU baseline run (3 runs)")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the reverse-mode measurement")
    ap.add_argument("--no-extra", action="store_true", help="skip the mixed-precision and jet lines")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="only run an elementwise FP32 copy (read + write) of this many MiB five times; "
                         "its rocprofv3 FETCH_SIZE / WRITE_SIZE calibrate the traffic counters")
    ap.add_argument("--adjoint-only", choices=["f32", "f64"], default=None,
                    help="only run the reverse-mode measurement at this precision (for kernel traces)")
    ap.add_argument("--adjoint-c5-only", choices=["f32", "f64"], default=None,
                    help="only run the 3D jet reverse-mode measurement at this precision")
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
    if args.adjoint_c5_only:
        print(json.dumps(adjoint_c5_measure(args.adjoint_c5_only)), flush=True)
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
    copy_torch = copy_bandwidth()
    copy_gbs = copy_rate_16b()

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
                     "copy_gbs": round(copy_gbs, 1), "frac_of_copy": round(achieved / copy_gbs, 4),
                     "copy_gbs_torch": round(copy_torch, 1)},
        "step_hbm_frac": round(mlups * 1e6 * step_bytes / 1e9 / HBM_PEAK_GBS, 4),
        "kernel_ms_per_step": {k: round(v / tsteps, 4) for k, v in tm.items()},
        "phi_sum": health,
    }
    if not args.no_extra:
        other = "mixed_f16" if args.precision == "f64" else "f64"
        line["extra"] = [measure(args.workload, other, against_f64=True),
                         measure("jet3d192x192x768", "mixed_f16", against_f64=True),
                         measure("jet3d192x192x768", "f64"),
                         measure("slab192x192x96", "mixed_f16"),
                         measure("slab192x192x96", "f64"),
                         measure("drop2d128", "f64", steps=1000, warmup=20),
                         measure("drop2d128", "f64", steps=1000, warmup=20, graph=True),
                         measure("rti2d", "f64", steps=1000, warmup=20),
                         measure("rti2d", "f64", steps=1000, warmup=20, graph=True),
                         measure("rti2d", "f32", steps=1000, warmup=20),
                         measure("rti2d", "f32", steps=1000, warmup=20, graph=True)]
    if not args.no_adjoint:
        line["adjoint"] = [adjoint_measure("f32"), adjoint_measure("f64")] + adjoint_c5_measure("f32")
    if not args.no_cpu_baseline:
        progress("CPU baseline (FP64 OpenMP oracle)")
        line["cpu_baseline"] = cpu_baseline(text, args.cpu_steps)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
