"""Benchmark of the per-step agent hot path on one MI355X.

Workload (default): 3-D cell mechanics, plain FP64, N = 1,048,576 agents
uniform in a cube sized for packing fraction 0.3 (side 244.64), r = 1,
repulsion k = 1 plus a continuous adhesion well of depth 0.1 out to 2.4 r,
Heun dt = 0.005, Verlet list radius 2.7 rebuilt when an agent moved 0.15.
A "step" is one step! of the whole community: both Heun stages (neighbor
gather + pair law + integrator), the displacement trigger and the
(conditional) neighbor rebuild.  Inputs are resident on the device before
the timed region.  Prints one JSON line.

--workload em selects the Euler-Maruyama variant (isotropic noise D = 0.01,
Philox keyed by (seed, step, agent id)).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "cellbasedmodels.jl_amd"))

HBM_PEAK_GBS = 8000.0


def workload(kind, n):
    import cbm._abi as A
    L = (n * (4.0 / 3.0) * np.pi / 0.3) ** (1.0 / 3.0)
    m = A.ModelDesc()
    m.dims = 3
    m.integrator = A.HEUN if kind == "heun" else A.EM
    m.neighbors = A.VERLET_DISPLACEMENT
    m.law = A.LAW_REP_ADH
    m.max_neighbors = 32
    m.dt = 0.005
    m.skin = 2.7
    m.rebuild_disp = 0.15
    m.noise = np.sqrt(2 * 0.01) if kind == "em" else 0.0
    m.seed = 2
    s = A.SimDesc()
    s.n = n
    s.n_max = n
    for a in range(3):
        s.box_lo[a] = 0.0
        s.box_hi[a] = L
    for q, v in enumerate([1.0, 1.0, 0.1, 2.4]):
        s.p[q] = v
    rng = np.random.default_rng(2)
    pos = [rng.uniform(0.0, L, n) for _ in range(3)]
    return m, s, pos, L


def stage_bytes(kind, kbar):
    """Algorithmic HBM bytes per agent for each stage kernel (FP64 SoA, int32 ELL list)."""
    lst = 4.0 + 4.0 * kbar                      # list length + entries
    if kind == "heun":
        return {2: 24 + lst + 24,               # stage 1: read u, list; write h
                3: 24 + 24 + lst + 24 + 24}     # stage 2: read h, u, list, x_ref; write u
    return {1: 24 + lst + 8 + 24 + 24}          # EM: read u, list, id, x_ref; write u


def cpu_baseline(m, s, pos, budget_s):
    """The oracle (C++ OpenMP restatement of the reference CPU algorithm) on the
    same workload, timed on this host for a bounded number of steps."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    o = O.Oracle(m, s)
    for a in range(3):
        o.set(a, pos[a])
    o.init()
    o.step(1)     # warm-up (first step after the initial list build)
    steps, t0 = 0, time.perf_counter()
    while True:
        o.step(1)
        steps += 1
        el = time.perf_counter() - t0
        if el > budget_s or steps >= 50:
            break
    return {"value": s.n * steps / el, "unit": "agent-steps/s", "cores": O.threads(), "kind": "port",
            "sample": f"{steps} steps of the same N={s.n} workload after 2 warm-up steps "
                      f"(oracle/cbm_oracle.cpp, OpenMP, FP64)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=1048576)
    ap.add_argument("--workload", choices=["heun", "em"], default="heun")
    ap.add_argument("--graph", type=int, default=0, help="capture this many steps per hipGraph replay")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--profile-steps", type=int, default=50)
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="first run a plain device-to-device copy of MIB MiB (calibrates FETCH_SIZE/WRITE_SIZE)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; multi-GPU runs are not part of it",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    if args.copy_calibration:
        import torch
        src = torch.ones(args.copy_calibration * (1 << 20) // 8, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    from cbm.api import Context
    m, s, pos, L = workload(args.workload, args.n)
    ctx = Context(m, [s])
    for a in range(3):
        ctx.upload(a, pos[a])
    ctx.neighbors()
    ctx.step(args.warmup)
    ctx.sync()
    rb0 = ctx.rebuilds()

    t0 = time.perf_counter()
    if args.graph:
        ctx.step_graph(args.graph, args.steps // args.graph)
        steps = (args.steps // args.graph) * args.graph
    else:
        ctx.step(args.steps)
        steps = args.steps
    ctx.sync()
    t1 = time.perf_counter()
    dev_ms, _, _ = ctx.last_timing()
    rebuilds = ctx.rebuilds() - rb0
    wall = t1 - t0
    value = s.n * steps / wall

    # per-kernel times from HIP events on the library stream
    ms, nl = ctx.profile(args.profile_steps)
    lens = ctx.count_neighbors(0.0)
    kbar = float(lens.mean())
    sb = stage_bytes(args.workload, kbar)
    dom = max(sb, key=lambda k: ms[k])
    achieved = sb[dom] * s.n / (ms[dom] * 1e-3) / 1e9
    step_bytes = sum(sb.values()) * s.n
    rec = {
        "metric": "agent-steps/s @1M 3D",
        "value": value,
        "unit": "agent-steps/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": wall / steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"C2 3-D cell mechanics, {args.workload.upper()}, Verlet r_list 2.7, rebuild at 0.15",
                   "N": s.n, "box": L, "dt": m.dt, "graph_steps": args.graph},
        "roofline": {
            "bound": "hbm",
            "kernel": {2: "k_stage Heun stage 1", 3: "k_stage Heun stage 2", 1: "k_stage EM"}[dom],
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": None,
            "alg_bytes_per_launch": sb[dom] * s.n,
            "avg_launch_ms": ms[dom],
            "mean_list_len": kbar,
        },
        "step_roofline": {
            "alg_bytes_per_step": step_bytes,
            "achieved_GBs": step_bytes / (wall / steps) / 1e9,
            "frac": step_bytes / (wall / steps) / 1e9 / HBM_PEAK_GBS,
            "neighbor_update_ms": ms[8],
            "device_ms_per_step": dev_ms / steps if not args.graph else None,
            "rebuilds_in_timed_steps": rebuilds,
        },
    }
    if not args.no_cpu:
        rec["cpu_baseline"] = cpu_baseline(m, s, pos, args.cpu_seconds)
    print(json.dumps(rec), flush=True)


if __name__ == "__main__":
    main()
