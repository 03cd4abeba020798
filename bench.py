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
Philox keyed by (seed, step, agent id)); --workload heun_cl runs the same
Heun model on a cell-linked grid of edge 2.4 (CellLinked loop order, walked
on a skin list of radius 2.7).
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
# HBM bytes per launch of the stage kernels, from rocprofv3 FETCH_SIZE and
# WRITE_SIZE runs scaled by what the same counters read for a device copy of
# known size; written next to the kernel-trace summaries.
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic.json")


def measured_traffic(kernel, n):
    """HBM bytes per launch of `kernel` at N = n, or None when not measured."""
    try:
        with open(TRAFFIC_FILE) as f:
            table = json.load(f)
    except (OSError, ValueError):
        return None
    e = table.get(kernel)
    if not e or e.get("N") != n:
        return None
    return e.get("bytes_per_launch")


def workload(kind, n):
    import cbm._abi as A
    L = (n * (4.0 / 3.0) * np.pi / 0.3) ** (1.0 / 3.0)
    m = A.ModelDesc()
    m.dims = 3
    m.integrator = A.EM if kind == "em" else A.HEUN
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
    if kind == "heun_cl":
        # the same model on a cell grid of edge 2.4 (the law's cutoff), walked
        # in CellLinked order on a skin list
        m.neighbors = A.CELL_LINKED
        m.cell_edge = 2.4
    rng = np.random.default_rng(2)
    pos = [rng.uniform(0.0, L, n) for _ in range(3)]
    return m, s, pos, L


# Cell-linked rebuild of every step, per agent (D = 3): cell key from the
# positions (8D), key write and read (8), the sort's passes over (key, row)
# pairs (16 x 3), the permutation of the positions (16D) and the cell starts (8):
# 24 D + 64 bytes.
CL_REBUILD_BYTES = 24 * 3 + 64


def stage_bytes(kind, kbar):
    """Algorithmic HBM bytes per agent for each stage kernel (FP64 SoA, int32 ELL list)."""
    lst = 4.0 + 4.0 * kbar                      # list length + entries
    if kind == "heun_cl":
        # CellLinked on a skin list: stage 1 reads u, the skin list, the
        # agent's cell and x_ref (its distance from the list's positions) and
        # writes h and the ordered stage-2 list; stage 2 reads h, u, that
        # list and x_ref and writes u' and the next cell
        return {2: 24 + lst + 4 + 24 + 24 + lst, 3: 24 + 24 + lst + 24 + 24 + 4}
    if kind == "heun":
        return {2: 24 + lst + 24,               # stage 1: read u, list; write h
                3: 24 + 24 + lst + 24 + 24}     # stage 2: read h, u, list, x_ref; write u
    return {1: 24 + lst + 8 + 24 + 24}          # EM: read u, list, id, x_ref; write u


def measure(kind, n, steps, warmup, graph, profile_steps):
    """The C2 model (Verlet "heun"/"em", or cell-linked "heun_cl") at N = n:
    agent-steps/s over `steps` timed steps after `warmup`, the dominant stage
    kernel's roofline from HIP events on the library stream, and the whole
    step against its algorithmic bytes."""
    from cbm.api import Context
    m, s, pos, L = workload(kind, n)
    ctx = Context(m, [s])
    for a in range(3):
        ctx.upload(a, pos[a])
    ctx.neighbors()
    ctx.step(warmup)
    ctx.sync()
    rb0 = ctx.rebuilds()

    t0 = time.perf_counter()
    if graph:
        ctx.step_graph(graph, steps // graph)
        steps = (steps // graph) * graph
    else:
        ctx.step(steps)
    ctx.sync()
    t1 = time.perf_counter()
    dev_ms, _, _ = ctx.last_timing()
    rebuilds = ctx.rebuilds() - rb0
    wall = t1 - t0

    # per-kernel times from HIP events on the library stream
    ms, nl = ctx.profile(profile_steps)
    # mean list length: the Verlet list, or the skin list of the cell-linked
    # model (its candidates within the list radius, reach 2.4 x 1.125)
    kbar = float(ctx.count_neighbors(2.4 * 1.125 if kind == "heun_cl" else 0.0).mean())
    sb = stage_bytes(kind, kbar)
    dom = max(sb, key=lambda k: ms[k])
    achieved = sb[dom] * s.n / (ms[dom] * 1e-3) / 1e9
    # the skin-list rebuilds, amortised: the cell sort (24 D + 64), the list
    # (4 + 4 kbar) and x_ref (8 D) per agent and rebuild
    rebuild_bytes = (CL_REBUILD_BYTES + 4 + 4 * kbar + 24) * rebuilds / steps
    step_bytes = (sum(sb.values()) + rebuild_bytes) * s.n
    kern = "k_clv" if kind == "heun_cl" else "k_stage"
    dom_name = {2: f"{kern} Heun stage 1", 3: f"{kern} Heun stage 2", 1: f"{kern} EM"}[dom]
    ctx.close()
    return {
        "value": s.n * steps / wall,
        "steps": steps,
        "ms_per_step": wall / steps * 1e3,
        "config": {"workload": ("C2 3-D cell mechanics, HEUN, cell-linked (edge 2.4) on a skin list "
                                "(radius 2.7, rebuilt at 0.15 displacement)"
                                if kind == "heun_cl" else
                                f"C2 3-D cell mechanics, {kind.upper()}, Verlet r_list 2.7, rebuild at 0.15"),
                   "N": s.n, "box": L, "dt": m.dt, "graph_steps": graph},
        "roofline": {
            "bound": "hbm",
            "kernel": dom_name,
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": measured_traffic(dom_name, s.n) if kind in ("heun", "heun_cl") else None,
            "alg_bytes_per_launch": sb[dom] * s.n,
            "avg_launch_ms": ms[dom],
            "mean_list_len": kbar,
        },
        "step_roofline": {
            "alg_bytes_per_step": step_bytes,
            "alg_bytes_per_agent_step": step_bytes / s.n,
            "achieved_GBs": step_bytes / (wall / steps) / 1e9,
            "frac": step_bytes / (wall / steps) / 1e9 / HBM_PEAK_GBS,
            "neighbor_update_ms": ms[8],
            "device_ms_per_step": dev_ms / steps if not graph else None,
            "rebuilds_in_timed_steps": rebuilds,
        },
        # the same time against the bytes of a cell-linked step that sorts
        # the grid every step (both stages, 52 + 76 B, plus 24 D + 64 B)
        **({"step_roofline_grid_sorted_every_step": {
            "alg_bytes_per_agent_step": 52 + 76 + CL_REBUILD_BYTES,
            "frac": (52 + 76 + CL_REBUILD_BYTES) * s.n / (wall / steps) / 1e9 / HBM_PEAK_GBS}}
           if kind == "heun_cl" else {}),
    }


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
    cores = O.threads()
    rec = {"value": s.n * steps / el, "unit": "agent-steps/s", "cores": cores, "kind": "port",
           "sample": f"{steps} steps of the same N={s.n} workload after 2 warm-up steps "
                     f"(oracle/cbm_oracle.cpp, OpenMP, FP64)",
           "cpu_model": cpu_model()}
    # the same oracle on one thread, a few steps
    O.set_threads(1)
    try:
        steps1, t0 = 0, time.perf_counter()
        while True:
            o.step(1)
            steps1 += 1
            el1 = time.perf_counter() - t0
            if el1 > budget_s / 4 or steps1 >= 5:
                break
    finally:
        O.set_threads(cores)
    rec["single_thread"] = {"value": s.n * steps1 / el1, "cores": 1, "sample": f"{steps1} further steps"}
    return rec


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def run_em(n, steps=200):
    """The C2 Euler-Maruyama variant (sigma = sqrt(2 D), D = 0.01, Philox noise)."""
    from cbm.api import Context
    m, s, pos, L = workload("em", n)
    ctx = Context(m, [s])
    for a in range(3):
        ctx.upload(a, pos[a])
    ctx.neighbors()
    ctx.step(20)
    ctx.sync()
    t0 = time.perf_counter()
    ctx.step(steps)
    ctx.sync()
    wall = time.perf_counter() - t0
    ms, nl = ctx.profile(20)
    ctx.close()
    return {"value": s.n * steps / wall, "unit": "agent-steps/s", "steps": steps,
            "ms_per_step": wall / steps * 1e3, "stage_ms": ms[1],
            "config": f"C2 EM variant, N={s.n}, sigma=sqrt(0.02), Verlet r_list 2.7, FP64"}


def run_c1(steps=1000):
    """C1, the reference's CPU-sized case: 2-D N=1000 periodic 50 x 50, linear
    repulsion k=1 for d < 2r, Euler dt=0.01, cell-linked (edge 2), 1000 steps,
    on the device and on the oracle (all host threads)."""
    import cbm._abi as A
    from cbm.api import Context
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    n, L = 1000, 50.0
    m = A.ModelDesc()
    m.dims, m.integrator, m.neighbors, m.law, m.periodic = 2, A.EULER, A.CELL_LINKED, A.LAW_LINEAR, 1
    m.dt, m.cell_edge, m.max_neighbors = 0.01, 2.0, 32
    s = A.SimDesc()
    s.n = s.n_max = n
    s.box_hi[0] = s.box_hi[1] = L
    s.p[0], s.p[1] = 1.0, 2.0
    rng = np.random.default_rng(1)
    pos = [rng.uniform(0.0, L, n), rng.uniform(0.0, L, n)]
    ctx = Context(m, [s])
    o = O.Oracle(m, s)
    for a in range(2):
        ctx.upload(a, pos[a])
        o.set(a, pos[a])
    ctx.neighbors()
    o.init()
    ctx.sync()
    t0 = time.perf_counter()
    ctx.step(steps)
    ctx.sync()
    tg = time.perf_counter() - t0
    t0 = time.perf_counter()
    o.step(steps)
    tc = time.perf_counter() - t0
    same = all(np.array_equal(ctx.download(a, n), o.get(a)) for a in range(2))
    ctx.close()
    # the same run with the step loop captured in a hipGraph (100 steps, replayed)
    gg = Context(m, [s])
    for a in range(2):
        gg.upload(a, pos[a])
    gg.neighbors()
    gg.sync()
    t0 = time.perf_counter()
    gg.step_graph(100, steps // 100)
    gg.sync()
    tgr = time.perf_counter() - t0
    same_g = all(np.array_equal(gg.download(a, n), o.get(a)) for a in range(2))
    gg.close()
    return {"gpu_value": n * steps / tg, "gpu_graph_value": n * steps / tgr, "cpu_value": n * steps / tc,
            "unit": "agent-steps/s", "cpu_cores": O.threads(), "steps": steps,
            "bit_identical": bool(same), "graph_bit_identical": bool(same_g),
            "config": "C1 2-D N=1000 periodic 50x50, linear k=1, Euler dt=0.01, CL edge 2, FP64"}


def sweep_workload(n_sims=512, seed=4):
    """512 independent 2-D periodic soft-sphere simulations, N log-uniform in
    [500, 50000], packing fraction U[0.2, 0.5], k U[0.5, 2], adhesion U[0, 0.2]."""
    import cbm._abi as A
    rng = np.random.default_rng(seed)
    ns = np.round(np.exp(rng.uniform(np.log(500), np.log(50000), n_sims))).astype(np.int64)
    sims, states = [], []
    for n in ns:
        phi = rng.uniform(0.2, 0.5)
        L = float(np.sqrt(n * np.pi / phi))
        s = A.SimDesc()
        s.n = s.n_max = int(n)
        s.box_hi[0] = s.box_hi[1] = L
        for q, v in enumerate([rng.uniform(0.5, 2.0), 1.0, rng.uniform(0.0, 0.2), 2.4]):
            s.p[q] = v
        sims.append(s)
        states.append([rng.uniform(0.0, L, n), rng.uniform(0.0, L, n)])
    m = A.ModelDesc()
    m.dims, m.integrator, m.law, m.periodic, m.dt = 2, A.EULER, A.LAW_REP_ADH, 1, 0.01
    m.neighbors, m.cell_edge = A.CELL_LINKED, 2.4
    return m, sims, states


def run_sweep(steps, plain_sims, plain_steps):
    """Batched (one context, one launch per kernel for all simulations) against
    the plain path (one context per simulation, stepped one after another as
    the reference's fitting loops do)."""
    import tempfile
    from cbm import autotune
    from cbm.api import Context
    m, sims, states = sweep_workload()

    def upload(ctx, idx=None):
        for q, st in enumerate(states if idx is None else [states[idx]]):
            for a in range(2):
                ctx.upload(a, st[a], sim=q)

    cache = os.path.join(tempfile.gettempdir(), "cbm_tune.json")
    mt, info = autotune.tune(m, sims, upload, cutoff=2.4, steps=10, warm=3, cache_path=cache)
    ctx = Context(mt, sims)
    upload(ctx)
    ctx.neighbors()
    block = autotune.tune_block(ctx, cache + ".blocks")
    ctx.step(10)
    ctx.sync()
    t0 = time.perf_counter()
    ctx.step(steps)
    ctx.sync()
    tb = time.perf_counter() - t0
    total = sum(s.n for s in sims)
    batched = total * steps / tb
    ctx.close()
    # plain path on a size-stratified sample of the same simulations
    order = np.argsort([s.n for s in sims])
    pick = order[np.linspace(0, len(order) - 1, plain_sims).astype(int)]
    work, tp = 0, 0.0
    for q in pick:
        c1 = Context(mt, [sims[q]])
        upload(c1, q)
        c1.neighbors()
        c1.step(5)
        c1.sync()
        t0 = time.perf_counter()
        c1.step(plain_steps)
        c1.sync()
        tp += time.perf_counter() - t0
        work += sims[q].n * plain_steps
        c1.close()
    plain = work / tp
    return {"value": batched, "unit": "agent-steps/s", "plain_value": plain, "ratio": batched / plain,
            "n_sims": len(sims), "agents": total, "steps": steps, "ms_per_step": tb / steps * 1e3,
            "structure": info["choice"], "block": block,
            "plain_sample": f"{plain_sims} simulations spanning the N range, {plain_steps} steps each",
            "config": "C4 512 2-D periodic sims, N log-uniform [500, 50000], Euler dt=0.01, FP64"}


def adjoint_workload(n=65536, sets=64, steps=200, saves=(50, 100, 150, 200), seed=5):
    """C5: 3-D overdamped Euler at packing fraction 0.3, dt 0.01; the target is
    the run from (k, r_eq, eps) = (1.3, 1.1, 0.08); the gradient is taken at
    `sets` parameter sets around (1, 1, 0.05)."""
    import cbm._abi as A
    from cbm.api import Context
    L = (n * 4 / 3 * np.pi / 0.3) ** (1 / 3)
    m = A.ModelDesc()
    m.dims, m.integrator, m.law, m.dt = 3, A.EULER, A.LAW_REP_ADH, 0.01
    m.neighbors, m.skin, m.max_neighbors, m.rebuild_disp = A.VERLET_DISPLACEMENT, 3.0, 48, 0.15
    rng = np.random.default_rng(seed)
    pos = [rng.uniform(0.0, L, n) for _ in range(3)]

    def simd(theta):
        s = A.SimDesc()
        s.n = s.n_max = n
        for a in range(3):
            s.box_hi[a] = L
        for q, v in enumerate(list(theta) + [2.4]):
            s.p[q] = v
        return s

    tctx = Context(m, [simd((1.3, 1.1, 0.08))])
    for a in range(3):
        tctx.upload(a, pos[a])
    tctx.neighbors()
    target, done = [], 0
    for st in saves:
        tctx.step(st - done)
        done = st
        target.append(np.stack([tctx.download(a, n) for a in range(3)]))
    tctx.close()
    thetas = [np.array([1.0, 1.0, 0.05]) * rng.uniform(0.9, 1.1, 3) for _ in range(sets)]
    return m, [simd(t) for t in thetas], pos, np.stack(target), thetas


def run_adjoint(steps=200, every=20):
    from cbm.api import Context
    saves = (50, 100, 150, 200)
    m, sims, pos, target, thetas = adjoint_workload(steps=steps, saves=saves)

    def fresh():
        c = Context(m, sims)
        for q in range(len(sims)):
            for a in range(3):
                c.upload(a, pos[a], sim=q)
        c.neighbors()
        c.sync()
        return c

    cf = fresh()
    t0 = time.perf_counter()
    cf.step(steps)
    cf.sync()
    t_fwd = time.perf_counter() - t0
    cf.close()
    ca = fresh()
    # first call: uploads the targets (64 x 4 x 3 x 65536 doubles) and sizes
    # the work buffers; the timed call then finds its inputs in HBM, as the
    # forward run does (the adjoint leaves the community at its initial state)
    t0 = time.perf_counter()
    ca.adjoint(steps, every, saves, [target] * len(sims))
    t_first = time.perf_counter() - t0
    t0 = time.perf_counter()
    loss, grad = ca.adjoint(steps, every, saves, None)
    t_adj = time.perf_counter() - t0
    ca.close()
    agents = sum(s.n for s in sims)
    return {"adjoint_ms": t_adj * 1e3, "forward_ms": t_fwd * 1e3, "ratio": t_adj / t_fwd,
            "first_call_ms": t_first * 1e3,
            "agents": agents, "sets": len(sims), "steps": steps, "checkpoint_every": every,
            "forward_agent_steps_per_s": agents * steps / t_fwd,
            "loss0": float(loss[0]), "grad0": [float(v) for v in grad[0]],
            "config": "C5 3-D overdamped Euler, N=65536 x 64 parameter sets, 200 steps, dt 0.01, FP64"}


def run_colony(target_n=2_000_000, max_steps=8000):
    """C3: 1000 agents (r = 1, k = 1) at the centre of a 2048 x 2048 medium
    (spacing 1, zero flux, D = 1, uptake 0.1 c per agent, FTCS dt = 0.2, c0 = 1);
    volumes grow at 0.01 +-10 % per unit time and divide at twice the birth
    volume.  Runs until target_n agents.  Division outpaces the weak
    repulsion, so the colony compresses far past disc packing (a core of
    ~7 agents per unit area at 623k agents, radius ~210): 2M agents stay
    well inside the box."""
    import cbm._abi as A
    from cbm.api import Context
    m = A.ModelDesc()
    m.dims, m.integrator, m.law, m.dt = 2, A.EULER, A.LAW_LINEAR, 0.2
    m.neighbors, m.cell_edge = A.CELL_LINKED, 2.0
    m.rule, m.seed = A.RULE_GROWTH_DIVISION, 3
    for q, v in enumerate((0.01, 0.1, 0.5, 2.0)):
        m.rule_p[q] = v
    m.medium_n[0] = m.medium_n[1] = 2048
    m.medium_D, m.medium_consume, m.medium_c0 = 1.0, 0.1, 1.0
    n0 = 1000
    s = A.SimDesc()
    s.n, s.n_max = n0, int(target_n * 2.1)
    s.box_hi[0] = s.box_hi[1] = 2048.0
    s.p[0], s.p[1] = 1.0, 2.0
    rng = np.random.default_rng(3)
    r = np.sqrt(n0 / 0.9) * np.sqrt(rng.uniform(0, 1, n0))
    th = rng.uniform(0, 2 * np.pi, n0)
    ctx = Context(m, [s])
    ctx.upload(A.F_X, 1024.0 + r * np.cos(th))
    ctx.upload(A.F_Y, 1024.0 + r * np.sin(th))
    ctx.upload(A.F_VOL, rng.uniform(1.0, 2.0, n0))
    ctx.upload(A.F_VBIRTH, np.ones(n0))
    ctx.upload(A.F_GROWTH, rng.uniform(0.009, 0.011, n0))
    ctx.neighbors()
    ctx.sync()
    steps, agent_steps, chunk = 0, 0, 100
    t0 = time.perf_counter()
    n = n0
    while n < target_n and steps < max_steps:
        ctx.step(chunk)
        n_new = ctx.n()
        agent_steps += chunk * (n + n_new) / 2      # N grows smoothly within a chunk
        n = n_new
        steps += chunk
    wall = time.perf_counter() - t0
    ctx.close()
    return {"value": agent_steps / wall, "unit": "agent-steps/s (N_t summed over steps)", "steps": steps,
            "final_n": n, "wall_s": wall, "medium_node_steps_per_s": 2048 * 2048 * steps / wall,
            "config": f"C3 2-D colony 1000 -> {target_n} agents with a 2048^2 FTCS medium, Euler dt 0.2, CL edge 2, FP64"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=1048576)
    ap.add_argument("--workload", choices=["heun", "em", "heun_cl"], default="heun")
    ap.add_argument("--graph", type=int, default=0, help="capture this many steps per hipGraph replay")
    ap.add_argument("--cl", type=int, default=500, metavar="STEPS",
                    help="also time the same model on a cell-linked grid of edge 2.4 (0 skips)")
    ap.add_argument("--em", type=int, default=1, help="also time the C2 Euler-Maruyama variant")
    ap.add_argument("--c1", type=int, default=1, help="also time C1 on the device and the oracle")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--profile-steps", type=int, default=50)
    ap.add_argument("--sweep-steps", type=int, default=500, help="batched sweep steps (0 skips the sweep)")
    ap.add_argument("--adjoint", type=int, default=1, help="1: also time the C5 adjoint against its forward run")
    ap.add_argument("--colony", type=int, default=1, help="1: also run the C3 growing colony with its medium")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="first run a plain device-to-device copy of MIB MiB (calibrates FETCH_SIZE/WRITE_SIZE)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "bench.py measures one GPU; runs on several GPUs are not supported",
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

    print(f"bench: C2 {args.workload} N={args.n}", file=sys.stderr, flush=True)
    r = measure(args.workload, args.n, args.steps, args.warmup, args.graph, args.profile_steps)
    rec = {
        "metric": "agent-steps/s @1M 3D",
        "value": r["value"],
        "unit": "agent-steps/s",
        "n_gpus": 1,
        "steps": r["steps"],
        "warmup": args.warmup,
        "ms_per_step": r["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": r["config"],
        "roofline": r["roofline"],
        "step_roofline": r["step_roofline"],
    }
    m, s, pos, L = workload(args.workload, args.n)

    def phase(name):
        print(f"bench: {name}", file=sys.stderr, flush=True)

    if args.cl and args.workload != "heun_cl":
        phase("C2 on a cell-linked grid")
        rec["cl"] = measure("heun_cl", args.n, args.cl, args.warmup, 0, args.profile_steps)
    if args.em:
        phase("C2 Euler-Maruyama variant")
        rec["em"] = run_em(args.n)
    if args.c1:
        phase("C1 (2-D, N=1000, 1000 steps) on the device and the oracle")
        rec["c1"] = run_c1()
    if args.sweep_steps:
        phase("batched sweep (512 sims)")
        rec["batched"] = run_sweep(args.sweep_steps, plain_sims=16, plain_steps=100)
    if args.adjoint:
        phase("adjoint (64 x 65536 agents, 200 steps)")
        rec["adjoint"] = run_adjoint()
    if args.colony:
        phase("growing colony with medium")
        rec["colony"] = run_colony()
    if not args.no_cpu:
        phase("cpu baseline")
        rec["cpu_baseline"] = cpu_baseline(m, s, pos, args.cpu_seconds)
    print(json.dumps(rec), flush=True)


if __name__ == "__main__":
    main()
