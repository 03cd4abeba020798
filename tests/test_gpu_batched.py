"""Batched simulations (many independent communities in one launch, the
caller side of CBMFitting's parameter loops, src/fitting/gridSearch.jl:44):
every simulation of a batched context must equal its own oracle run."""
import numpy as np
import pytest

import cbm._abi as A
import oracle as O
from cbm.api import Context
from helpers import model, sim

pytestmark = pytest.mark.gpu


def sweep_sims(ns, seed=0, dims=2):
    rng = np.random.default_rng(seed)
    out = []
    for n in ns:
        phi = rng.uniform(0.2, 0.5)
        L = np.sqrt(n * np.pi / phi) if dims == 2 else (n * 4 / 3 * np.pi / phi) ** (1 / 3)
        k = rng.uniform(0.5, 2.0)
        eps = rng.uniform(0.0, 0.2)
        s = sim(n, [[0.0, L]] * dims, [k, 1.0, eps, 2.4])
        pos = [rng.uniform(0.0, L, n) for _ in range(dims)]
        out.append((s, pos))
    return out


@pytest.mark.parametrize("nb", [A.CELL_LINKED, A.VERLET_DISPLACEMENT])
def test_batched_sims_equal_their_oracles(gpu_available, nb):
    kw = dict(cell_edge=2.4) if nb == A.CELL_LINKED else dict(skin=2.8, max_neighbors=48, rebuild_disp=0.2)
    m = model(2, integrator=A.EULER, neighbors=nb, law=A.LAW_REP_ADH, dt=0.01, periodic=1, **kw)
    sims = sweep_sims([300, 1000, 2500, 700, 64, 4000, 257], seed=3)
    g = Context(m, [s for s, _ in sims])
    oracles = []
    for q, (s, pos) in enumerate(sims):
        o = O.Oracle(m, s)
        for a in range(2):
            o.set(a, pos[a])
            g.upload(a, pos[a], sim=q)
        o.init()
        oracles.append(o)
    g.neighbors()
    for rounds in range(3):
        g.step(20)
        for q, o in enumerate(oracles):
            o.step(20)
            n = o.n
            assert g.n(q) == n
            for a in range(2):
                assert np.array_equal(g.download(a, n, sim=q), o.get(a)), (q, a)


def test_batched_graph_replay(gpu_available):
    m = model(2, integrator=A.EULER, neighbors=A.CELL_LINKED, law=A.LAW_REP_ADH, dt=0.01, periodic=1, cell_edge=2.4)
    sims = sweep_sims([500, 1500, 900], seed=5)
    ctxs = [Context(m, [s for s, _ in sims]) for _ in range(2)]
    for g in ctxs:
        for q, (s, pos) in enumerate(sims):
            for a in range(2):
                g.upload(a, pos[a], sim=q)
        g.neighbors()
    ctxs[0].step(24)
    ctxs[1].step_graph(4, 6)
    for q, (s, _) in enumerate(sims):
        for a in range(2):
            assert np.array_equal(ctxs[0].download(a, s.n, sim=q), ctxs[1].download(a, s.n, sim=q))


def test_autotuner_choices_and_no_state_change(gpu_available, tmp_path):
    from cbm import autotune
    m = model(2, integrator=A.EULER, neighbors=A.CELL_LINKED, law=A.LAW_REP_ADH, dt=0.01, periodic=1, cell_edge=2.4)
    sims = sweep_sims([800, 3000, 1200], seed=9)

    def upload(ctx):
        for q, (s, pos) in enumerate(sims):
            for a in range(2):
                ctx.upload(a, pos[a], sim=q)

    cache = str(tmp_path / "tune.json")
    mt, info = autotune.tune(m, [s for s, _ in sims], upload, cutoff=2.4, steps=5, warm=2, cache_path=cache)
    assert info["choice"][0] in ("cell", "verlet")
    mt2, info2 = autotune.tune(m, [s for s, _ in sims], upload, cutoff=2.4, cache_path=cache)
    assert info2["cached"] and info2["choice"] == info["choice"]
    g = Context(mt, [s for s, _ in sims])
    upload(g)
    g.neighbors()
    g.step(3)
    before = [g.download(a, sims[1][0].n, sim=1) for a in range(2)]
    b = autotune.tune_block(g, cache + ".blocks")
    assert b in (64, 128, 256)
    after = [g.download(a, sims[1][0].n, sim=1) for a in range(2)]
    assert all(np.array_equal(x, y) for x, y in zip(before, after))
    assert autotune.tune_block(g, cache + ".blocks") == b
    # the tuned structure computes the same trajectory as the oracle
    o = O.Oracle(mt, sims[1][0])
    for a in range(2):
        o.set(a, sims[1][1][a])
    o.init()
    o.step(3)
    g.step(7)
    o.step(7)
    for a in range(2):
        assert np.array_equal(g.download(a, sims[1][0].n, sim=1), o.get(a))


def test_c4_sweep_sample_equals_its_oracles(gpu_available):
    # 128 simulations drawn as the C4 sweep draws its 512 (N log-uniform in
    # [500, 50000], packing 0.2 to 0.5, k and adhesion per simulation, 2-D
    # periodic, Euler, CellLinked 2.4): 30 steps, every simulation bit for bit
    import os
    import sys
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    import bench
    m, sims, states = bench.sweep_workload(n_sims=128, seed=6)
    print(f"C4 sample: {len(sims)} simulations, {sum(s.n for s in sims)} agents, 30 steps", flush=True)
    g = Context(m, sims)
    for q, st in enumerate(states):
        for a in range(2):
            g.upload(a, st[a], sim=q)
    g.neighbors()
    g.step(30)
    for q, (s, st) in enumerate(zip(sims, states)):
        o = O.Oracle(m, s)
        for a in range(2):
            o.set(a, st[a])
        assert o.init() == 0
        o.step(30)
        for a in range(2):
            assert np.array_equal(g.download(a, s.n, sim=q), o.get(a)), (q, a)
