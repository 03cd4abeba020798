"""evolve!(saveEach) snapshots kept on the device (cbm_snapshot_*): each
snapshot equals a download taken after the same step, in a growing colony
(N changes between saves), for every simulation of a batched context, and
for the medium; a save schedule that does not fit is refused before any
step runs."""
import numpy as np
import pytest

import cbm._abi as A
import oracle as O
from cbm.api import ABM, CBMIntegrators, CBMModels, CBMNeighbors, Community, Context, evolve_
from helpers import model, rel_err, sim, uniform_positions
from test_oracle_medium import medium_model

pytestmark = pytest.mark.gpu


def colony_model():
    return model(2, integrator=A.EULER, neighbors=A.CELL_LINKED, law=A.LAW_LINEAR, dt=0.05, cell_edge=2.0,
                 rule=A.RULE_GROWTH_DIVISION, rule_p=(0.5, 0.1, 0.5, 2.0), seed=4)


def colony(m, n0=200, cap=5000):
    rng = np.random.default_rng(8)
    fields = {A.F_X: rng.uniform(34, 46, n0), A.F_Y: rng.uniform(34, 46, n0), A.F_VOL: rng.uniform(1, 2, n0),
              A.F_VBIRTH: np.ones(n0), A.F_GROWTH: rng.uniform(0.45, 0.55, n0)}
    g = Context(m, [sim(n0, [[0.0, 80.0], [0.0, 80.0]], [1.0, 2.0], n_max=cap)])
    for f, v in fields.items():
        g.upload(f, v)
    g.neighbors()
    return g, fields


def test_snapshots_equal_downloads_in_a_growing_colony(gpu_available):
    m = colony_model()
    fl = [A.F_X, A.F_Y, A.F_ID, A.F_VOL]
    a, _ = colony(m)
    a.snapshot_config(fl, 7)
    a.snapshot_take()
    a.step_save(60, 10)
    assert a.snapshot_count() == 7
    b, _ = colony(m)
    ref = [{f: b.download(f, b.n()) for f in fl} | {"t": b.time()[0]}]
    for _ in range(6):
        b.step(10)
        ref.append({f: b.download(f, b.n()) for f in fl} | {"t": b.time()[0]})
    for k in range(7):
        for f in fl:
            v, n, t = a.snapshot(k, f)
            assert n == len(ref[k][f]) and t == ref[k]["t"]
            assert np.array_equal(v, ref[k][f])
    assert len(ref[-1][A.F_ID]) > len(ref[0][A.F_ID])
    # the step state after the saves is the same as without them
    for f in fl:
        assert np.array_equal(a.download(f, a.n()), ref[-1][f])
    with pytest.raises(A.CBMError) as e:     # one snapshot too many: refused before stepping
        a.step_save(10, 10)
    assert e.value.code == A.ERR_CAPACITY
    assert a.time()[1] == 60


def test_snapshots_of_every_simulation(gpu_available):
    m = model(2, integrator=A.EULER, neighbors=A.VERLET_DISPLACEMENT, law=A.LAW_LINEAR, dt=0.01, periodic=1,
              skin=2.4, max_neighbors=64, rebuild_disp=0.2)
    sims, pos = [], []
    for s, (n, L) in enumerate([(300, 30.0), (800, 50.0), (150, 20.0)]):
        sims.append(sim(n, [[0.0, L], [0.0, L]], [1.0, 2.0]))
        pos.append(uniform_positions(n, [[0.0, L], [0.0, L]], 40 + s))
    g = Context(m, sims)
    for s in range(3):
        for a in range(2):
            g.upload(a, pos[s][a], sim=s)
    g.neighbors()
    g.snapshot_config([A.F_X, A.F_Y], 3)
    g.step_save(30, 10)
    o = [O.Oracle(m, sims[s]) for s in range(3)]
    for s in range(3):
        for a in range(2):
            o[s].set(a, pos[s][a])
        assert o[s].init() == 0
    for k in range(3):
        for s in range(3):
            o[s].step(10)
            for a in range(2):
                v, n, t = g.snapshot(k, a, sim=s)
                assert n == sims[s].n
                assert rel_err(v, o[s].get(a)) <= 1e-12


def test_snapshots_in_evolve_with_medium(gpu_available):
    n = 120
    pos = uniform_positions(n, [[0.0, 32.0]] * 2, 3)
    abm = ABM(2, agentODE=CBMModels.linearRepulsion(1.0, 2.0), neighborsAlg=CBMNeighbors.CellLinked(cellEdge=2.0),
              agentAlg=CBMIntegrators.Euler(), periodic=True,
              medium={"NMedium": [32, 32], "D": 0.5, "consume": 0.1, "c0": 1.0})
    com = Community(abm, N=n, dt=0.1, simBox=[[0.0, 32.0]] * 2, x=pos[0], y=pos[1])
    evolve_(com, steps=12, saveEach=4, saveCurrentState=True)
    assert [round(s["t"], 12) for s in com.pastTimes] == [0.0, 0.4, 0.8, 1.2]
    ref = Community(abm, N=n, dt=0.1, simBox=[[0.0, 32.0]] * 2, x=pos[0], y=pos[1])
    evolve_(ref, steps=0)
    for k in range(1, 4):
        evolve_(ref, steps=4)
        got = np.empty(32 * 32)
        A.check(ref.ctx.lib.cbm_download_medium(ref.ctx.h, got.ctypes.data, got.size), ref.ctx.h)
        assert np.array_equal(com.pastTimes[k]["medium"], got)
        assert np.array_equal(com.pastTimes[k]["x"], ref.x) and np.array_equal(com.pastTimes[k]["id"], ref.id)
    assert np.array_equal(com.x, ref.x)
