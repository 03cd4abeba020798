"""Model-scope bodies (cbm_compile_model_scope: modelRule, modelODE and
modelSDE, run by modelStepRule! and modelStepDE! after the medium step,
src/CommunityStructure/step.jl:96-107) on the device against the oracle:
the parameters and the agents that read them equal the oracle bit for bit
(rule, Euler and Euler-Maruyama parameters, the linear law's stiffness
moved by the model), for one and for several simulations of a context;
graph replays equal eager steps (the step's time lives on the device);
the reference KAT's model part with agents reading gf; parameters set from
the host keep the device's other values; contexts running CellLinked on a
skin list refuse a model body."""
import numpy as np
import pytest

import cbm._abi as A
import model_bodies as MB
import oracle as O
from cbm.api import Context
from helpers import model, sim, uniform_positions

pytestmark = pytest.mark.gpu

P0 = [0.5, 1.0, 3.0, 5.0, 0.0, 0.0, 0.4, 0.0]     # k, c of the linear law; free slots


def setup(S=1, n=600, steps_seed=3, walk=1, graph=False):
    L = 20.0
    m = model(2, integrator=A.HEUN, neighbors=A.CELL_LINKED, law=A.LAW_LINEAR, dt=0.01, cell_edge=1.0, seed=41)
    m.walk = walk
    sims = []
    for s in range(S):
        p = list(P0)
        p[0] = 0.5 + 0.25 * s
        sims.append(sim(n, [[0, L]] * 2, p))
    g = Context(m, sims)
    g.compile_model_scope(MB.COMBINED)
    oracles = []
    for s in range(S):
        pos = uniform_positions(n, [[0, L]] * 2, steps_seed + s)
        o = O.Oracle(m, sims[s])
        for a in range(2):
            o.set(a, pos[a])
            g.upload(a, pos[a], sim=s)
        o.set_model_body(**MB.combined())
        o.set_sim(s)
        assert o.init() == 0
        oracles.append(o)
    g.neighbors()
    return g, oracles


def check(g, oracles, n):
    for s, o in enumerate(oracles):
        assert np.array_equal(g.params(sim=s), o.params()), s
        for a in range(2):
            assert np.array_equal(g.download(a, n, sim=s), o.get(a)), (s, a)


def test_model_scope_single_simulation(gpu_available):
    g, (o,) = setup()
    for chunk in (1, 7, 22):
        g.step(chunk)
        o.step(chunk)
        check(g, [o], 600)
    p = o.params()
    assert p[3] == 5.0 + 30 and p[6] != 0.4 and abs(p[0] - 2.0) < 1.5


def test_model_scope_batched_simulations(gpu_available):
    g, oracles = setup(S=3)
    g.step(25)
    for o in oracles:
        o.step(25)
    check(g, oracles, 600)
    assert len({float(o.params()[6]) for o in oracles}) == 3        # each simulation draws its own noise


def test_model_scope_graph_replays(gpu_available):
    g, (o,) = setup()
    g.step(2)
    g.step_graph(4, 5)
    g.step(1)
    o.step(23)
    check(g, [o], 600)


def test_reference_kat_agents_read_gf(gpu_available):
    # test/testCommunity.jl:35-200: dt(x) = gf for the agents, dt(gf) = 1,
    # gi += 1, ga[1,2] = 7; gf = 1 at the start, dt = 0.1
    m = model(2, integrator=A.EULER, neighbors=A.FULL, law=A.LAW_CUSTOM, dt=0.1)
    g = Context(m, [sim(10, [[0, 1]] * 2, [0.0, 1.0, 0.0, 0.0])])
    g.compile_model(MB.SELF_GF)
    g.compile_model_scope(MB.KAT)
    g.upload(A.F_X, np.zeros(10))
    g.neighbors()
    g.step(1)
    p = g.params()
    assert p[2] == 1.0 and p[3] == 7.0 and abs(p[1] - 1.1) < 1e-15
    assert np.allclose(g.download(A.F_X, 10), 0.1, rtol=0, atol=1e-15)      # the agents read gf = 1
    g.step(1)
    assert np.allclose(g.download(A.F_X, 10), 0.1 + 0.11, rtol=0, atol=1e-15)
    assert g.params()[2] == 2.0


def test_set_params_keeps_the_device_values(gpu_available):
    g, (o,) = setup()
    g.step(5)
    o.step(5)
    before = g.params()
    g.set_params([1.5])                 # k only
    after = g.params()
    assert after[0] == 1.5 and np.array_equal(after[1:], before[1:])


def test_skin_list_contexts_refuse_a_model_body(gpu_available):
    m = model(2, integrator=A.EULER, neighbors=A.CELL_LINKED, law=A.LAW_LINEAR, dt=0.01, cell_edge=1.0)
    g = Context(m, [sim(100, [[0, 10]] * 2, P0)])
    with pytest.raises(RuntimeError, match="walk = 1"):
        g.compile_model_scope(MB.COMBINED)
