"""agentRule bodies with a neighbor loop (RuleAgent::forNeighbors, the
reference's @loopOverNeighbors in a rule): the neighbor KAT of
test/testCommunity.jl:361-415 for every neighbor algorithm, and a
contact-inhibition rule (neighbor counts, nearest distances, sums of the
neighbors' parameters, additions decided by the count, agents moved by it)
against the oracle running its Python restatement."""
import numpy as np
import pytest

import cbm._abi as A
import oracle as O
import rule_bodies as RB
from cbm.api import Context
from helpers import golden, lattice, model, rel_err, sim, uniform_positions

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dims", [1, 2, 3])
def test_neighbor_loop_lattice_kat(gpu_available, dims):
    from test_oracle import lattice_case
    g = golden("ref_neighbors_lattice.json")
    for alg in g["algorithms"]:
        m, s = lattice_case(dims, alg)
        m.rule = A.RULE_CUSTOM
        m.rule_p[0] = g["radius"]
        ctx = Context(m, [s])
        ctx.compile_model(rule=RB.COUNT_NEIGHBORS)
        for a, v in enumerate(lattice(dims)):
            ctx.upload(a, v)
        ctx.neighbors()
        ctx.step(1)
        got = ctx.download(A.F_U0, 3 ** dims).tolist()
        assert got == [float(v) for v in g["expected"][str(dims)]], alg["name"]


@pytest.mark.parametrize("nb", [A.FULL, A.CELL_LINKED, A.VERLET_DISPLACEMENT])
def test_contact_inhibition_matches_oracle(gpu_available, nb):
    n0, L = 150, 30.0
    rp = (1.5, 1.0, 0.0, 0.0)
    kw = dict(cell_edge=2.0) if nb == A.CELL_LINKED else (
        dict(skin=2.5, max_neighbors=128, rebuild_disp=0.2) if nb == A.VERLET_DISPLACEMENT else {})
    m = model(2, integrator=A.EULER, neighbors=nb, law=A.LAW_LINEAR, rule=A.RULE_CUSTOM, rule_p=rp, dt=0.02,
              periodic=1, **kw)
    s = sim(n0, [[0.0, L], [0.0, L]], [1.0, 2.0], n_max=3000)
    pos = uniform_positions(n0, [[0.0, L], [0.0, L]], 77)
    u3 = np.random.default_rng(1).uniform(0.0, 1.0, n0)
    o = O.Oracle(m, s)
    o.set_rule(RB.inhibit(rp))
    g = Context(m, [s])
    g.compile_model(rule=RB.INHIBIT)
    for f, v in ((A.F_X, pos[0]), (A.F_Y, pos[1]), (A.F_U3, u3)):
        o.set(f, v)
        g.upload(f, v)
    assert o.init() == 0
    g.neighbors()
    for chunk in (1, 2, 3):
        o.step(chunk)
        g.step(chunk)
        n = o.n
        assert g.n() == n and g.id_max() == o.id_max
        for f in (A.F_ID, A.F_LF):
            assert np.array_equal(g.download(f, n), o.get(f))
        for f in (A.F_U0, A.F_U1, A.F_U2, A.F_U3):
            assert np.array_equal(g.download(f, n), o.get(f))
        for f in (A.F_X, A.F_Y):
            assert rel_err(g.download(f, n), o.get(f)) <= 1e-12
    assert o.n > n0


def test_eight_user_parameters(gpu_available):
    # a rule reading and writing U4..U7 (neighbors' u(4) through the loop's
    # view, daughters with their own u[7]) equals its oracle restatement bit
    # for bit
    n0, L = 200, 25.0
    m = model(2, integrator=A.EULER, neighbors=A.CELL_LINKED, law=A.LAW_LINEAR, rule=A.RULE_CUSTOM, dt=0.02,
              cell_edge=2.0)
    s = sim(n0, [[0.0, L], [0.0, L]], [1.0, 2.0], n_max=400)
    pos = uniform_positions(n0, [[0.0, L], [0.0, L]], 78)
    rng = np.random.default_rng(2)
    fields = [(A.F_X, pos[0]), (A.F_Y, pos[1]), (A.F_U4, rng.uniform(0.0, 1.0, n0)), (A.F_U5, np.ones(n0)),
              (A.F_U2, rng.uniform(0.0, 1.0, n0))]
    o = O.Oracle(m, s)
    o.set_rule(RB.eight)
    g = Context(m, [s])
    g.compile_model(rule=RB.EIGHT)
    for f, v in fields:
        o.set(f, v)
        g.upload(f, v)
    assert o.init() == 0
    g.neighbors()
    for chunk in (1, 2):
        o.step(chunk)
        g.step(chunk)
        n = o.n
        assert g.n() == n
        assert np.array_equal(g.download(A.F_ID, n), o.get(A.F_ID))
        for k in range(8):
            f = A.user_field(k)
            assert np.array_equal(g.download(f, n), o.get(f)), k
        for f in (A.F_X, A.F_Y):
            assert rel_err(g.download(f, n), o.get(f)) <= 1e-12
    assert o.n > n0 and np.all(o.get(A.F_U5)[:n0] == 8.0)
