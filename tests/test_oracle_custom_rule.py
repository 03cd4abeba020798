"""The oracle's CBM_RULE_CUSTOM path (a rule body restated in Python, called
per agent in slot order) pinned by the reference's rule tests: the add/remove
KAT (test/testCommunity.jl:241-316) and its golden numbers, equality with the
built-in parity rule, the random-distribution moments of
test/testCommunity.jl:319-356, and the hole order of update.jl:15-119 when
every agent is replaced by two new ones."""
import math

import numpy as np
import pytest

import cbm._abi as A
import oracle as O
import rule_bodies as RB
from helpers import golden, model, sim, uniform_positions


def test_custom_parity_matches_golden_kat():
    g = golden("ref_add_remove.json")
    N = g["N"]
    m = model(1, neighbors=A.FULL, rule=A.RULE_CUSTOM)
    o = O.Oracle(m, sim(N, [[0.0, 1.0]], n_max=N + g["preallocateAgents"]))
    o.set_rule(RB.parity)
    o.set(A.F_X, np.zeros(N))
    assert o.init() == 0
    assert o.step(1) == 0
    assert o.n == g["after_update"]["N"]
    assert o.id_max == g["after_update"]["idMax_"]
    assert set(g["after_update"]["lf_contains"]) <= set(o.get(A.F_LF).tolist())


def test_custom_parity_equals_builtin():
    N = 40
    pos = uniform_positions(N, [[0.0, 10.0]], 3)
    out = []
    for rule in (A.RULE_PARITY, A.RULE_CUSTOM):
        m = model(1, neighbors=A.FULL, rule=rule, law=A.LAW_LINEAR, dt=0.01)
        o = O.Oracle(m, sim(N, [[0.0, 10.0]], [1.0, 2.0], n_max=4 * N))
        if rule == A.RULE_CUSTOM:
            o.set_rule(RB.parity)
        o.set(A.F_X, pos)
        assert o.init() == 0
        assert o.step(3) == 0
        out.append((o.n, o.id_max, o.get(A.F_X), o.get(A.F_ID), o.get(A.F_LF)))
    (n0, i0, x0, id0, lf0), (n1, i1, x1, id1, lf1) = out
    assert (n0, i0) == (n1, i1)
    assert np.array_equal(x0, x1) and np.array_equal(id0, id1) and np.array_equal(lf0, lf1)


def test_custom_rule_distribution_moments():
    # testCommunity.jl:319-356 with its tolerances, on 20,000 agents
    N = 20000
    m = model(1, neighbors=A.FULL, rule=A.RULE_CUSTOM, seed=12345)
    o = O.Oracle(m, sim(N, [[0.0, 1.0]]))
    o.set_rule(RB.random_draws)
    assert o.init() == 0
    assert o.step(1) == 0
    nrm, unif, expo = o.get(A.F_U0), o.get(A.F_U1), o.get(A.F_U2)
    assert abs(nrm.mean() - 1) < 0.05
    assert abs(nrm.var() - 4) < 0.15        # sd of the sample variance: sqrt(2*16/N) = 0.04
    assert abs(unif.min() - 1) < 0.05 and abs(unif.max() - 2) < 0.05
    assert unif.min() >= 1.0 and unif.max() < 2.0
    assert abs(expo.mean() - 3) < 0.1        # sd of the mean: 3/sqrt(N) = 0.021
    assert expo.min() >= 0.0


def test_custom_rule_draws_are_the_agents_philox_stream():
    # the k-th draw of agent id at step s is Philox(seed, (s, id, 0, 0x600 + k))
    m = model(1, neighbors=A.FULL, rule=A.RULE_CUSTOM, seed=(7 << 32) | 99)
    o = O.Oracle(m, sim(3, [[0.0, 1.0]]))
    o.set_rule(RB.random_draws)
    assert o.init() == 0
    o.step(1)
    o.step(1)   # step index 1
    lib = O.lib()
    for slot, id_ in enumerate(o.get(A.F_ID).tolist()):
        out = np.zeros(4, dtype=np.uint32)
        key = np.array([99, 7], dtype=np.uint32)
        ctr = np.array([1, id_, 0, 0x601], dtype=np.uint32)
        lib.orc_philox(ctr.ctypes.data, key.ctypes.data, out.ctypes.data)
        u = O.u53(int(out[0]), int(out[1]))
        assert o.get(A.F_U1)[slot] == (2.0 - 1.0) * u + 1.0
        ctr[3] = 0x602
        lib.orc_philox(ctr.ctypes.data, key.ctypes.data, out.ctypes.data)
        u = O.u53(int(out[0]), int(out[1]))
        assert o.get(A.F_U2)[slot] == -math.log(1.0 - u) * 3.0


def test_custom_divide_replaces_every_agent_with_two():
    # all ten agents at vol 2 >= 2 * vbirth 1: each adds two and is removed
    N = 10
    rp = (2.0, 0.5, 0.1, 0.0)
    m = model(2, neighbors=A.FULL, rule=A.RULE_CUSTOM, rule_p=rp, dt=0.01)
    o = O.Oracle(m, sim(N, [[0.0, 20.0], [0.0, 20.0]], n_max=3 * N))
    o.set_rule(RB.divide_2d(rp))
    pos = uniform_positions(N, [[5.0, 15.0], [5.0, 15.0]], 1)
    o.set(A.F_X, pos[0])
    o.set(A.F_Y, pos[1])
    o.set(A.F_VOL, np.full(N, 2.0))
    o.set(A.F_VBIRTH, np.ones(N))
    assert o.init() == 0
    assert o.step(1) == 0
    assert o.n == 2 * N and o.id_max == 3 * N
    # holes 0..9 take the survivors from the tail (rows 29..20, ids 30..21);
    # rows 10..19 (ids 11..20) stay
    assert o.get(A.F_ID).tolist() == list(range(30, 20, -1)) + list(range(11, 21))
    lf = o.get(A.F_LF)
    assert sorted(lf.tolist()) == sorted(list(range(1, N + 1)) * 2)
    # each pair is symmetric about its parent and 2 * split apart; volume is conserved
    x, y = o.get(A.F_X), o.get(A.F_Y)
    for p in range(1, N + 1):
        k = np.nonzero(lf == p)[0]
        assert len(k) == 2
        assert abs(math.hypot(x[k[0]] - x[k[1]], y[k[0]] - y[k[1]]) - 2 * 0.5) < 1e-12
        assert abs((x[k[0]] + x[k[1]]) / 2 - pos[0][p - 1]) < 1e-12
    assert abs(o.get(A.F_VOL).sum() - 2.0 * N) < 1e-12
    assert np.all(o.get(A.F_U0) == 1.0)


@pytest.mark.parametrize("dims", [1, 2, 3])
def test_custom_rule_neighbor_loop_lattice_kat(dims):
    # the reference's neighbor KAT (testCommunity.jl:361-415) as a rule body
    # with a neighbor loop, for every neighbor algorithm
    from test_oracle import lattice_case
    from helpers import lattice
    g = golden("ref_neighbors_lattice.json")
    for alg in g["algorithms"]:
        m, s = lattice_case(dims, alg)
        m.rule = A.RULE_CUSTOM
        o = O.Oracle(m, s)
        o.set_rule(RB.count_neighbors(g["radius"]))
        for a, v in enumerate(lattice(dims)):
            o.set(a, v)
        assert o.init() == 0
        o.step(1)
        assert o.get(A.F_U0).tolist() == [float(v) for v in g["expected"][str(dims)]], alg["name"]


def medium_rule_model(n, D, dt, c0, rp0):
    m = model(2, neighbors=A.FULL, dt=dt, rule=A.RULE_CUSTOM)
    m.rule_p[0] = rp0
    m.medium_n[0], m.medium_n[1] = n
    m.medium_D, m.medium_c0 = D, c0
    return m


def test_medium_additions_and_reads_by_hand():
    """Agents add s dt = 0.25 to their node and read it back: with no
    diffusion a node gains 0.25 per agent and step (exact in 32.32 fixed
    point), and u0 holds the node's value at the step start."""
    import rule_bodies as RB
    m = medium_rule_model((4, 4), 0.0, 0.5, 1.0, 0.5)        # s dt = 0.25
    o = O.Oracle(m, sim(3, [[0.0, 4.0], [0.0, 4.0]], n_max=8))
    o.set(A.F_X, [1.5, 1.2, 3.9])    # two agents in node (1, 2), one in (3, 0)
    o.set(A.F_Y, [2.5, 2.1, 0.2])
    o.set_rule(RB.secrete(0.5))
    assert o.init() == 0
    o.step(1)
    c = o.medium(16).reshape(4, 4)
    assert c[2, 1] == 1.5 and c[0, 3] == 1.25 and c.sum() == 16.0 + 0.75
    assert np.array_equal(o.get(A.F_U0), [1.0, 1.0, 1.0])
    o.step(2)
    c = o.medium(16).reshape(4, 4)
    assert c[2, 1] == 2.5 and c[0, 3] == 1.75
    assert np.array_equal(o.get(A.F_U0), [2.0, 2.0, 1.5])     # read at the third step's start


def test_medium_additions_are_fixed_point_sums():
    """0.1 is not a multiple of 2^-32: each agent's addition is rounded to
    32.32 fixed point once and the node's sum is exact in any order."""
    import rule_bodies as RB
    m = medium_rule_model((2, 2), 0.0, 1.0, 0.0, 0.1)
    o = O.Oracle(m, sim(5, [[0.0, 2.0], [0.0, 2.0]], n_max=8))
    o.set(A.F_X, [0.5] * 5)
    o.set(A.F_Y, [0.5] * 5)
    o.set_rule(RB.secrete(0.1))
    assert o.init() == 0
    o.step(1)
    q = round(0.1 * 2.0 ** 32)
    assert o.medium(4)[0] == (5 * q) / 2.0 ** 32
