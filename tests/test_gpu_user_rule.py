"""User agentRule bodies compiled with hipRTC (cbm_compile_model rule_body)
against the oracle's CBM_RULE_CUSTOM path running the same rule restated in
Python: the add/remove KAT, the random-distribution KAT, and a colony whose
agents are replaced by two new ones on a random axis (examples/Bacteries
agentRule) under repulsion and growth.  Integer fields (ids, lf, N) are
compared exactly; positions at 1e-9 (the rule's cos/sin and the normals go
through the device's math library)."""
import numpy as np
import pytest

import cbm._abi as A
import oracle as O
import rule_bodies as RB
from cbm.api import ABM, CBMIntegrators, CBMModels, CBMNeighbors, Community, Context, evolve_
from helpers import model, rel_err, sim, uniform_positions

pytestmark = pytest.mark.gpu


def pair(m, s, fields, body_src, body_py):
    o = O.Oracle(m, s)
    o.set_rule(body_py)
    g = Context(m, [s])
    g.compile_model(rule=body_src)
    for f, v in fields.items():
        o.set(f, v)
        g.upload(f, v)
    assert o.init() == 0
    g.neighbors()
    return o, g


def test_user_parity_rule_kat_and_builtin(gpu_available):
    N = 100
    m = model(1, neighbors=A.FULL, rule=A.RULE_CUSTOM)
    s = sim(N, [[0.0, 1.0]], n_max=N + 50)
    o, g = pair(m, s, {A.F_X: np.zeros(N)}, RB.PARITY, RB.parity)
    g.step(1)
    o.step(1)
    assert g.n() == o.n == 100 and g.id_max() == o.id_max == 150
    for f in (A.F_ID, A.F_LF, A.F_X):
        assert np.array_equal(g.download(f, N), o.get(f))
    assert set(range(2, 101, 2)) <= set(g.download(A.F_LF, N).tolist())
    # the same rule as the built-in CBM_RULE_PARITY kernel, over several steps
    n0 = 40
    pos = uniform_positions(n0, [[0.0, 10.0]], 3)
    got = []
    for rule in (A.RULE_PARITY, A.RULE_CUSTOM):
        m = model(1, neighbors=A.FULL, rule=rule, law=A.LAW_LINEAR, dt=0.01)
        c = Context(m, [sim(n0, [[0.0, 10.0]], [1.0, 2.0], n_max=4 * n0)])
        if rule == A.RULE_CUSTOM:
            c.compile_model(rule=RB.PARITY)
        c.upload(A.F_X, pos)
        c.neighbors()
        c.step(3)
        n = c.n()
        got.append((n, c.id_max(), c.download(A.F_X, n), c.download(A.F_ID, n), c.download(A.F_LF, n)))
    assert got[0][:2] == got[1][:2]
    for a, b in zip(got[0][2:], got[1][2:]):
        assert np.array_equal(a, b)


def test_user_random_rule_moments_and_oracle(gpu_available):
    # testCommunity.jl:319-356 on N = 100,000 with its tolerances
    N = 100000
    m = model(1, neighbors=A.FULL, rule=A.RULE_CUSTOM, seed=2024)
    g = Context(m, [sim(N, [[0.0, 1.0]])])
    g.compile_model(rule=RB.RANDOM)
    g.neighbors()
    g.step(1)
    nrm, unif, expo = (g.download(f, N) for f in (A.F_U0, A.F_U1, A.F_U2))
    assert abs(nrm.mean() - 1) < 0.05
    assert abs(nrm.var() - 4) < 0.05
    assert abs(unif.min() - 1) < 0.05 and abs(unif.max() - 2) < 0.05
    assert abs(expo.mean() - 3) < 0.05
    # the first 2,000 agents against the oracle's restatement: same draws
    n = 2000
    o = O.Oracle(m, sim(n, [[0.0, 1.0]]))
    o.set_rule(RB.random_draws)
    assert o.init() == 0
    o.step(1)
    assert np.array_equal(unif[:n], o.get(A.F_U1))
    assert rel_err(nrm[:n], o.get(A.F_U0)) <= 1e-12
    assert rel_err(expo[:n], o.get(A.F_U2)) <= 1e-12


@pytest.mark.parametrize("nb", [A.CELL_LINKED, A.VERLET_DISPLACEMENT])
def test_user_divide_rule_colony_matches_oracle(gpu_available, nb):
    # 60 agents grow (rate 0.5 +- 20 %) and split in two at twice their birth volume
    n0, L = 60, 40.0
    rp = (2.0, 0.5, 0.5, 0.2)
    kw = dict(cell_edge=2.0) if nb == A.CELL_LINKED else dict(skin=2.4, max_neighbors=64, rebuild_disp=0.2)
    m = model(2, integrator=A.EULER, neighbors=nb, law=A.LAW_LINEAR, rule=A.RULE_CUSTOM, rule_p=rp, dt=0.05,
              periodic=1, seed=77, **kw)
    s = sim(n0, [[0.0, L], [0.0, L]], [1.0, 2.0], n_max=1200)
    pos = uniform_positions(n0, [[10.0, 30.0], [10.0, 30.0]], 5)
    rng = np.random.default_rng(6)
    fields = {A.F_X: pos[0], A.F_Y: pos[1], A.F_VOL: rng.uniform(1.0, 2.0, n0), A.F_VBIRTH: np.ones(n0),
              A.F_GROWTH: rng.uniform(0.4, 0.6, n0)}
    o, g = pair(m, s, fields, RB.DIVIDE_2D, RB.divide_2d(rp))
    for chunk in (1, 9, 30):
        o.step(chunk)
        g.step(chunk)
        n = o.n
        assert g.n() == n and g.id_max() == o.id_max
        for f in (A.F_ID, A.F_LF):
            assert np.array_equal(g.download(f, n), o.get(f))
        assert np.array_equal(g.download(A.F_U0, n), o.get(A.F_U0))
        for f in (A.F_X, A.F_Y, A.F_VOL, A.F_GROWTH, A.F_VBIRTH):
            assert rel_err(g.download(f, n), o.get(f)) <= 1e-9
    assert o.n > 2 * n0                     # every agent divided, some twice


def test_user_rule_errors(gpu_available):
    box = [[0.0, 10.0]]
    m = model(1, neighbors=A.FULL, rule=A.RULE_CUSTOM)
    g = Context(m, [sim(4, box, n_max=20)])
    g.neighbors()
    with pytest.raises(A.CBMError):          # no rule body compiled yet
        g.step(1)
    with pytest.raises(A.CBMError):          # a rule body with a built-in rule
        Context(model(1, neighbors=A.FULL, rule=A.RULE_DEATH, rule_p=(0.1, 0, 0, 0)),
                [sim(4, box)]).compile_model(rule=RB.PARITY)
    g = Context(m, [sim(4, box, n_max=20)])
    g.compile_model(rule=RB.THREE_ADDS)
    g.neighbors()
    g.step(1)
    with pytest.raises(A.CBMError) as e:     # three additions by one agent, reported at the next sync
        g.sync()
    assert e.value.code == A.ERR_CONFIG and "more than 2" in str(e.value)
    g = Context(m, [sim(4, box, n_max=5)])
    g.compile_model(rule=RB.PARITY)
    g.neighbors()
    g.step(1)
    with pytest.raises(A.CBMError) as e:     # ids 2 and 4 add one each: N + 2 = 6 > 5
        g.sync()
    assert e.value.code == A.ERR_CAPACITY


def test_user_rule_python_mirror(gpu_available):
    body = """
struct UserRule {
  double ratio;
  __device__ UserRule(const double* rp, const double*) : ratio(rp[0]) {}
  __device__ void operator()(RuleAgent& a) const {
    a.nw.u[agent::age] = a.old.u[agent::age] + a.dt;
    if (a.old.u[agent::age] >= ratio) {
      RuleRec d = a.child();
      d.u[agent::age] = 0.0;
      d.u[agent::gen] = a.old.u[agent::gen] + 1.0;
      a.add(d);
      a.nw.u[agent::age] = 0.0;
    }
  }
};
"""
    n = 50
    pos = uniform_positions(n, [[0.0, 30.0]] * 2, 2)
    abm = ABM(2, agentODE=CBMModels.linearRepulsion(1.0, 2.0), agentRule=(body, [0.25, 0, 0, 0]),
              agent={"age": float, "gen": float}, neighborsAlg=CBMNeighbors.CellLinked(cellEdge=2.0),
              agentAlg=CBMIntegrators.Euler(), periodic=True)
    com = Community(abm, N=n, dt=0.1, simBox=[[0.0, 30.0], [0.0, 30.0]], x=pos[0], y=pos[1])
    com.age = np.linspace(0.0, 0.2, n)
    evolve_(com, steps=4, preallocateAgents=200)
    # ages below 0.25 take ceil((0.25 - age0)/0.1) steps to reach it, then the agent adds one
    assert com.N > n
    assert set(np.unique(com.gen).tolist()) <= {0.0, 1.0}
    assert np.count_nonzero(com.gen == 1.0) == com.N - n
