"""User DE variables of new agents.  An agent added by a rule starts with the
values @addAgent gives it (src/AgentStructure/functionRule.jl:9-79: the
parent's for every parameter it does not set) and is first integrated in
the next step, whatever buffer the integrator writes: Euler and EM step
out of place and flip the DE state, user DE variables included.

The law's user variable has dt(u) = -a u, so one step multiplies it by
(1 - a dt) exactly as restated below; new agents join that recurrence one
step late."""
import numpy as np
import pytest

import cbm._abi as A
from cbm.api import Context
from helpers import model, sim, uniform_positions

pytestmark = pytest.mark.gpu

LAW = """
struct UserLaw {
  static constexpr bool active = true, per_axis_division = false, walls = false;
  double a;
  __device__ explicit UserLaw(const double* p) : a(p[0]) {}
  __device__ bool operator()(double, double& s) const { s = 0.0; return false; }
  __device__ void de(const double*, const double* u, double* du) const { du[0] = -a * u[0]; }
};
"""
# every even-id agent divides once (lf marks the agents that have divided):
# the daughter gets u = 2 u_parent and x shifted by 1/4
DIVIDE = """
struct UserRule {
  __device__ UserRule(const double*, const double*) {}
  __device__ void operator()(RuleAgent& a) const {
    if (a.old.lf == 0 && a.id % 2 == 0) {
      RuleRec d = a.child();
      d.x[0] = a.old.x[0] + 0.25;
      d.u[0] = 2.0 * a.old.u[0];
      d.lf = 1;
      a.add(d);
      a.nw.lf = 1;
    }
  }
};
"""


def decay(u, a, dt, steps, integ=A.EULER):
    """dt(u) = -a u stepped by Euler / EM (no noise off the positions) or
    Heun (src/integrators.jl:496-498, :724-727, :551-566)."""
    for _ in range(steps):
        if integ == A.HEUN:
            h = u + (0.0 - a * u) * dt
            u = (h + (u + (0.0 - a * h) * dt)) / 2
        else:
            u = u + (0.0 - a * u) * dt
    return u


@pytest.mark.parametrize("integ", [A.EULER, A.EM, A.HEUN], ids=["euler", "em", "heun"])
def test_daughters_of_a_user_rule_carry_their_user_de_values(gpu_available, integ):
    n, a, dt, steps = 300, 0.7, 0.05, 6
    box = [[0.0, 10.0]] * 2
    m = model(2, integrator=integ, neighbors=A.FULL, law=A.LAW_CUSTOM, dt=dt, rule=A.RULE_CUSTOM, seed=4)
    m.user_de = 1
    g = Context(m, [sim(n, box, [a], n_max=2 * n)])
    g.compile_model(LAW, None, DIVIDE)
    pos = uniform_positions(n, box, 8)
    u0 = np.linspace(0.5, 3.0, n)
    for c, v in enumerate(pos):
        g.upload(c, v)
    g.upload(A.F_U0, u0)
    g.upload(A.F_LF, np.zeros(n, dtype=np.int64))
    g.neighbors()
    g.step(steps)
    n2 = g.n()
    assert n2 == n + n // 2
    ids = g.download(A.F_ID, n2)
    u = g.download(A.F_U0, n2)
    x = g.download(A.F_X, n2)
    # parents: slots 0..n-1, integrated every step
    assert np.array_equal(ids[:n], np.arange(1, n + 1))
    assert np.array_equal(u[:n], decay(u0, a, dt, steps, integ))
    # daughters: slot n + k for the k-th even id, added in the first step
    # with 2 u_parent at its start, then integrated steps - 1 times
    par = np.arange(1, n, 2)
    assert np.array_equal(ids[n:], n + 1 + np.arange(n // 2))
    assert np.array_equal(u[n:], decay(2.0 * u0[par], a, dt, steps - 1, integ))
    assert np.array_equal(x[n:], pos[0][par] + 0.25)


@pytest.mark.parametrize("integ", [A.EULER, A.EM], ids=["euler", "em"])
def test_growth_division_daughters_inherit_user_de_values(gpu_available, integ):
    """The built-in growth/division rule copies the parent's user variables
    into the daughter; with dt(u) = 0 every agent keeps its founder's u."""
    n, dt = 200, 0.1
    box = [[0.0, 40.0]] * 2
    m = model(2, integrator=integ, neighbors=A.FULL, law=A.LAW_CUSTOM, dt=dt,
              rule=A.RULE_GROWTH_DIVISION, rule_p=(0.5, 0.1, 0.3, 2.0), seed=6)
    m.user_de = 1
    g = Context(m, [sim(n, box, [0.0], n_max=8 * n)])
    g.compile_model(LAW)
    pos = uniform_positions(n, box, 9)
    for c, v in enumerate(pos):
        g.upload(c, v)
    founder = np.arange(1, n + 1, dtype=np.int64)
    g.upload(A.F_LF, founder)
    g.upload(A.F_U0, 1.5 * founder)
    g.upload(A.F_VOL, np.linspace(1.0, 1.9, n))
    g.upload(A.F_VBIRTH, np.ones(n))
    g.upload(A.F_GROWTH, np.full(n, 0.5))
    g.neighbors()
    g.step(30)
    n2 = g.n()
    assert n2 > 2 * n
    lf = g.download(A.F_LF, n2)
    u = g.download(A.F_U0, n2)
    assert np.array_equal(u, 1.5 * lf)
