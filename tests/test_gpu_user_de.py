"""User DE variables of an agentODE body (UserLaw::de, the dt(u) lines of
src/AgentStructure/functionDE.jl:13-116): integrated beside the positions
by the model's integrator.  A variable with dt(u) = -a u started at x equals
the position under drift a bit for bit, and a variable coupled to the
positions, dt(u) = x, follows the reference's Euler, Heun and RK4 updates
(src/integrators.jl:496-498, :551-566, :620-661) restated below in numpy."""
import numpy as np
import pytest

import cbm._abi as A
from cbm.api import Context
from helpers import model, sim, uniform_positions

pytestmark = pytest.mark.gpu

BODY = """
struct UserLaw {
  static constexpr bool active = true, per_axis_division = false, walls = false;
  double a;
  __device__ explicit UserLaw(const double* p) : a(p[0]) {}
  __device__ bool operator()(double, double& s) const { s = 0.0; return false; }
  __device__ void de(const double* x, const double* u, double* du) const {
    du[0] = x[0];
    du[1] = -a * u[1];
  }
};
"""
NOOP_RULE = ("struct UserRule { __device__ UserRule(const double*, const double*) {} "
             "__device__ void operator()(RuleAgent&) const {} };")


def restated(integ, x, u, a, dt, steps):
    def f(x, u):
        return 0.0 - a * x, x
    for _ in range(steps):
        kx, ku = f(x, u)
        if integ == A.EULER:
            x, u = x + kx * dt, u + ku * dt
        elif integ == A.HEUN:
            hx, hu = x + kx * dt, u + ku * dt
            lx, lu = f(hx, hu)
            x, u = (hx + (x + lx * dt)) / 2, (hu + (u + lu * dt)) / 2
        else:
            ax, au = kx, ku
            k2 = f(x + kx * dt / 2, u + ku * dt / 2)
            ax, au = ax + 2 * k2[0], au + 2 * k2[1]
            k3 = f(x + k2[0] * dt / 2, u + k2[1] * dt / 2)
            ax, au = ax + 2 * k3[0], au + 2 * k3[1]
            k4 = f(x + k3[0] * dt, u + k3[1] * dt)
            x, u = x + (ax + k4[0]) * dt / 6, u + (au + k4[1]) * dt / 6
    return x, u


@pytest.mark.parametrize("rule", [False, True])
@pytest.mark.parametrize("integ", [A.EULER, A.HEUN, A.RK4])
def test_user_de_variables(gpu_available, integ, rule):
    n, a, dt, steps = 700, 0.7, 0.05, 40
    box = [[0.0, 10.0]] * 3
    m = model(3, integrator=integ, neighbors=A.FULL, law=A.LAW_CUSTOM, dt=dt, drift=a,
              rule=A.RULE_CUSTOM if rule else A.RULE_NONE)
    m.user_de = 2
    g = Context(m, [sim(n, box, [a])])
    g.compile_model(BODY, None, NOOP_RULE if rule else None)
    pos = uniform_positions(n, box, 5)
    u0 = np.linspace(-1.0, 2.0, n)
    for c, v in enumerate(pos):
        g.upload(c, v)
    g.upload(A.F_U0, u0)
    g.upload(A.F_U1, pos[0])
    g.neighbors()
    g.step(steps)
    x = [g.download(c, n) for c in range(3)]
    u = g.download(A.F_U0, n), g.download(A.F_U1, n)
    assert np.array_equal(u[1], x[0])
    for c in range(3):
        xe, _ = restated(integ, pos[c], pos[c], a, dt, steps)
        assert np.array_equal(x[c], xe)
    _, ue = restated(integ, pos[0], u0, a, dt, steps)
    assert np.array_equal(u[0], ue)


def test_user_de_needs_de_in_the_body(gpu_available):
    m = model(2, neighbors=A.FULL, law=A.LAW_CUSTOM)
    m.user_de = 1
    g = Context(m, [sim(4, [[0, 8]] * 2, [1.0])])
    no_de = BODY.replace("de(const double* x", "other(const double* x")
    with pytest.raises(RuntimeError, match="UserLaw::de"):
        g.compile_model(no_de)
