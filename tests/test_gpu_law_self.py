"""Per-agent terms of a user agentODE body (UserLaw::self, compiled with
hipRTC): a body restating the linear law plus -k x equals the built-in law
with drift k bit for bit, and a harmonic trap around a point follows the
closed form of the explicit Euler step, x_t - c = (x_0 - c)(1 - k dt)^t."""
import numpy as np
import pytest

import cbm._abi as A
from cbm.api import Context
from helpers import model, rel_err, sim, uniform_positions

pytestmark = pytest.mark.gpu

LINEAR_DRIFT = """
struct UserLaw {
  static constexpr bool active = true, per_axis_division = false, walls = false;
  double k, c, a;
  __device__ explicit UserLaw(const double* p) : k(p[0]), c(p[1]), a(p[2]) {}
  __device__ bool operator()(double d, double& s) const {
    s = k * (c - d);
    return d > 0.0 && d < c;
  }
  __device__ void self(const double* x, double* f) const {
    for (int q = 0; q < 3; ++q) f[q] = -a * x[q];
  }
};
"""

TRAP = """
struct UserLaw {
  static constexpr bool active = true, per_axis_division = false, walls = false;
  double k, cx, cy;
  __device__ explicit UserLaw(const double* p) : k(p[0]), cx(p[1]), cy(p[2]) {}
  __device__ bool operator()(double, double& s) const { s = 0.0; return false; }
  __device__ void self(const double* x, double* f) const {
    f[0] = -k * (x[0] - cx);
    f[1] = -k * (x[1] - cy);
  }
};
"""


def test_self_term_restating_drift_is_bit_identical(gpu_available):
    n, L = 3000, 30.0
    box = [[0.0, L]] * 3
    pos = uniform_positions(n, box, 12)
    out = []
    for user in (False, True):
        m = model(3, integrator=A.HEUN, neighbors=A.VERLET_DISPLACEMENT, law=A.LAW_CUSTOM if user else A.LAW_LINEAR,
                  dt=0.01, skin=2.4, max_neighbors=64, rebuild_disp=0.2, drift=0.0 if user else 0.3)
        g = Context(m, [sim(n, box, [1.0, 2.0, 0.3])])
        if user:
            g.compile_model(LINEAR_DRIFT)
        for a, v in enumerate(pos):
            g.upload(a, v)
        g.neighbors()
        g.step(25)
        out.append([g.download(a, n) for a in range(3)])
    for a in range(3):
        assert np.array_equal(out[0][a], out[1][a])


def test_harmonic_trap_closed_form(gpu_available):
    n, k, dt, steps = 500, 2.0, 0.01, 40
    box = [[0.0, 20.0], [0.0, 20.0]]
    pos = uniform_positions(n, box, 3)
    m = model(2, integrator=A.EULER, neighbors=A.FULL, law=A.LAW_CUSTOM, dt=dt)
    g = Context(m, [sim(n, box, [k, 12.0, 7.0])])
    g.compile_model(TRAP)
    for a, v in enumerate(pos):
        g.upload(a, v)
    g.neighbors()
    g.step(steps)
    f = (1 - k * dt) ** steps
    for a, c in ((0, 12.0), (1, 7.0)):
        want = c + (pos[a] - c) * f
        assert rel_err(g.download(a, n), want) < 1e-13
