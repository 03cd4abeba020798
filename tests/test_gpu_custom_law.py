"""User interaction bodies compiled with hipRTC (cbm_compile) into the same
stage skeleton as the built-in laws."""
import numpy as np
import pytest

import cbm._abi as A
import oracle as O
from cbm.api import Context
from helpers import model, sim, uniform_positions

pytestmark = pytest.mark.gpu

# the LINEAR law restated as a user body
LINEAR_BODY = """
struct UserLaw {
  static constexpr bool active = true, per_axis_division = false, walls = false;
  double k, c;
  __device__ explicit UserLaw(const double* p) : k(p[0]), c(p[1]) {}
  __device__ bool operator()(double d, double& s) const {
    s = k * (c - d);
    return d > 0.0 && d < c;
  }
};
"""

# Hertz-like repulsion s = k (c - d)^(3/2)
HERTZ_BODY = """
struct UserLaw {
  static constexpr bool active = true, per_axis_division = false, walls = false;
  double k, c;
  __device__ explicit UserLaw(const double* p) : k(p[0]), c(p[1]) {}
  __device__ bool operator()(double d, double& s) const {
    const double o = c - d;
    s = k * o * sqrt(o > 0.0 ? o : 0.0);
    return d > 0.0 && d < c;
  }
};
"""


def test_user_linear_law_equals_builtin_and_oracle(gpu_available):
    n, L = 1000, 50.0
    box = [[0.0, L], [0.0, L]]
    pos = uniform_positions(n, box, 1)
    base = dict(integrator=A.EULER, neighbors=A.CELL_LINKED, dt=0.01, periodic=1, cell_edge=2.0)
    m_user = model(2, law=A.LAW_CUSTOM, **base)
    m_ref = model(2, law=A.LAW_LINEAR, **base)
    s = sim(n, box, [1.0, 2.0])
    g = Context(m_user, [s])
    g.compile(LINEAR_BODY)
    o = O.Oracle(m_ref, s)
    for a in range(2):
        g.upload(a, pos[a])
        o.set(a, pos[a])
    g.neighbors()
    o.init()
    g.step(50)
    o.step(50)
    for a in range(2):
        assert np.array_equal(g.download(a, n), o.get(a))


def test_user_hertz_law_two_body_closed_form(gpu_available):
    k, c, d0, dt = 2.0, 2.0, 1.5, 0.01
    m = model(1, integrator=A.EULER, neighbors=A.FULL, law=A.LAW_CUSTOM, dt=dt)
    g = Context(m, [sim(2, [[-10.0, 10.0]], [k, c])])
    g.compile(HERTZ_BODY)
    g.upload(A.F_X, [-d0 / 2, d0 / 2])
    g.neighbors()
    d = d0
    for _ in range(20):
        g.step(1)
        s = k * (c - d) * np.sqrt(c - d)
        d = d + 2 * s * dt
        x = g.download(A.F_X, 2)
        assert abs((x[1] - x[0]) - d) < 1e-13


def test_compile_errors_are_reported(gpu_available):
    m = model(2, integrator=A.EULER, neighbors=A.FULL, law=A.LAW_CUSTOM, dt=0.01)
    g = Context(m, [sim(4, [[0.0, 5.0]] * 2, [1.0, 2.0])])
    with pytest.raises(A.CBMError) as e:
        g.step(1)
    assert e.value.code == A.ERR_CONFIG
    with pytest.raises(A.CBMError) as e:
        g.compile("struct UserLaw { this is not C++ };")
    assert e.value.code == A.ERR_COMPILE
    assert "error" in str(e.value)
