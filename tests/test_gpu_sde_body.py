"""User agentSDE bodies compiled with hipRTC (cbm_compile_model): a body
restating the constant noise equals the built-in path bit for bit, and a
state-dependent g(x) = g0 + s x matches the oracle's affine noise (EM and
EulerHeun, Verlet 3-D and cell-linked 2-D), at 1e-12 per step and 1e-9 after
50 steps (the normals go through the device's log/sin/cos)."""
import numpy as np
import pytest

import cbm._abi as A
import oracle as O
from cbm.api import ABM, CBMIntegrators, CBMModels, CBMNeighbors, Community, Context, evolve_
from helpers import model, rel_err, sim, uniform_positions

pytestmark = pytest.mark.gpu

REP_ADH = [1.0, 1.0, 0.1, 2.4, 0.0, 0.0, 0.0, 0.0]


def noise_body(dims):
    # g of the positions from p[6] + p[7] * x; nothing on other variables
    return """
struct UserNoise {
  double g0, g1;
  __device__ explicit UserNoise(const double* p) : g0(p[6]), g1(p[7]) {}
  __device__ double operator()(int v, const double* u) const { return v < %d ? g0 + g1 * u[v] : 0.0; }
};
""" % dims


def run(m, s, pos, steps, body=None):
    g = Context(m, [s])
    if body is not None:
        g.compile_model(None, body)
    for a, v in enumerate(pos):
        g.upload(a, v)
    g.neighbors()
    g.step(steps)
    return [g.download(a, s.n) for a in range(len(pos))]


@pytest.mark.parametrize("integ", [A.EM, A.EULER_HEUN])
def test_constant_noise_body_equals_builtin(gpu_available, integ):
    n, L = 2000, 30.0
    box = [[0.0, L]] * 3
    pos = uniform_positions(n, box, 3)
    m = model(3, integrator=integ, neighbors=A.VERLET_DISPLACEMENT, law=A.LAW_REP_ADH, dt=0.005, periodic=1,
              skin=2.7, max_neighbors=64, rebuild_disp=0.15, noise=0.2, seed=9)
    p = list(REP_ADH)
    p[6] = 0.2
    builtin = run(m, sim(n, box, p), pos, 30)
    user = run(m, sim(n, box, p), pos, 30, noise_body(3))
    for a in range(3):
        assert np.array_equal(builtin[a], user[a])


@pytest.mark.parametrize("integ", [A.EM, A.EULER_HEUN])
@pytest.mark.parametrize("dims,nb", [(3, A.VERLET_DISPLACEMENT), (2, A.CELL_LINKED)])
def test_state_dependent_noise_matches_oracle(gpu_available, integ, dims, nb):
    n = 1500
    L = 30.0 if dims == 3 else 60.0
    box = [[0.0, L]] * dims
    pos = uniform_positions(n, box, 4 + dims)
    kw = dict(skin=2.7, max_neighbors=64, rebuild_disp=0.15) if nb == A.VERLET_DISPLACEMENT else dict(cell_edge=2.4)
    g0, slope = 0.05, 0.01
    m = model(dims, integrator=integ, neighbors=nb, law=A.LAW_REP_ADH, dt=0.005, periodic=1, noise=g0, seed=17,
              **kw)
    p = list(REP_ADH)
    p[6], p[7] = g0, slope
    s = sim(n, box, p)
    o = O.Oracle(m, s)
    o.set_noise_slope(slope)
    g = Context(m, [s])
    g.compile_model(None, noise_body(dims))
    for a, v in enumerate(pos):
        o.set(a, v)
        g.upload(a, v)
    assert o.init() == 0
    g.neighbors()
    o.step(1)
    g.step(1)
    assert max(rel_err(g.download(a, n), o.get(a)) for a in range(dims)) <= 1e-12
    o.step(49)
    g.step(49)
    assert max(rel_err(g.download(a, n), o.get(a)) for a in range(dims)) <= 1e-9


def test_sde_body_errors(gpu_available):
    box = [[0.0, 20.0]] * 2
    euler = model(2, integrator=A.EULER, neighbors=A.FULL, law=A.LAW_LINEAR, dt=0.01)
    with pytest.raises(A.CBMError):
        Context(euler, [sim(10, box, [1.0, 2.0])]).compile_model(None, noise_body(2))
    em = model(2, integrator=A.EM, neighbors=A.FULL, law=A.LAW_LINEAR, dt=0.01, noise=0.1)
    with pytest.raises(A.CBMError):       # a law body needs law CBM_LAW_CUSTOM
        Context(em, [sim(10, box, [1.0, 2.0])]).compile_model("struct UserLaw {};", noise_body(2))
    with pytest.raises(A.CBMError) as e:  # hipRTC's log comes back in the error
        Context(em, [sim(10, box, [1.0, 2.0])]).compile_model(None, "struct UserNoise { syntax error };")
    assert "hipRTC" in str(e.value)


def test_sde_body_python_mirror(gpu_available):
    n, L = 300, 40.0
    pos = uniform_positions(n, [[0.0, L]] * 2, 8)
    abm = ABM(2, agentODE=CBMModels.linearRepulsion(1.0, 2.0), agentSDE=noise_body(2),
              neighborsAlg=CBMNeighbors.CellLinked(cellEdge=2.0), agentAlg=CBMIntegrators.EM(), periodic=True,
              seed=5)
    com = Community(abm, N=n, dt=0.01, simBox=[[0.0, L], [0.0, L]], x=pos[0], y=pos[1],
                    lawParameters=[1.0, 2.0, 0.0, 0.0, 0.0, 0.0, 0.1, 0.02])
    evolve_(com, steps=20)
    m = model(2, integrator=A.EM, neighbors=A.CELL_LINKED, law=A.LAW_LINEAR, dt=0.01, periodic=1, cell_edge=2.0,
              noise=0.1, seed=5)
    o = O.Oracle(m, sim(n, [[0.0, L]] * 2, [1.0, 2.0]))
    o.set_noise_slope(0.02)
    o.set(A.F_X, pos[0])
    o.set(A.F_Y, pos[1])
    assert o.init() == 0
    o.step(20)
    assert rel_err(com.x, o.get(A.F_X)) <= 1e-11
