"""Pins of the oracle's general agentODE body path (a body's neighbor sums
and equations, src/AgentStructure/functionDE.jl:13-116): the soft-sphere
model restated as a body over user parameters equals the oracle's built-in
soft-sphere law bit for bit under every walk and integrator; a crowding
variable fed by neighbor sums follows the closed-form one-step maps of
Euler, Heun and RK4 for its linear system (which fixes that neighbors are
read at the stage state); pair forces of a symmetric body leave the mean
position unchanged."""
import numpy as np
import pytest

import cbm._abi as A
import law_bodies as LB
import oracle as O
from helpers import model, sim

P_SOFT = [0.5, 1.4, 2.0, 1.0, 0.5, 3.0, 1.3]     # b, mu, f0 (2 x 2), m


def soft_agents(n, L, seed):
    rng = np.random.default_rng(seed)
    return (rng.uniform(0, L, n), rng.uniform(0, L, n), rng.normal(0, 0.5, n), rng.normal(0, 0.5, n),
            rng.uniform(0.8, 1.2, n), rng.integers(1, 3, n).astype(np.int64))


@pytest.mark.parametrize("neigh", [A.FULL, A.CELL_LINKED, A.VERLET_DISPLACEMENT])
@pytest.mark.parametrize("integ", [A.EULER, A.HEUN, A.RK4])
def test_soft_spheres_body_equals_the_builtin_law(neigh, integ):
    # Full walks every pair through the Python body: a smaller population
    n, L, steps = (80, 12.0, 10) if neigh == A.FULL else (150, 17.0, 8)
    kw = dict(integrator=integ, neighbors=neigh, dt=0.01, cell_edge=3.4, skin=3.8, max_neighbors=64,
              rebuild_disp=0.2)
    x, y, vx, vy, r, fate = soft_agents(n, L, 3)
    built = O.Oracle(model(2, law=A.LAW_SOFT_SPHERES, **kw), sim(n, [[0, L]] * 2, P_SOFT))
    for f, v in ((A.F_X, x), (A.F_Y, y), (A.F_VX, vx), (A.F_VY, vy), (A.F_RADIUS, r), (A.F_FATE, fate),
                 (A.F_MASS, np.full(n, P_SOFT[6]))):
        built.set(f, v)
    m = model(2, law=A.LAW_CUSTOM, **kw)
    m.user_de = 2
    body = O.Oracle(m, sim(n, [[0, L]] * 2, P_SOFT))
    for f, v in ((A.F_X, x), (A.F_Y, y), (A.F_U0, vx), (A.F_U1, vy), (A.F_U2, r), (A.F_U3, fate.astype(float))):
        body.set(f, v)
    body.set_law_body(*LB.soft2(P_SOFT))
    for o in (built, body):
        assert o.init() == 0
        o.step(steps)
        assert o.error() == 0
    assert np.array_equal(body.get(A.F_X), built.get(A.F_X))
    assert np.array_equal(body.get(A.F_Y), built.get(A.F_Y))
    assert np.array_equal(body.get(A.F_U0), built.get(A.F_VX))
    assert np.array_equal(body.get(A.F_U1), built.get(A.F_VY))
    assert not np.array_equal(body.get(A.F_U0), vx)


P_CROWD = [0.0, 1.0, 0.6]            # k (no forces: fixed positions), R, c


def two_crowded(integ, q0, steps, dt):
    m = model(2, integrator=integ, neighbors=A.FULL, law=A.LAW_CUSTOM, dt=dt)
    m.user_de = 1
    o = O.Oracle(m, sim(2, [[0, 10]] * 2, P_CROWD))
    o.set(A.F_X, [4.0, 4.9])
    o.set(A.F_Y, [5.0, 5.0])
    o.set(A.F_U0, q0)
    o.set(A.F_U1, [0.6, 0.9])            # reach R (0.6 + 0.9) = 1.5 at distance 0.9: w = 0.4
    o.set_law_body(*LB.crowd(P_CROWD))
    assert o.init() == 0
    o.step(steps)
    assert o.error() == 0
    return o


@pytest.mark.parametrize("integ", [A.EULER, A.HEUN, A.RK4])
def test_crowding_follows_the_linear_one_step_map(integ):
    # dq/dt = A q + 1 with A = [[-c, w], [w, -c]]; one step of each scheme is
    # q' = T(hA) q + h S(hA) 1 with T, S the scheme's polynomials
    w, c, dt, steps = 0.4, P_CROWD[2], 0.1, 7
    q0 = np.array([0.3, 2.0])
    Amat = np.array([[-c, w], [w, -c]])
    hA = dt * Amat
    Id = np.eye(2)
    if integ == A.EULER:
        T, S = Id + hA, Id
    elif integ == A.HEUN:
        T, S = Id + hA + hA @ hA / 2, Id + hA / 2
    else:
        hA2, hA3 = hA @ hA, hA @ hA @ hA
        T, S = Id + hA + hA2 / 2 + hA3 / 6 + hA3 @ hA / 24, Id + hA / 2 + hA2 / 6 + hA3 / 24
    q = q0.copy()
    for _ in range(steps):
        q = T @ q + dt * (S @ np.ones(2))
    o = two_crowded(integ, q0, steps, dt)
    assert np.allclose(o.get(A.F_U0), q, rtol=1e-14, atol=1e-14)
    assert np.array_equal(o.get(A.F_X), [4.0, 4.9])


def test_symmetric_pair_forces_keep_the_mean_position():
    n, L = 200, 12.0
    rng = np.random.default_rng(8)
    m = model(3, integrator=A.EULER, neighbors=A.CELL_LINKED, law=A.LAW_CUSTOM, dt=0.01, cell_edge=2.2)
    m.user_de = 1
    o = O.Oracle(m, sim(n, [[0, L]] * 3, [2.0, 1.0, 0.3]))
    x = [rng.uniform(3, L - 3, n) for _ in range(3)]
    for a in range(3):
        o.set(a, x[a])
    o.set(A.F_U1, rng.uniform(0.7, 1.1, n))
    o.set_law_body(*LB.crowd([2.0, 1.0, 0.3]))
    assert o.init() == 0
    o.step(10)
    for a in range(3):
        assert abs(o.get(a).mean() - x[a].mean()) < 1e-13
    assert np.std(o.get(A.F_X) - x[0]) > 1e-4
    assert np.all(o.get(A.F_U0) > 0)
