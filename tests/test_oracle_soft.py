"""Oracle checks of the inertial soft-sphere model (src/models/softSpheres.jl:1-81):
closed forms for a free damped particle, the sign and range of the pair force
(repulsive below r_i + r_j, attractive up to mu (r_i + r_j), zero beyond), the
fate matrix f0 picked per ordered pair, and momentum under pair forces."""
import numpy as np

import cbm._abi as A
import oracle as O
from helpers import model, sim

B, MU = 0.5, 1.4
F0 = [2.0, 1.0, 0.5, 3.0]      # f0[1][1], f0[1][2], f0[2][1], f0[2][2]


def soft(n, dims, box, integ=A.EULER, dt=0.01, nb=A.FULL, **kw):
    m = model(dims, integrator=integ, neighbors=nb, law=A.LAW_SOFT_SPHERES, dt=dt, **kw)
    return O.Oracle(m, sim(n, box, [B, MU] + F0))


def load(o, pos, vel=None, r=None, mass=None, fate=None):
    n = len(pos[0])
    for a, v in enumerate(pos):
        o.set(a, v)
        o.set(A.F_VX + a, np.zeros(n) if vel is None else vel[a])
    o.set(A.F_RADIUS, np.ones(n) if r is None else r)
    o.set(A.F_MASS, np.ones(n) if mass is None else mass)
    o.set(A.F_FATE, np.ones(n, dtype=np.int64) if fate is None else np.asarray(fate, dtype=np.int64))
    assert o.init() == 0


def test_free_damped_particle_closed_forms():
    # no neighbour within mu (r_i + r_j): dv/dt = -b v / m, dx/dt = v
    m_, v0, x0, dt = 2.0, 1.5, 0.25, 0.1
    for integ, vf, xf in [
            (A.EULER, v0 * (1 - B * dt / m_), x0 + v0 * dt),
            (A.HEUN, v0 * (1 - B * dt / m_ + (B * dt / m_) ** 2 / 2), x0 + v0 * dt * (1 - B * dt / m_ / 2))]:
        o = soft(2, 2, [[-50.0, 50.0]] * 2, integ=integ, dt=dt)
        load(o, [np.array([x0, 30.0]), np.array([0.0, 0.0])], vel=[np.array([v0, 0.0]), np.zeros(2)],
             mass=np.array([m_, m_]))
        o.step(1)
        assert np.isclose(o.get(A.F_VX)[0], vf, rtol=1e-15)
        assert np.isclose(o.get(A.F_X)[0], xf, rtol=1e-15)


def pair_accel(d, r=(1.0, 1.0), fate=(1, 1)):
    o = soft(2, 2, [[-50.0, 50.0]] * 2, dt=1e-3)
    load(o, [np.array([0.0, d]), np.zeros(2)], r=np.array(r), fate=fate)
    o.step(1)
    return o.get(A.F_VX) / 1e-3      # a = F / m with m = 1, v0 = 0


def test_pair_force_sign_and_range():
    rij = 2.0
    a = pair_accel(0.8 * rij)                    # overlap: pushed apart
    assert a[0] < 0 < a[1] and a[0] == -a[1]
    a = pair_accel(0.5 * (rij + MU * rij))       # between r_ij and mu r_ij: pulled together
    assert a[0] > 0 > a[1] and a[0] == -a[1]
    assert np.all(pair_accel(rij) == 0.0)        # the zero at d = r_ij
    assert np.all(pair_accel(MU * rij * 1.0001) == 0.0)   # out of range
    a = pair_accel(1.0, r=(0.5, 0.3))            # r_ij = 0.8 < d < mu r_ij = 1.12
    assert a[0] > 0 > a[1]


def test_fate_matrix_is_read_per_ordered_pair():
    d = 1.5
    base = pair_accel(d, fate=(1, 1))[0]          # f0[1][1] = 2
    a12 = pair_accel(d, fate=(1, 2))              # agent 0 reads f0[1][2] = 1, agent 1 f0[2][1] = 0.5
    assert np.isclose(a12[0], base * F0[1] / F0[0], rtol=1e-14)
    assert np.isclose(a12[1], -base * F0[2] / F0[0], rtol=1e-14)
    a22 = pair_accel(d, fate=(2, 2))
    assert np.isclose(a22[0], base * F0[3] / F0[0], rtol=1e-14)


def test_momentum_decays_only_by_friction():
    # symmetric f0 so pair forces cancel: sum m v follows v (1 - b dt / m) per Euler step
    rng = np.random.default_rng(4)
    n, L, dt = 400, 20.0, 0.01
    m = model(2, integrator=A.EULER, neighbors=A.CELL_LINKED, law=A.LAW_SOFT_SPHERES, dt=dt, cell_edge=3.0)
    o = O.Oracle(m, sim(n, [[0.0, L]] * 2, [B, MU, 2.0, 1.0, 1.0, 2.0]))
    vel = [rng.normal(0, 1, n), rng.normal(0, 1, n)]
    load(o, [rng.uniform(0, L, n), rng.uniform(0, L, n)], vel=vel, r=rng.uniform(0.8, 1.2, n),
         fate=rng.integers(1, 3, n))
    p0 = np.array([v.sum() for v in vel])
    o.step(1)
    p1 = np.array([o.get(A.F_VX + a).sum() for a in range(2)])
    assert np.allclose(p1, p0 * (1 - B * dt), rtol=0, atol=1e-10)
