"""General agentODE bodies (UserLaw::pair + UserLaw::ode, hipRTC) on the
device.  The soft-sphere model of src/models/softSpheres.jl:19-36 restated
as a body over user parameters equals the built-in LAW_SOFT_SPHERES kernels
bit for bit (2-D CellLinked with mixed radii and fates, 3-D Verlet); a body
whose neighbor sums feed a non-position DE variable (the crowding level of
tests/law_bodies.py, reading the neighbors' values at the stage state)
equals the oracle running its CPU restatement bit for bit under Full,
CellLinked and Verlet and each integrator; graph replays equal eager steps;
lagged forces are refused."""
import numpy as np
import pytest

import cbm._abi as A
import law_bodies as LB
import oracle as O
from cbm.api import Context
from helpers import model, sim

pytestmark = pytest.mark.gpu

P_SOFT = [0.5, 1.4, 2.0, 1.0, 0.5, 3.0, 1.3]     # b, mu, f0 (2 x 2), m


def soft_agents(n, dims, L, seed):
    rng = np.random.default_rng(seed)
    return {"pos": [rng.uniform(0, L, n) for _ in range(dims)], "vel": [rng.normal(0, 0.5, n) for _ in range(dims)],
            "r": rng.uniform(0.8, 1.2, n), "fate": rng.integers(1, 3, n).astype(np.int64)}


def builtin_and_body(kw, dims, n, L, ag, body):
    b = Context(model(dims, law=A.LAW_SOFT_SPHERES, **kw), [sim(n, [[0, L]] * dims, P_SOFT)])
    for a in range(dims):
        b.upload(a, ag["pos"][a])
        b.upload(A.F_VX + a, ag["vel"][a])
    b.upload(A.F_RADIUS, ag["r"])
    b.upload(A.F_FATE, ag["fate"] if dims == 2 else np.ones(n, dtype=np.int64))
    b.upload(A.F_MASS, np.full(n, P_SOFT[6]))
    m = model(dims, law=A.LAW_CUSTOM, **kw)
    m.user_de = dims
    u = Context(m, [sim(n, [[0, L]] * dims, P_SOFT)])
    u.compile_model(body)
    for a in range(dims):
        u.upload(a, ag["pos"][a])
        u.upload(A.F_U0 + a, ag["vel"][a])
    u.upload(A.F_U0 + dims, ag["r"])
    if dims == 2:
        u.upload(A.F_U3, ag["fate"].astype(np.float64))
    b.neighbors()
    u.neighbors()
    return b, u


@pytest.mark.parametrize("integ", [A.EULER, A.HEUN, A.RK4])
def test_soft_spheres_body_2d_cell_linked_bit_exact(gpu_available, integ):
    n, L = 1500, 40.0
    kw = dict(integrator=integ, neighbors=A.CELL_LINKED, dt=0.01, cell_edge=3.4)
    b, u = builtin_and_body(kw, 2, n, L, soft_agents(n, 2, L, 3), LB.SOFT2)
    b.step(100)
    u.step(100)
    for a in range(2):
        assert np.array_equal(u.download(a, n), b.download(a, n)), a
        assert np.array_equal(u.download(A.F_U0 + a, n), b.download(A.F_VX + a, n)), a


def test_soft_spheres_body_3d_verlet_bit_exact(gpu_available):
    n = 3000
    L = (n * 4.0 / 3.0 * np.pi / 0.3) ** (1 / 3)
    kw = dict(integrator=A.HEUN, neighbors=A.VERLET_DISPLACEMENT, dt=0.005, skin=3.8, max_neighbors=64,
              rebuild_disp=0.2)
    b, u = builtin_and_body(kw, 3, n, L, soft_agents(n, 3, L, 5), LB.SOFT3)
    b.step(60)
    u.step(60)
    for a in range(3):
        assert np.array_equal(u.download(a, n), b.download(a, n)), a
        assert np.array_equal(u.download(A.F_U0 + a, n), b.download(A.F_VX + a, n)), a
    assert u.rebuilds() == b.rebuilds()


P_CROWD = [2.0, 1.0, 0.3]            # k, R, c


def crowd_pair(dims, neigh, integ, n, seed):
    L = (n * 1.2) ** (1.0 / dims) * 1.3
    m = model(dims, integrator=integ, neighbors=neigh, law=A.LAW_CUSTOM, dt=0.01, cell_edge=2.3, skin=2.6,
              max_neighbors=48, rebuild_disp=0.1)
    m.user_de = 1
    s = sim(n, [[0, L]] * dims, P_CROWD)
    rng = np.random.default_rng(seed)
    fields = [(a, rng.uniform(0, L, n)) for a in range(dims)]
    fields += [(A.F_U0, rng.uniform(0.0, 1.0, n)), (A.F_U1, rng.uniform(0.7, 1.1, n))]
    o = O.Oracle(m, s)
    g = Context(m, [s])
    g.compile_model(LB.CROWD)
    for f, v in fields:
        o.set(f, v)
        g.upload(f, v)
    o.set_law_body(*LB.crowd(P_CROWD))
    assert o.init() == 0
    g.neighbors()
    return o, g


@pytest.mark.parametrize("dims,neigh,integ", [
    (2, A.FULL, A.EULER), (2, A.CELL_LINKED, A.HEUN), (3, A.CELL_LINKED, A.RK4),
    (3, A.VERLET_DISPLACEMENT, A.HEUN), (2, A.VERLET_DISPLACEMENT, A.RK4)])
def test_crowding_body_equals_the_oracle(gpu_available, dims, neigh, integ):
    n = 150 if neigh == A.FULL else 400
    o, g = crowd_pair(dims, neigh, integ, n, 7 + dims)
    o.step(15)
    g.step(15)
    for f in list(range(dims)) + [A.F_U0, A.F_U1]:
        assert np.array_equal(g.download(f, n), o.get(f)), f
    assert np.std(o.get(A.F_U0)) > 0.01


def test_crowding_body_graph_replays_equal_eager_steps(gpu_available):
    n = 400
    _, g = crowd_pair(2, A.CELL_LINKED, A.HEUN, n, 21)
    _, h = crowd_pair(2, A.CELL_LINKED, A.HEUN, n, 21)
    g.step(12)
    h.step_graph(4, 3)
    for f in (0, 1, A.F_U0):
        assert np.array_equal(g.download(f, n), h.download(f, n)), f


def test_general_body_refuses_lagged_forces(gpu_available):
    m = model(2, neighbors=A.FULL, law=A.LAW_CUSTOM, lagged=1)
    m.user_de = 1
    g = Context(m, [sim(4, [[0, 8]] * 2, P_CROWD)])
    with pytest.raises(RuntimeError, match="lagged"):
        g.compile_model(LB.CROWD)
