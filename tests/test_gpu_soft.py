"""The inertial soft-sphere model (src/models/softSpheres.jl:1-81) on the device
against the oracle: 2-D cell-linked and 3-D Verlet, mixed radii, masses and
fates, Euler/Heun/RK4; plus the Python mirror (CBMModels.softSpheres) end to end."""
import numpy as np
import pytest

import cbm._abi as A
import oracle as O
from cbm.api import (ABM, CBMIntegrators, CBMModels, CBMNeighbors, Community, Context, bringFromPlatform_,
                     evolve_)
from helpers import model, rel_err, sim

pytestmark = pytest.mark.gpu

P = [0.5, 1.4, 2.0, 1.0, 0.5, 3.0]      # b, mu, f0 (2 x 2, row-major)


def agents(n, dims, L, seed):
    rng = np.random.default_rng(seed)
    return {
        "pos": [rng.uniform(0, L, n) for _ in range(dims)],
        "vel": [rng.normal(0, 0.5, n) for _ in range(dims)],
        "r": rng.uniform(0.8, 1.2, n),
        "m": rng.uniform(0.5, 2.0, n),
        "fate": rng.integers(1, 3, n).astype(np.int64),
    }


def both(m, s, ag, dims):
    o = O.Oracle(m, s)
    g = Context(m, [s])
    fields = [(a, ag["pos"][a]) for a in range(dims)] + [(A.F_VX + a, ag["vel"][a]) for a in range(dims)]
    for f, v in fields + [(A.F_RADIUS, ag["r"]), (A.F_MASS, ag["m"]), (A.F_FATE, ag["fate"])]:
        o.set(f, v)
        g.upload(f, v)
    assert o.init() == 0
    g.neighbors()
    return o, g


def assert_same(o, g, dims, tol, what):
    n = o.n
    for f in list(range(dims)) + [A.F_VX + a for a in range(dims)]:
        e = rel_err(g.download(f, n), o.get(f))
        assert e <= tol, (what, f, e)


@pytest.mark.parametrize("integ", [A.EULER, A.HEUN, A.RK4])
def test_soft_spheres_2d_cell_linked_parity(gpu_available, integ):
    n, L = 1500, 40.0
    m = model(2, integrator=integ, neighbors=A.CELL_LINKED, law=A.LAW_SOFT_SPHERES, dt=0.01, cell_edge=3.4)
    s = sim(n, [[0.0, L]] * 2, P)
    o, g = both(m, s, agents(n, 2, L, 3), 2)
    o.step(1)
    g.step(1)
    assert_same(o, g, 2, 1e-12, "step 1")
    o.step(99)
    g.step(99)
    assert_same(o, g, 2, 1e-9, "step 100")


def test_soft_spheres_3d_verlet_parity(gpu_available):
    n = 3000
    L = (n * 4.0 / 3.0 * np.pi / 0.3) ** (1 / 3)
    m = model(3, integrator=A.HEUN, neighbors=A.VERLET_DISPLACEMENT, law=A.LAW_SOFT_SPHERES, dt=0.005,
              skin=3.8, max_neighbors=64, rebuild_disp=0.2)
    s = sim(n, [[0.0, L]] * 3, P)
    o, g = both(m, s, agents(n, 3, L, 5), 3)
    o.step(1)
    g.step(1)
    assert_same(o, g, 3, 1e-12, "step 1")
    o.step(59)
    g.step(59)
    assert_same(o, g, 3, 1e-9, "step 60")
    assert g.rebuilds() == o.rebuilds


def test_soft_spheres_rejects_bad_fate(gpu_available):
    m = model(2, integrator=A.EULER, neighbors=A.FULL, law=A.LAW_SOFT_SPHERES, dt=0.01)
    g = Context(m, [sim(4, [[0.0, 10.0]] * 2, P)])
    with pytest.raises(Exception):
        g.upload(A.F_FATE, np.array([1, 2, 3, 1], dtype=np.int64))


def test_soft_spheres_python_mirror(gpu_available):
    n, L = 400, 25.0
    ag = agents(n, 2, L, 9)
    abm = ABM(2, agentODE=CBMModels.softSpheres(P[0], P[1], [[P[2], P[3]], [P[4], P[5]]]),
              neighborsAlg=CBMNeighbors.CellLinked(cellEdge=3.4), agentAlg=CBMIntegrators.Euler())
    com = Community(abm, N=n, dt=0.01, simBox=[[0.0, L], [0.0, L]], x=ag["pos"][0], y=ag["pos"][1],
                    vx=ag["vel"][0], vy=ag["vel"][1], r=ag["r"], m=ag["m"], cellFate=ag["fate"])
    evolve_(com, steps=20)
    m = model(2, integrator=A.EULER, neighbors=A.CELL_LINKED, law=A.LAW_SOFT_SPHERES, dt=0.01, cell_edge=3.4)
    o = O.Oracle(m, sim(n, [[0.0, L]] * 2, P))
    for a in range(2):
        o.set(a, ag["pos"][a])
        o.set(A.F_VX + a, ag["vel"][a])
    o.set(A.F_RADIUS, ag["r"])
    o.set(A.F_MASS, ag["m"])
    o.set(A.F_FATE, ag["fate"])
    assert o.init() == 0
    o.step(20)
    assert rel_err(com.x, o.get(A.F_X)) <= 1e-12
    assert rel_err(com.vy, o.get(A.F_VY)) <= 1e-12
