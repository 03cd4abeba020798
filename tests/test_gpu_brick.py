"""The LDS-tiled cell-linked stage against the oracle on the cases that stress
its tiling: a brick whose halo overflows the LDS capacity (it walks global
memory instead), periodic grids of 3-5 cells per axis (a brick's halo wraps
onto itself), 1-D grids, and block sizes smaller than a brick's population.
The walk order is the reference's cellPosNeigh order (src/neighbors.jl), so
the built-in laws are expected bit-identical to the oracle.

The same cases also run with walk = 0, where an Euler or Heun model takes
CellLinked on a skin list (no per-step cell sort; each agent's skin-list
entries are put in the cell walk's order): there the dense clusters give
lists past the skin list's capacity and large stage-1 steps move agents past
the list margin, and those agents take the exact cell-order walk."""
import numpy as np
import pytest

import cbm._abi as A
import oracle as O
from cbm.api import Context
from helpers import model, rel_err, sim, uniform_positions

pytestmark = pytest.mark.gpu

LONG_TOL = 1e-9
REP_ADH = [1.0, 1.0, 0.1, 2.4]


WALKS = pytest.mark.parametrize("walk", [1, 0], ids=["bricks", "skinlist"])


def pair(m, s, pos, block=None, walk=1):
    m.walk = walk
    o = O.Oracle(m, s)
    g = Context(m, [s])
    if block:
        g.set_block(block)
    for a, v in enumerate(pos):
        o.set(a, v)
        g.upload(a, v)
    assert o.init() == 0
    g.neighbors()
    return o, g


def err(o, g, dims):
    n = o.n
    return max(rel_err(g.download(a, n), o.get(a)) for a in range(dims))


def test_brick_geometry_is_morton_aligned(gpu_available):
    for dims, L in [(1, 2000.0), (2, 60.0), (3, 30.0)]:
        n = 2000
        box = [[0.0, L]] * dims
        m = model(dims, integrator=A.EULER, neighbors=A.CELL_LINKED, law=A.LAW_REP_ADH, dt=0.01, cell_edge=2.4)
        g = Context(m, [sim(n, box, REP_ADH)])
        b = g.bricks()
        bits = [int(np.log2(v)) for v in b["bdim"][:dims]]
        assert all(2 ** q == v for q, v in zip(bits, b["bdim"][:dims]))
        # x takes the extra bits first: x >= y >= z >= x - 1 (in bits)
        assert bits == sorted(bits, reverse=True) and bits[0] - bits[-1] <= 1
        assert b["cap"] >= 256 and b["nbricks"] >= 1 and b["overflows"] == 0


@WALKS
def test_brick_overflow_falls_back_bit_exact(gpu_available, walk):
    # 1,500 of 2,000 agents packed in a 4.8-wide cube: the bricks around it
    # hold far more agents than the capacity sized for the mean density
    rng = np.random.default_rng(11)
    n, L = 2000, 40.0
    box = [[0.0, L]] * 3
    pos = [rng.uniform(0.0, L, n) for _ in range(3)]
    for a in range(3):
        pos[a][:1500] = rng.uniform(17.6, 22.4, 1500)
    m = model(3, integrator=A.EULER, neighbors=A.CELL_LINKED, law=A.LAW_REP_ADH, dt=1e-4, cell_edge=2.4)
    o, g = pair(m, sim(n, box, REP_ADH), pos, walk=walk)
    assert g.bricks()["cap"] < 1500
    o.step(1)
    g.step(1)
    assert err(o, g, 3) == 0.0
    assert g.bricks()["overflows"] > 0
    o.step(19)
    g.step(19)
    assert err(o, g, 3) <= LONG_TOL


@WALKS
@pytest.mark.parametrize("dims", [2, 3])
def test_brick_dense_cluster_overflows_lane_queue(gpu_available, dims, walk):
    # 60 agents within 1.0 of a point: each has 59 candidates within the law's
    # reach, more than a lane queues, while the brick still fits in LDS
    rng = np.random.default_rng(7 + dims)
    n, L = 460, 40.0
    box = [[0.0, L]] * dims
    pos = [rng.uniform(0.0, L, n) for _ in range(dims)]
    u = rng.normal(size=(dims, 60))
    u *= rng.uniform(0.0, 1.0, 60) ** (1.0 / dims) / np.linalg.norm(u, axis=0)
    for a in range(dims):
        pos[a][:60] = 20.0 + u[a]
    m = model(dims, integrator=A.HEUN, neighbors=A.CELL_LINKED, law=A.LAW_REP_ADH, dt=1e-3, periodic=1,
              cell_edge=2.4)
    o, g = pair(m, sim(n, box, REP_ADH), pos, walk=walk)
    o.step(1)
    g.step(1)
    assert err(o, g, dims) == 0.0
    o.step(19)
    g.step(19)
    assert err(o, g, dims) <= LONG_TOL
    if walk:
        assert g.bricks()["overflows"] == 0


@WALKS
@pytest.mark.parametrize("gn", [3, 4, 5])
def test_brick_small_periodic_grid(gpu_available, gn, walk):
    # a periodic grid narrower than a brick: the halo wraps onto the brick
    L = gn * 2.4 + 0.01
    n = int(0.9 * gn ** 3) + 5
    box = [[0.0, L]] * 3
    m = model(3, integrator=A.HEUN, neighbors=A.CELL_LINKED, law=A.LAW_REP_ADH, dt=0.005, periodic=1,
              cell_edge=2.4)
    o, g = pair(m, sim(n, box, REP_ADH), uniform_positions(n, box, 20 + gn), walk=walk)
    if walk:
        assert g.bricks()["bdim"][0] > gn
    o.step(1)
    g.step(1)
    assert err(o, g, 3) == 0.0
    o.step(29)
    g.step(29)
    assert err(o, g, 3) <= LONG_TOL


@WALKS
@pytest.mark.parametrize("periodic", [0, 1])
def test_brick_1d(gpu_available, periodic, walk):
    n, L = 400, 300.0
    box = [[0.0, L]]
    m = model(1, integrator=A.EULER, neighbors=A.CELL_LINKED, law=A.LAW_LINEAR, dt=0.01, periodic=periodic,
              cell_edge=2.0)
    o, g = pair(m, sim(n, box, [1.0, 2.0]), uniform_positions(n, box, 3), walk=walk)
    o.step(1)
    g.step(1)
    assert err(o, g, 1) == 0.0
    o.step(49)
    g.step(49)
    assert err(o, g, 1) <= LONG_TOL


def test_brick_block_sizes_agree(gpu_available):
    # a brick of ~256 agents served by blocks of 64, 128 and 256 threads
    n = 6000
    box = [[0.0, 40.0]] * 3
    pos = uniform_positions(n, box, 5)
    m = model(3, integrator=A.RK4, neighbors=A.CELL_LINKED, law=A.LAW_REP_ADH, dt=0.01, periodic=1, cell_edge=2.4)
    out = []
    for block in (64, 128, 256):
        o, g = pair(m, sim(n, box, REP_ADH), pos, block=block)
        g.step(5)
        out.append([g.download(a, n) for a in range(3)])
    o.step(5)
    ref = [o.get(a) for a in range(3)]
    for got in out:
        for a in range(3):
            assert np.array_equal(got[a], out[0][a])
            assert rel_err(got[a], ref[a]) == 0.0


@pytest.mark.parametrize("cluster", [60, 1500])
def test_verlet_brick_build_matches_oracle(gpu_available, cluster):
    # Verlet lists built on brick tiles: 60 agents within 1.0 of a point give
    # lists longer than a lane's sorted queue; 1,500 in a 4.8-wide cube give a
    # halo beyond the tile capacity, left to the row-parallel build
    rng = np.random.default_rng(100 + cluster)
    n, L = 2000, 40.0
    box = [[0.0, L]] * 3
    pos = [rng.uniform(0.0, L, n) for _ in range(3)]
    if cluster == 60:
        u = rng.normal(size=(3, 60))
        u *= rng.uniform(0.0, 1.0, 60) ** (1.0 / 3.0) / np.linalg.norm(u, axis=0)
        for a in range(3):
            pos[a][:60] = 20.0 + u[a]
    else:
        for a in range(3):
            pos[a][:cluster] = rng.uniform(17.6, 22.4, cluster)
    m = model(3, integrator=A.EULER, neighbors=A.VERLET_DISPLACEMENT, law=A.LAW_REP_ADH, dt=1e-4, periodic=1,
              skin=2.7, max_neighbors=2048, rebuild_disp=0.05)
    o, g = pair(m, sim(n, box, REP_ADH), pos)
    b = g.bricks()
    assert b["nbricks"] > 0
    if cluster > 60:
        assert b["cap"] < cluster and b["overflows"] > 0
    else:
        assert b["overflows"] == 0
    o.step(1)
    g.step(1)
    assert err(o, g, 3) == 0.0
    o.step(29)
    g.step(29)
    assert err(o, g, 3) <= LONG_TOL
    assert g.rebuilds() == o.rebuilds


@WALKS
@pytest.mark.parametrize("integ,dt", [(A.HEUN, 0.005), (A.HEUN, 0.4), (A.EULER_HEUN, 0.005)])
def test_cell_linked_two_stage_lists(gpu_available, integ, dt, walk):
    # stage 1 writes each agent's candidates within 1.05x the law's reach and
    # stage 2 walks them as a list while no stage-1 displacement exceeds half
    # the margin; dt = 0.4 moves overlapping agents farther than that, so
    # stage 2 falls back to the brick tiles
    n, L = 3000, 30.0
    box = [[0.0, L]] * 3
    m = model(3, integrator=integ, neighbors=A.CELL_LINKED, law=A.LAW_REP_ADH, dt=dt, periodic=1, cell_edge=2.4,
              noise=0.05 if integ == A.EULER_HEUN else 0.0, seed=11)
    o, g = pair(m, sim(n, box, REP_ADH), uniform_positions(n, box, 31), walk=walk)
    o.step(1)
    g.step(1)
    # the noise goes through the device's log/sin/cos: 1e-12 per step, not bits
    assert err(o, g, 3) <= (1e-12 if integ == A.EULER_HEUN else 0.0)
    o.step(19)
    g.step(19)
    assert err(o, g, 3) <= LONG_TOL


@pytest.mark.parametrize("law", [A.LAW_LINEAR, A.LAW_REP_ADH], ids=["linear", "rep_adh"])
def test_dense_core_runs_on_sub_bricks_bit_exact(gpu_available, law):
    # With rules the main bricks are sized for the capacity's mean density; a
    # core of ~30 agents per cell (edge 2) overflows them and the sub-bricks
    # (sized for 48 per cell) stage it; a core of ~64 per cell (edge 2.4)
    # overflows some sub-bricks too, whose rows the row kernel takes.  A death
    # rule at rate 0 keeps the population.
    rng = np.random.default_rng(5)
    n, L, core = 6000, 200.0, 4000
    box = [[0.0, L]] * 2
    pos = [rng.uniform(0.0, L, n) for _ in range(2)]
    p, edge, half = ([1.0, 2.0], 2.0, 11.5) if law == A.LAW_LINEAR else (REP_ADH, 2.4, 9.5)
    for a in range(2):
        pos[a][:core] = rng.uniform(100.0 - half, 100.0 + half, core)
    m = model(2, integrator=A.EULER, neighbors=A.CELL_LINKED, law=law, dt=1e-3, cell_edge=edge,
              rule=A.RULE_DEATH, rule_p=(0.0, 0, 0, 0), seed=3)
    m.walk = 1
    s = sim(n, box, p, n_max=8000)
    o = O.Oracle(m, s)
    g = Context(m, [s])
    for a in range(2):
        o.set(a, pos[a])
        g.upload(a, pos[a])
    assert o.init() == 0
    g.neighbors()
    sb, b = g.sub_bricks(), g.bricks()
    assert sb["nbricks"] > b["nbricks"] and sb["cap"] > b["cap"]
    assert np.prod(sb["bdim"]) < np.prod(b["bdim"])
    o.step(1)
    g.step(1)
    assert err(o, g, 2) == 0.0
    assert g.bricks()["overflows"] > 0
    o.step(9)
    g.step(9)
    assert g.n() == o.n == n
    assert err(o, g, 2) == 0.0
