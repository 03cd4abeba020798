"""CellLinked on a skin list: a cell-linked model with a built-in law under
Euler or Heun runs without a per-step cell sort.  Its skin list (radius
reach + reach/8, rebuilt when an agent has moved reach/16) holds every pair
within the law's reach, and each stage orders an agent's entries by the
position of the entry's cell among the agent's 3^D neighbor cells, then by
slot: the reference CellLinked loop order (src/neighbors.jl:715-747).  The
results must equal the oracle's cell walk bit for bit, and the per-step
rebuild path (walk = 1) too, for every export."""
import numpy as np
import pytest

import cbm._abi as A
import oracle as O
from cbm.api import Context
from helpers import model, sim, uniform_positions

pytestmark = pytest.mark.gpu
REP_ADH = [1.0, 1.0, 0.1, 2.4]


def device(m, sims, pos, walk):
    m.walk = walk
    g = Context(m, sims)
    for q, p in enumerate(pos):
        for a, v in enumerate(p):
            g.upload(a, v, sim=q)
    g.neighbors()
    return g


def oracle(m, s, pos):
    o = O.Oracle(m, s)
    for a, v in enumerate(pos):
        o.set(a, v)
    assert o.init() == 0
    return o


def c2_cl(n, seed=12):
    L = (n * (4.0 / 3.0) * np.pi / 0.3) ** (1.0 / 3.0)
    m = model(3, integrator=A.HEUN, neighbors=A.CELL_LINKED, law=A.LAW_REP_ADH, dt=0.005, cell_edge=2.4)
    s = sim(n, [[0.0, L]] * 3, REP_ADH)
    return m, s, uniform_positions(n, [[0.0, L]] * 3, seed)


def test_heun_3d_equals_the_oracle_and_the_rebuilt_grid(gpu_available):
    m, s, pos = c2_cl(30000)
    print("skin-list CellLinked: 3-D Heun N=30000, 100 steps", flush=True)
    g = device(m, [s], [pos], 0)
    b = device(m, [s], [pos], 1)
    o = oracle(m, s, pos)
    for k in (1, 9, 90):
        g.step(k)
        b.step(k)
        o.step(k)
        for a in range(3):
            x = g.download(a, s.n)
            assert np.array_equal(x, o.get(a)), (k, a)
            assert np.array_equal(x, b.download(a, s.n)), (k, a)
    # the skin list was rebuilt a few times; the grid of walk = 1 every step
    assert 1 <= g.rebuilds() <= 30 < b.rebuilds()


@pytest.mark.parametrize("law,p,lagged", [(A.LAW_LINEAR, [1.0, 2.0], 0),
                                          (A.LAW_AGGREGATION, [0.8, 1.0, 1.6, 0.5, 0.0], 1),
                                          (A.LAW_AGGREGATION, [0.8, 1.0, 1.6, 0.5, 1.0], 1)],
                         ids=["linear", "aggregation", "aggregation-walls"])
def test_euler_2d_periodic_and_walled(gpu_available, law, p, lagged):
    n, L = 1500, 50.0
    periodic = 1 if law == A.LAW_LINEAR else 0
    m = model(2, integrator=A.EULER, neighbors=A.CELL_LINKED, law=law, dt=0.01, periodic=periodic,
              lagged=lagged, cell_edge=2.0)
    s = sim(n, [[0.0, L]] * 2, p)
    pos = uniform_positions(n, [[0.0, L]] * 2, 21)
    g = device(m, [s], [pos], 0)
    o = oracle(m, s, pos)
    g.step(200)
    o.step(200)
    for a in range(2):
        assert np.array_equal(g.download(a, n), o.get(a)), a
    if lagged:
        for f in (A.F_FX, A.F_FY):
            assert np.array_equal(g.download(f, n), o.get(f)), f
    assert g.rebuilds() < 200


@pytest.mark.parametrize("radius", [1.5, 2.4, 4.0])
def test_exports_follow_the_cell_walk(gpu_available, radius):
    """Neighbor counts and ordered lists: radii below, at and beyond the
    law's reach (the last needs the cells two away in the list grid)."""
    n, L = 3000, 30.0
    m = model(3, integrator=A.EULER, neighbors=A.CELL_LINKED, law=A.LAW_REP_ADH, dt=0.01, cell_edge=2.4)
    s = sim(n, [[0.0, L]] * 3, REP_ADH)
    pos = uniform_positions(n, [[0.0, L]] * 3, 4)
    g = device(m, [s], [pos], 0)
    b = device(m, [s], [pos], 1)
    o = oracle(m, s, pos)
    g.step(7)
    b.step(7)
    o.step(7)
    cg = g.count_neighbors(radius)
    assert np.array_equal(cg, b.count_neighbors(radius))
    assert np.array_equal(cg, o.count_neighbors(radius))
    cap = int(cg.max()) + 1
    lg, lb = g.neighbor_lists(radius, cap), b.neighbor_lists(radius, cap)
    assert np.array_equal(lg[0], lb[0]) and np.array_equal(lg[1], lb[1])


def test_batched_simulations_each_equal_their_oracle(gpu_available):
    rng = np.random.default_rng(9)
    m = model(2, integrator=A.EULER, neighbors=A.CELL_LINKED, law=A.LAW_REP_ADH, dt=0.01, periodic=1,
              cell_edge=2.4)
    sims, pos = [], []
    for q in range(5):
        n = int(rng.integers(300, 3000))
        L = float(np.sqrt(n * np.pi / rng.uniform(0.2, 0.5)))
        sims.append(sim(n, [[0.0, L]] * 2, [rng.uniform(0.5, 2.0), 1.0, rng.uniform(0.0, 0.2), 2.4]))
        pos.append(uniform_positions(n, [[0.0, L]] * 2, 40 + q))
    g = device(m, sims, pos, 0)
    g.step(60)
    for q, s in enumerate(sims):
        o = oracle(m, s, pos[q])
        o.step(60)
        for a in range(2):
            assert np.array_equal(g.download(a, s.n, sim=q), o.get(a)), (q, a)


def test_law_parameters_may_shrink_but_not_outgrow_the_lists(gpu_available):
    m, s, pos = c2_cl(4000)
    g = device(m, [s], [pos], 0)
    g.step(3)
    with pytest.raises(RuntimeError, match="reach"):
        g.set_params([1.0, 1.2, 0.1, 2.4])      # r_eq 1.2: reach 2.88 > 2.4
    g.set_params([1.0, 0.9, 0.1, 2.4])          # reach 2.16
    s2 = sim(s.n, [[s.box_lo[a], s.box_hi[a]] for a in range(3)], [1.0, 0.9, 0.1, 2.4])
    o = oracle(m, s2, [g.download(a, s.n) for a in range(3)])
    g.step(20)
    o.step(20)
    for a in range(3):
        assert np.array_equal(g.download(a, s.n), o.get(a)), a


def test_eager_steps_and_graph_replays_agree(gpu_available):
    """Eager steps schedule the list builds from the host (the record the
    order kernel writes), graph replays from the device flag: the two may
    build at different steps, and the positions must still be equal."""
    m, s, pos = c2_cl(6000, seed=31)
    g = device(m, [s], [pos], 0)
    h = device(m, [s], [pos], 0)
    g.step(120)
    h.step_graph(4, 30)
    for a in range(3):
        assert np.array_equal(g.download(a, s.n), h.download(a, s.n)), a
    assert g.rebuilds() >= 2 and h.rebuilds() >= 2


def test_late_builds_stay_exact(gpu_available):
    """Agents that cover a fifth of the skin per step outrun any build the
    host can schedule a few steps ahead: stage 1 bounds every displacement
    by the reported maximum and the agents whose list may miss a pair take
    the exact walk, so the result is the oracle's cell walk bit for bit."""
    n, L = 1500, 45.0
    m = model(2, integrator=A.EULER, neighbors=A.CELL_LINKED, law=A.LAW_LINEAR, dt=0.02, periodic=1,
              cell_edge=2.0)
    s = sim(n, [[0.0, L]] * 2, [3.0, 2.0])
    pos = uniform_positions(n, [[0.0, L]] * 2, 33)
    g = device(m, [s], [pos], 0)
    o = oracle(m, s, pos)
    for k in (3, 17, 40):
        g.step(k)
        o.step(k)
        for a in range(2):
            assert np.array_equal(g.download(a, n), o.get(a)), (k, a)
    assert g.rebuilds() >= 4


def test_heun_steps_ten_times_longer(gpu_available):
    """With a step ten times C2's, agents move past half the skin within a
    step or two: stage 2 sends the agents whose list may miss a pair to the
    exact walk, builds come every few steps, and the result is still the
    oracle's cell walk, bit for bit."""
    n = 4000
    L = (n * (4.0 / 3.0) * np.pi / 0.3) ** (1.0 / 3.0)
    m = model(3, integrator=A.HEUN, neighbors=A.CELL_LINKED, law=A.LAW_REP_ADH, dt=0.05, cell_edge=2.4)
    s = sim(n, [[0.0, L]] * 3, REP_ADH)
    pos = uniform_positions(n, [[0.0, L]] * 3, 41)
    g = device(m, [s], [pos], 0)
    o = oracle(m, s, pos)
    for k in (1, 4, 10):
        g.step(k)
        o.step(k)
        for a in range(3):
            assert np.array_equal(g.download(a, n), o.get(a)), (k, a)


def test_fast_agents_switch_to_builds_every_step(gpu_available):
    # four 2-D periodic simulations of random discs at packing 0.5 under
    # REP_ADH: overlapping agents cross half the skin within a step or two of
    # every build, so builds come 4 steps apart (the host's lead) and the
    # context switches to a build after every step; each simulation stays
    # its oracle's bit for bit
    m = model(2, integrator=A.EULER, neighbors=A.CELL_LINKED, law=A.LAW_REP_ADH, dt=0.01, cell_edge=2.4,
              periodic=1)
    rng = np.random.default_rng(5)
    sims, pos = [], []
    for q in range(4):
        n = 4000
        L = float(np.sqrt(n * np.pi / 0.5))
        sims.append(sim(n, [[0.0, L]] * 2, [2.0, 1.0, 0.1, 2.4]))
        pos.append([rng.uniform(0.0, L, n), rng.uniform(0.0, L, n)])
    print("skin-list CellLinked: 4 x 2-D periodic REP_ADH N=4000, packing 0.5, 40 steps", flush=True)
    g = device(m, sims, pos, 0)
    g.step(40)
    for q in range(4):
        o = oracle(m, sims[q], pos[q])
        o.step(40)
        for a in range(2):
            assert np.array_equal(g.download(a, 4000, sim=q), o.get(a)), (q, a)
    assert g.rebuilds() >= 20
