"""The flattened list stage (flat_body: a block's pair work spread over all
of its lanes, each agent's terms added in list order) against the per-agent
list walk and the oracle.  Both device forms must give the same bits, for
every round size, with lists that fill several rounds, lists at the
flattened form's capacity and lists beyond it (which take the per-agent
walk), periodic boxes, the per-axis Aggregation law with its walls and
lagged accumulator, EM noise, and stage 2 of the cell-linked Heun step,
which walks the lists stage 1 wrote."""
import numpy as np
import pytest

import cbm._abi as A
import oracle as O
from cbm.api import Context
from helpers import model, sim, uniform_positions

pytestmark = pytest.mark.gpu


def run(m, s, pos, steps, walk, fields=None):
    m.walk = walk
    g = Context(m, [s])
    for c, v in enumerate(pos):
        g.upload(c, v)
    g.neighbors()
    g.step(steps)
    out = [g.download(c, s.n) for c in (fields or range(m.dims))]
    g.close()
    return out


def c2_like(n, phi=0.3, cap=32, integ=A.HEUN, noise=0.0):
    L = (n * (4.0 / 3.0) * np.pi / phi) ** (1.0 / 3.0)
    m = model(3, integrator=integ, neighbors=A.VERLET_DISPLACEMENT, law=A.LAW_REP_ADH, dt=0.005,
              skin=2.7, max_neighbors=cap, rebuild_disp=0.15, noise=noise, seed=2)
    s = sim(n, [[0.0, L]] * 3, [1.0, 1.0, 0.1, 2.4])
    return m, s, uniform_positions(n, [[0.0, L]] * 3, 11)


@pytest.mark.parametrize("walk", [0, 512, 1536, 2048])
def test_flat_rounds_equal_the_per_agent_walk_and_the_oracle(gpu_available, walk):
    m, s, pos = c2_like(40000)
    print(f"flat walk={walk}: C2-like Verlet Heun N=40000, 40 steps", flush=True)
    ref = run(m, s, pos, 40, 1)
    got = run(m, s, pos, 40, walk)
    for a in range(3):
        assert np.array_equal(got[a], ref[a]), a
    o = O.Oracle(m, s)
    for a in range(3):
        o.set(a, pos[a])
    assert o.init() == 0
    o.step(40)
    for a in range(3):
        assert np.array_equal(got[a], o.get(a)), a


@pytest.mark.parametrize("phi,cap", [(1.6, 64), (2.6, 96)])
def test_dense_lists_at_and_beyond_the_flat_capacity(gpu_available, phi, cap):
    """phi 1.6: lists of ~31 entries on average (up to ~57) fill about eight
    rounds per block; phi 2.6 with capacity 96 exceeds the flattened form's
    64, so both settings walk per agent and must still agree."""
    m, s, pos = c2_like(12000, phi=phi, cap=cap)
    ref = run(m, s, pos, 10, 1)
    for walk in (0, 512):
        got = run(m, s, pos, 10, walk)
        for a in range(3):
            assert np.array_equal(got[a], ref[a]), (walk, a)


def test_periodic_aggregation_lagged(gpu_available):
    n, Lb = 2500, 60.0
    m = model(2, integrator=A.EULER, neighbors=A.VERLET_DISPLACEMENT, law=A.LAW_AGGREGATION, dt=0.01,
              periodic=1, lagged=1, skin=2.5, max_neighbors=64, rebuild_disp=0.2)
    s = sim(n, [[0.0, Lb]] * 2, [0.8, 1.0, 1.6, 0.5, 0.0])
    pos = uniform_positions(n, [[0.0, Lb]] * 2, 5)
    ref = run(m, s, pos, 30, 1, fields=[A.F_X, A.F_Y, A.F_FX, A.F_FY])
    got = run(m, s, pos, 30, 0, fields=[A.F_X, A.F_Y, A.F_FX, A.F_FY])
    for a in range(4):
        assert np.array_equal(got[a], ref[a]), a


def test_em_noise_on_the_flat_stage(gpu_available):
    m, s, pos = c2_like(20000, integ=A.EM, noise=np.sqrt(0.02))
    ref = run(m, s, pos, 25, 1)
    got = run(m, s, pos, 25, 0)
    for a in range(3):
        assert np.array_equal(got[a], ref[a]), a


def test_cell_linked_heun_stage_two_lists(gpu_available):
    """Cell-linked Heun: the sorted grid on brick tiles (walk 1), and the skin
    list (stage 2 walks the ordered lists stage 1 wrote) per agent and
    flattened, bit for bit as the oracle's cell walk."""
    n = 30000
    L = (n * (4.0 / 3.0) * np.pi / 0.3) ** (1.0 / 3.0)
    m = model(3, integrator=A.HEUN, neighbors=A.CELL_LINKED, law=A.LAW_REP_ADH, dt=0.005, cell_edge=2.4)
    s = sim(n, [[0.0, L]] * 3, [1.0, 1.0, 0.1, 2.4])
    pos = uniform_positions(n, [[0.0, L]] * 3, 12)
    ref = run(m, s, pos, 20, 1)
    o = O.Oracle(m, s)
    for a in range(3):
        o.set(a, pos[a])
    assert o.init() == 0
    o.step(20)
    for walk in (0, 1024):          # skin list per agent, and flattened
        got = run(m, s, pos, 20, walk)
        for a in range(3):
            assert np.array_equal(got[a], ref[a]), (walk, a)
            assert np.array_equal(got[a], o.get(a)), (walk, a)


def test_walk_option_is_validated(gpu_available):
    m, s, pos = c2_like(100)
    for bad in (300, 4096):
        m.walk = bad
        with pytest.raises(RuntimeError, match="walk"):
            Context(m, [s])
