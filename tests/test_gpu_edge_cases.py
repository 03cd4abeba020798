"""Edge cases of the step: simulations with no agents and with one agent,
for every neighbor algorithm, alone and inside a batched context, where the
other simulations must step exactly as they do on their own."""
import numpy as np
import pytest

import cbm._abi as A
from cbm.api import Context
from helpers import model, sim, uniform_positions

pytestmark = pytest.mark.gpu

ALGS = [
    (A.FULL, {}),
    (A.CELL_LINKED, dict(cell_edge=2.0)),
    (A.VERLET_DISPLACEMENT, dict(skin=2.4, max_neighbors=64, rebuild_disp=0.2)),
    (A.CLVD, dict(skin=2.4, max_neighbors=64, rebuild_disp=0.2, cell_edge=2.4)),
]


@pytest.mark.parametrize("nb,kw", ALGS)
@pytest.mark.parametrize("n", [0, 1])
def test_empty_and_single_agent(gpu_available, nb, kw, n):
    box = [[0.0, 20.0]] * 3
    m = model(3, integrator=A.HEUN, neighbors=nb, law=A.LAW_REP_ADH, dt=0.01, periodic=1, **kw)
    g = Context(m, [sim(n, box, [1.0, 1.0, 0.1, 2.4], n_max=8)])
    pos = uniform_positions(n, box, 4)
    for a, v in enumerate(pos):
        g.upload(a, v)
    g.neighbors()
    g.step(10)
    g.sync()
    assert g.n() == n
    for a in range(3):
        # a lone agent feels no force: it does not move
        assert np.array_equal(g.download(a, n), pos[a])


@pytest.mark.parametrize("nb,kw", ALGS)
def test_batched_with_empty_and_single_simulations(gpu_available, nb, kw):
    box = [[0.0, 24.0]] * 2
    m = model(2, integrator=A.EULER, neighbors=nb, law=A.LAW_LINEAR, dt=0.01, periodic=1, **kw)
    sizes = [0, 400, 1, 0, 250]
    pos = [uniform_positions(n, box, 30 + s) for s, n in enumerate(sizes)]
    sims = [sim(n, box, [1.0, 2.0], n_max=max(n, 4)) for n in sizes]
    g = Context(m, sims)
    for s, p in enumerate(pos):
        for a in range(2):
            g.upload(a, p[a], sim=s)
    g.neighbors()
    g.step(20)
    for s, n in enumerate(sizes):
        assert g.n(s) == n
        alone = Context(m, [sim(n, box, [1.0, 2.0], n_max=max(n, 4))])
        for a in range(2):
            alone.upload(a, pos[s][a])
        alone.neighbors()
        alone.step(20)
        for a in range(2):
            assert np.array_equal(g.download(a, n, sim=s), alone.download(a, n))


def test_non_finite_uploads_are_refused(gpu_available):
    box = [[0.0, 20.0]] * 2
    m = model(2, integrator=A.EULER, neighbors=A.CELL_LINKED, law=A.LAW_LINEAR, dt=0.01, cell_edge=2.0)
    g = Context(m, [sim(5, box, [1.0, 2.0])])
    for bad in (np.nan, np.inf, -np.inf):
        x = np.full(5, 3.0)
        x[2] = bad
        with pytest.raises(A.CBMError) as e:
            g.upload(A.F_X, x)
        assert e.value.code == A.ERR_NONFINITE
    g.upload(A.F_X, np.full(5, 3.0))   # the context stays usable
    assert np.array_equal(g.download(A.F_X, 5), np.full(5, 3.0))
