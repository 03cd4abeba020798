"""User mediumODE bodies compiled with hipRTC (cbm_compile_model
medium_body): a body restating the built-in diffusion with uptake equals the
built-in kernel bit for bit, and a secretion/decay body with a spatial
source matches the oracle running its Python restatement, bit for bit
(the step is plain FP64 arithmetic)."""
import numpy as np
import pytest

import cbm._abi as A
import medium_bodies as MB
import oracle as O
from cbm.api import ABM, CBMIntegrators, CBMModels, CBMNeighbors, Community, Context, evolve_
from helpers import sim, uniform_positions
from test_oracle_medium import medium_model

pytestmark = pytest.mark.gpu


def field(g, size):
    got = np.empty(size)
    A.check(g.lib.cbm_download_medium(g.h, got.ctypes.data, got.size), g.h)
    return got


def setup(m, s, pos, c0, body=None):
    g = Context(m, [s])
    if body is not None:
        g.compile_model(medium=body)
    for a, v in enumerate(pos):
        g.upload(a, v)
    A.check(g.lib.cbm_upload_medium(g.h, c0.ctypes.data, c0.size), g.h)
    g.neighbors()
    return g


def test_builtin_restatement_bit_exact(gpu_available):
    n = (64, 48)
    m = medium_model(2, n, 1.0, 0.1, 0.2)
    m.neighbors, m.law, m.cell_edge = A.CELL_LINKED, A.LAW_LINEAR, 2.0
    N = 500
    box = [[0.0, 64.0], [0.0, 48.0]]
    s = sim(N, box, [1.0, 2.0])
    rng = np.random.default_rng(3)
    pos = [rng.uniform(10, 54, N), rng.uniform(10, 38, N)]
    c0 = rng.uniform(0.5, 1.5, n[0] * n[1])
    gb = setup(m, s, pos, c0)
    gu = setup(m, s, pos, c0, MB.BUILTIN)
    gb.step(20)
    gu.step(20)
    assert np.array_equal(field(gb, c0.size), field(gu, c0.size))


@pytest.mark.parametrize("dims", [2, 3])
def test_secretion_body_matches_oracle(gpu_available, dims):
    n = (40, 30, 12)[:dims]
    box = [[0.0, 40.0], [0.0, 30.0], [0.0, 12.0]][:dims]
    m = medium_model(dims, n, 0.7, 0.0, 0.1)
    m.neighbors, m.law, m.cell_edge, m.periodic = A.CELL_LINKED, A.LAW_LINEAR, 2.0, 1
    N = 800
    p = [1.0, 2.0, 0, 0, 0, 0, 0.05, 0.3]
    s = sim(N, box, p)
    pos = uniform_positions(N, box, 9 + dims)
    c0 = np.random.default_rng(dims).uniform(0.0, 1.0, int(np.prod(n)))
    o = O.Oracle(m, s)
    o.set_medium_body(MB.secrete(0.7, 0.05, 0.3))
    for a, v in enumerate(pos):
        o.set(a, v)
    o.set_medium(c0)
    assert o.init() == 0
    g = setup(m, s, pos, c0, MB.SECRETE)
    for _ in range(3):
        o.step(5)
        g.step(5)
        assert np.array_equal(field(g, c0.size), o.medium(c0.size))
        for a in range(dims):
            assert np.array_equal(g.download(a, N), o.get(a))


def test_medium_body_errors_and_python_mirror(gpu_available):
    m = medium_model(2, (8, 8), 1.0, 0.0, 0.1)
    m.medium_n[0] = 0                       # no medium
    with pytest.raises(A.CBMError):
        Context(m, [sim(4, [[0.0, 8.0], [0.0, 8.0]])]).compile_model(medium=MB.BUILTIN)
    # the Python mirror compiles medium["mediumODE"] at loadToPlatform_
    N = 100
    pos = uniform_positions(N, [[0.0, 32.0], [0.0, 32.0]], 1)
    abm = ABM(2, agentODE=CBMModels.linearRepulsion(1.0, 2.0), neighborsAlg=CBMNeighbors.CellLinked(cellEdge=2.0),
              agentAlg=CBMIntegrators.Euler(), periodic=True,
              medium={"NMedium": [32, 32], "D": 0.5, "consume": 0.0, "c0": 0.0, "mediumODE": MB.GEOMETRY})
    com = Community(abm, N=N, dt=0.1, simBox=[[0.0, 32.0], [0.0, 32.0]], x=pos[0], y=pos[1])
    evolve_(com, steps=2)
    got = field(com.ctx, 32 * 32)
    i = np.tile(np.arange(32), 32)
    j = np.repeat(np.arange(32), 32)
    rate = (i + 0.5) + 1000.0 * j
    assert np.allclose(got, 2 * 0.1 * rate, rtol=1e-14, atol=0)
