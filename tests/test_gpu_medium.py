"""Medium (FTCS + uptake) and the growing-colony coupling on the device
against the oracle.  The stencil and the integer node counts are
deterministic, so the field is compared bit-exactly."""
import numpy as np
import pytest

import cbm._abi as A
import oracle as O
from cbm.api import Context
from helpers import model, rel_err, sim
from test_oracle_medium import medium_model

pytestmark = pytest.mark.gpu


def test_medium_step_bit_exact(gpu_available):
    n = (64, 48)
    m = medium_model(2, n, 1.0, 0.1, 0.2)
    m.neighbors, m.law, m.cell_edge = A.CELL_LINKED, A.LAW_LINEAR, 2.0
    N = 500
    box = [[0.0, 64.0], [0.0, 48.0]]
    s = sim(N, box, [1.0, 2.0])
    rng = np.random.default_rng(3)
    pos = [rng.uniform(10, 54, N), rng.uniform(10, 38, N)]
    c0 = rng.uniform(0.5, 1.5, n[0] * n[1])
    o, g = O.Oracle(m, s), Context(m, [s])
    for a in range(2):
        o.set(a, pos[a])
        g.upload(a, pos[a])
    o.set_medium(c0)
    A.check(g.lib.cbm_upload_medium(g.h, c0.ctypes.data, c0.size), g.h)
    o.init()
    g.neighbors()
    for _ in range(5):
        o.step(4)
        g.step(4)
        got = np.empty(c0.size)
        A.check(g.lib.cbm_download_medium(g.h, got.ctypes.data, got.size), g.h)
        assert np.array_equal(got, o.medium(c0.size))
        for a in range(2):
            assert np.array_equal(g.download(a, N), o.get(a))


def test_growing_colony_with_medium(gpu_available):
    n = (128, 128)
    m = medium_model(2, n, 1.0, 0.1, 0.2)
    m.neighbors, m.law, m.cell_edge = A.CELL_LINKED, A.LAW_LINEAR, 2.0
    m.integrator, m.rule, m.seed = A.EULER, A.RULE_GROWTH_DIVISION, 9
    for q, v in enumerate((0.05, 0.1, 0.5, 2.0)):
        m.rule_p[q] = v
    N0 = 300
    s = sim(N0, [[0.0, 128.0], [0.0, 128.0]], [1.0, 2.0], n_max=8000)
    rng = np.random.default_rng(5)
    r = 8 * np.sqrt(rng.uniform(0, 1, N0))
    th = rng.uniform(0, 2 * np.pi, N0)
    fields = {A.F_X: 64 + r * np.cos(th), A.F_Y: 64 + r * np.sin(th), A.F_VOL: rng.uniform(1, 2, N0),
              A.F_VBIRTH: np.ones(N0), A.F_GROWTH: rng.uniform(0.045, 0.055, N0)}
    o, g = O.Oracle(m, s), Context(m, [s])
    for f, v in fields.items():
        o.set(f, v)
        g.upload(f, v)
    o.init()
    g.neighbors()
    for _ in range(6):
        o.step(10)
        g.step(10)
        N = o.n
        assert g.n() == N
        assert np.array_equal(g.download(A.F_ID, N), o.get(A.F_ID))
        for a in range(2):
            assert rel_err(g.download(a, N), o.get(a)) < 1e-9
        got = np.empty(n[0] * n[1])
        A.check(g.lib.cbm_download_medium(g.h, got.ctypes.data, got.size), g.h)
        assert rel_err(got, o.medium(got.size)) < 1e-12
    assert o.n > 1.5 * N0
