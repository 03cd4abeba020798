"""DGADI medium solver on the device against the oracle's restatement, bit
for bit (the sweeps are serial Thomas solves in the same order of
operations): the built-in uptake law with moving agents in 2-D and 3-D, a
user mediumODE body, and the configuration check."""
import numpy as np
import pytest

import cbm._abi as A
import medium_bodies as MB
import oracle as O
from cbm.api import Context
from helpers import sim, uniform_positions
from test_oracle_medium_adi import adi_model

pytestmark = pytest.mark.gpu


def field(g, size):
    got = np.empty(size)
    A.check(g.lib.cbm_download_medium(g.h, got.ctypes.data, got.size), g.h)
    return got


@pytest.mark.parametrize("dims,user", [(1, False), (2, False), (3, False), (2, True)])
def test_adi_matches_oracle(gpu_available, dims, user):
    n = (64, 48, 12)[:dims]
    box = [[0.0, 64.0], [0.0, 48.0], [0.0, 12.0]][:dims]
    m = adi_model(dims, n, 0.9, 0.1, 0.25)
    m.neighbors, m.law, m.cell_edge, m.periodic = A.CELL_LINKED, A.LAW_LINEAR, 2.0, 1
    N = 600 if dims > 1 else 40
    p = [1.0, 2.0, 0, 0, 0, 0, 0.05, 0.3]
    s = sim(N, box, p)
    pos = uniform_positions(N, box, 20 + dims)
    c0 = np.random.default_rng(dims).uniform(0.5, 1.5, int(np.prod(n)))
    o = O.Oracle(m, s)
    g = Context(m, [s])
    if user:
        o.set_medium_body(MB.secrete(0.9, 0.05, 0.3))
        g.compile_model(medium=MB.SECRETE)
    for a, v in enumerate(pos):
        o.set(a, v)
        g.upload(a, v)
    o.set_medium(c0)
    A.check(g.lib.cbm_upload_medium(g.h, c0.ctypes.data, c0.size), g.h)
    assert o.init() == 0
    g.neighbors()
    for _ in range(3):
        o.step(4)
        g.step(4)
        assert np.array_equal(field(g, c0.size), o.medium(c0.size))
        for a in range(dims):
            assert np.array_equal(g.download(a, N), o.get(a))


def test_adi_needs_four_nodes_per_axis(gpu_available):
    m = adi_model(2, (3, 8), 1.0, 0.0, 0.1)
    with pytest.raises(A.CBMError) as e:
        Context(m, [sim(4, [[0.0, 3.0], [0.0, 8.0]])])
    assert e.value.code == A.ERR_CONFIG
