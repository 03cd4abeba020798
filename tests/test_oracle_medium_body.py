"""The oracle's user-mediumODE path (a body restated in Python, called per
node) pinned by the built-in medium it must reproduce, by a closed form
(uniform decay), and by the node geometry of cell-centred nodes
(dx = L/n, communityStructure.jl:274).  CPU only."""
import numpy as np

import cbm._abi as A
import medium_bodies as MB
import oracle as O
from helpers import sim, uniform_positions
from test_oracle_medium import medium_model


def test_builtin_restatement_equals_builtin():
    n = (24, 16)
    box = [[0.0, 24.0], [0.0, 16.0]]
    N = 200
    pos = uniform_positions(N, box, 4)
    c0 = np.random.default_rng(2).uniform(0.5, 1.5, n[0] * n[1])
    out = []
    for user in (False, True):
        m = medium_model(2, n, 0.8, 0.1, 0.2)
        o = O.Oracle(m, sim(N, box))
        if user:
            o.set_medium_body(MB.builtin(0.8, 0.1))
        o.set(A.F_X, pos[0])
        o.set(A.F_Y, pos[1])
        o.set_medium(c0)
        assert o.init() == 0
        o.step(6)
        out.append(o.medium(c0.size))
    assert np.array_equal(out[0], out[1])


def test_uniform_decay_closed_form():
    # dt(c) = -k c on a uniform field: the mirrored Laplacian vanishes and
    # every node follows c0 (1 - k dt)^t
    n = (10, 6, 4)
    m = medium_model(3, n, 1.0, 0.0, 0.1)
    o = O.Oracle(m, sim(0, [[0.0, 10.0], [0.0, 6.0], [0.0, 4.0]]))
    o.set_medium_body(lambda nd: nd.laplacian() - 0.5 * nd.c)
    o.set_medium(np.full(np.prod(n), 2.0))
    assert o.init() == 0
    o.step(7)
    assert np.max(np.abs(o.medium(int(np.prod(n))) / (2.0 * (1 - 0.5 * 0.1) ** 7) - 1)) < 1e-14


def test_node_geometry():
    n = (8, 5)
    lo, hi = (-4.0, 1.0), (12.0, 11.0)
    m = medium_model(2, n, 0.0, 0.0, 0.5)
    o = O.Oracle(m, sim(0, [[lo[0], hi[0]], [lo[1], hi[1]]]))
    o.set_medium_body(MB.geometry)
    o.set_medium(np.zeros(n[0] * n[1]))
    assert o.init() == 0
    o.step(1)
    dx = (hi[0] - lo[0]) / n[0]
    i = np.tile(np.arange(n[0]), n[1])
    j = np.repeat(np.arange(n[1]), n[0])
    want = 0.5 * ((lo[0] + (i + 0.5) * dx) + 1000.0 * j)
    assert np.allclose(o.medium(n[0] * n[1]), want, rtol=1e-15, atol=0)
