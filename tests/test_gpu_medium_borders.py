"""Medium borders on the device against the oracle, bit for bit: Dirichlet
and copy-the-interior sides of the built-in diffusion with uptake (the
`@mediumBorder` assignments of a mediumODE body, with `@mediumInside()`
nodes integrating), in 2-D and 3-D with agents, and a user body that makes
the same assignments itself (struct UserMedium::assign).  Borders that the
solver cannot honour are refused."""
import numpy as np
import pytest

import cbm._abi as A
import medium_bodies as MB
import oracle as O
from cbm.api import Context
from helpers import sim, uniform_positions
from test_oracle_medium import bordered, medium_model

pytestmark = pytest.mark.gpu


def field(g, size):
    got = np.empty(size)
    A.check(g.lib.cbm_download_medium(g.h, got.ctypes.data, got.size), g.h)
    return got


def pair(m, s, pos, c0, body=None, oracle_body=None):
    o = O.Oracle(m, s)
    if oracle_body is not None:
        o.set_medium_body(*oracle_body)
    for a, v in enumerate(pos):
        o.set(a, v)
    o.set_medium(c0)
    assert o.init() == 0
    g = Context(m, [s])
    if body is not None:
        g.compile_model(medium=body)
    for a, v in enumerate(pos):
        g.upload(a, v)
    A.check(g.lib.cbm_upload_medium(g.h, c0.ctypes.data, c0.size), g.h)
    g.neighbors()
    return o, g


@pytest.mark.parametrize("dims,borders", [
    (2, ["copy", 2.0, "copy", "mirror"]),
    (2, [0.0, 0.0, 1.5, "copy"]),
    (3, ["copy", "copy", 0.25, "mirror", "copy", 3.0]),
], ids=["2d-copy-dirichlet", "2d-all-fixed", "3d-mixed"])
def test_builtin_borders_bit_exact(gpu_available, dims, borders):
    n = (40, 30, 12)[:dims]
    box = [[0.0, 40.0], [0.0, 30.0], [0.0, 12.0]][:dims]
    m = bordered(medium_model(dims, n, 0.6, 0.05, 0.2), borders)
    m.neighbors, m.law, m.cell_edge = A.CELL_LINKED, A.LAW_LINEAR, 2.0
    N = 600
    s = sim(N, box, [1.0, 2.0])
    pos = uniform_positions(N, box, 17 + dims)
    c0 = np.random.default_rng(dims).uniform(0.0, 1.0, int(np.prod(n)))
    o, g = pair(m, s, pos, c0)
    for k in (1, 9, 30):
        o.step(k)
        g.step(k)
        assert np.array_equal(field(g, c0.size), o.medium(c0.size)), k


def test_user_body_assignments_bit_exact(gpu_available):
    n = (48, 20)
    box = [[0.0, 48.0], [0.0, 20.0]]
    m = medium_model(2, n, 0.8, 0.0, 0.1)
    m.neighbors, m.law, m.cell_edge = A.CELL_LINKED, A.LAW_LINEAR, 2.0
    N = 300
    s = sim(N, box, [1.0, 2.0])
    pos = uniform_positions(N, box, 23)
    c0 = np.random.default_rng(4).uniform(0.0, 2.0, n[0] * n[1])
    o, g = pair(m, s, pos, c0, MB.IDIOM, MB.idiom(0.8))
    for k in (1, 10, 40):
        o.step(k)
        g.step(k)
        got = field(g, c0.size)
        assert np.array_equal(got, o.medium(c0.size)), k
    assert np.all(got.reshape(n[1], n[0])[:, -1] == 0.0)


def test_borders_the_solver_cannot_honour_are_refused(gpu_available):
    box = [[0.0, 16.0], [0.0, 16.0]]
    m = bordered(medium_model(2, (16, 16), 1.0, 0.0, 0.1), [1.0])
    m.medium_alg = A.MEDIUM_DGADI
    with pytest.raises(A.CBMError, match="FTCS"):
        Context(m, [sim(4, box)])
    m = medium_model(2, (16, 16), 1.0, 0.0, 0.1)
    m.medium_border[4] = A.BORDER_COPY          # z- of a 2-D model
    with pytest.raises(A.CBMError, match="axis"):
        Context(m, [sim(4, box)])
    m = medium_model(2, (16, 16), 1.0, 0.0, 0.1)
    m.medium_border[0] = 7
    with pytest.raises(A.CBMError, match="unknown"):
        Context(m, [sim(4, box)])
