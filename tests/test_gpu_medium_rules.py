"""Agent code reading and adding to the medium at its node (the medium
variable in agent code, src/AgentStructure/auxiliar.jl:353-365; the docs'
secretion idiom, docs/src/Usage_Agent.md:390-395): a user rule body's
RuleAgent::medium() and medium_add() on the device against the oracle, bit
for bit (additions are summed per node in fixed point, so their order does
not matter), with diffusion and uptake, in 2-D and 3-D.  Using the medium
in a model without one is reported."""
import numpy as np
import pytest

import cbm._abi as A
import oracle as O
import rule_bodies as RB
from cbm.api import Context
from helpers import model, sim, uniform_positions

pytestmark = pytest.mark.gpu


def field(g, size):
    got = np.empty(size)
    A.check(g.lib.cbm_download_medium(g.h, got.ctypes.data, got.size), g.h)
    return got


@pytest.mark.parametrize("dims", [2, 3])
def test_secreting_agents_match_the_oracle(gpu_available, dims):
    n = (32, 24, 8)[:dims]
    box = [[0.0, 32.0], [0.0, 24.0], [0.0, 8.0]][:dims]
    m = model(dims, neighbors=A.CELL_LINKED, law=A.LAW_LINEAR, dt=0.1, rule=A.RULE_CUSTOM, cell_edge=2.0)
    m.rule_p[0] = 0.37
    for a in range(dims):
        m.medium_n[a] = n[a]
    m.medium_D, m.medium_consume, m.medium_c0 = 0.8, 0.05, 0.2
    N = 700
    s = sim(N, box, [1.0, 2.0], n_max=N)
    pos = uniform_positions(N, box, 61 + dims)
    o = O.Oracle(m, s)
    o.set_rule(RB.secrete(0.37))
    g = Context(m, [s])
    g.compile_model(rule=RB.SECRETE)
    for a, v in enumerate(pos):
        o.set(a, v)
        g.upload(a, v)
    assert o.init() == 0
    g.neighbors()
    size = int(np.prod(n))
    for k in (1, 6, 20):
        o.step(k)
        g.step(k)
        assert np.array_equal(field(g, size), o.medium(size)), k
        assert np.array_equal(g.download(A.F_U0, N), o.get(A.F_U0)), k
        for a in range(dims):
            assert np.array_equal(g.download(a, N), o.get(a)), (k, a)


def test_medium_access_without_a_medium_is_reported(gpu_available):
    box = [[0.0, 16.0], [0.0, 16.0]]
    m = model(2, neighbors=A.FULL, dt=0.1, rule=A.RULE_CUSTOM)
    g = Context(m, [sim(10, box, n_max=10)])
    g.compile_model(rule=RB.SECRETE)
    for a, v in enumerate(uniform_positions(10, box, 3)):
        g.upload(a, v)
    g.neighbors()
    g.step(1)
    with pytest.raises(A.CBMError, match="medium"):
        g.sync()
