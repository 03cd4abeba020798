"""Pins of the oracle's model scope (modelStepRule! and modelStepDE!,
src/CommunityStructure/step.jl:21-30, 66-75, 96-107): a rule's counter and
its reads of t and N; Euler's closed form for a parameter relaxing to a
fixed point and an exact linear ramp; the Euler-Maruyama noise keyed as
documented; the model part of the reference KAT test/testCommunity.jl:
35-200; agents read the step-start parameters (a stiffness switched on by
the rule moves nothing until the next step)."""
import math

import numpy as np

import cbm._abi as A
import model_bodies as MB
import oracle as O
from helpers import model, sim

DT = 0.125


def run(body, steps, p, n=3, law=A.LAW_NONE, pos=None, seed=0):
    m = model(2, integrator=A.EULER, neighbors=A.FULL, law=law, dt=DT, seed=seed)
    o = O.Oracle(m, sim(n, [[0, 10]] * 2, p))
    if pos is not None:
        o.set(A.F_X, pos[0])
        o.set(A.F_Y, pos[1])
    o.set_model_body(**body)
    assert o.init() == 0
    o.step(steps)
    assert o.error() == 0
    return o


def test_rule_counts_and_reads_t_and_n():
    o = run(MB.combined(), 9, [1.0, 1.0, 0, 5.0, 0, 0, 0, 0])
    p = o.params()
    assert p[3] == 5.0 + 9
    t = 0.0
    for _ in range(8):            # the last step's start time
        t = t + DT
    assert p[4] == t
    assert p[5] == 3.0


def test_ode_closed_forms():
    n = 12
    o = run(MB.combined(), n, [0.5, 1.0, 3.0, 0, 0, 0, 0, 0])
    p = o.params()
    assert abs(p[0] - (2.0 + (0.5 - 2.0) * (1 - 0.5 * DT) ** n)) < 1e-14
    assert p[2] == 3.0 + n * DT         # exact: dt is a power of two


def test_noise_is_euler_maruyama_keyed_by_purpose():
    seed = 77
    o = run(MB.combined(), 1, [1.0, 1.0, 0, 0, 0, 0, 0.4, 0], seed=seed)
    z = O.normals(seed, 0, 0, 0, 0x600 + 3)[0]        # p6: pair 3, first normal
    expect = (0.4 + (-0.4) * DT) + 0.3 * z * math.sqrt(DT)
    assert o.params()[6] == expect
    assert o.params()[7] == 0.4                        # the rule read the step-start p6
    # over many seeds one step's p6 has mean p6 (1 - dt) and variance 0.09 dt
    vals = np.array([run(MB.combined(), 1, [1.0, 1.0, 0, 0, 0, 0, 0.4, 0], n=1, seed=s).params()[6]
                     for s in range(4000)])
    assert abs(vals.mean() - 0.4 * (1 - DT)) < 5 * 0.3 * math.sqrt(DT) / math.sqrt(4000)
    assert abs(vals.var() / (0.09 * DT) - 1.0) < 0.1


def test_reference_kat_model_part():
    # testCommunity.jl:79-88, 104-110: gi += 1, ga[1,2] = 7, dt(gf) = 1 from
    # gf = 1 with dt = 0.1
    m = model(2, integrator=A.EULER, neighbors=A.FULL, dt=0.1)
    o = O.Oracle(m, sim(10, [[0, 1]] * 2, [0.0, 1.0, 0.0, 0.0]))
    o.set_model_body(**MB.kat())
    assert o.init() == 0
    o.step(1)
    p = o.params()
    assert p[2] == 1.0 and p[3] == 7.0 and abs(p[1] - 1.1) < 1e-15


def test_agents_read_the_step_start_parameters():
    # two agents 0.5 apart under the linear law (contact 1.0) with k = 0;
    # the rule switches k to 1: the first step moves nothing, the second
    # pushes them apart by k (c - d) dt each
    body = dict(rule=lambda m: m.pn.__setitem__(0, 1.0))
    pos = ([4.0, 4.5], [5.0, 5.0])
    o = run(body, 1, [0.0, 1.0], n=2, law=A.LAW_LINEAR, pos=pos)
    assert np.array_equal(o.get(A.F_X), pos[0])
    o.step(1)
    assert np.array_equal(o.get(A.F_X), [4.0 - 0.5 * DT, 4.5 + 0.5 * DT])
