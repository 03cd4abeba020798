"""The 1,048,576-agent 3-D workload of bench.py on the device, at full size.

Checks:
- Verlet list lengths equal the oracle's (and a k-d tree's) right after the build.
- States stay bit-identical to the oracle over 100 Heun steps that include rebuilds.
- Over 200 steps, the size-independent invariants of the model hold. Pair forces are
  antisymmetric and there are no walls, so the centre of mass stays put to rounding.
  Every agent keeps its id, and N is constant.
"""
import os
import sys

import numpy as np
import pytest

import cbm._abi as A
import oracle as O
from cbm.api import Context

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
import bench  # noqa: E402  (the workload definition)

pytestmark = pytest.mark.gpu

N = 1_048_576


def device(m, s, pos):
    g = Context(m, [s])
    for a in range(3):
        g.upload(a, pos[a])
    g.neighbors()
    return g


def test_full_size_lists_and_steps_match_oracle(gpu_available):
    m, s, pos, L = bench.workload("heun", N)
    print(f"full-size parity: N={N}, box {L:.3f}", flush=True)
    o = O.Oracle(m, s)
    for a in range(3):
        o.set(a, pos[a])
    assert o.init() == 0
    g = device(m, s, pos)
    lens = g.count_neighbors(0.0)
    assert np.array_equal(lens, o.count_neighbors(m.skin))
    # independent pin on a sample: agents within the list radius by a k-d tree
    from scipy.spatial import cKDTree
    P = np.stack(pos, 1)
    rng = np.random.default_rng(3)
    pick = rng.choice(N, 2000, replace=False)
    got = cKDTree(P).query_ball_point(P[pick], r=m.skin)
    d_kd = np.array([sum(1 for j in js if j != i and np.linalg.norm(P[j] - P[i]) < m.skin)
                     for i, js in zip(pick, got)])
    assert np.array_equal(lens[pick], d_kd)
    for k in (1, 4, 15, 40, 40):
        o.step(k)
        g.step(k)
        for a in range(3):
            assert np.array_equal(g.download(a, N), o.get(a)), (k, a)
    assert g.rebuilds() == o.rebuilds
    assert o.rebuilds >= 2


def test_full_size_invariants_over_200_steps(gpu_available):
    m, s, pos, L = bench.workload("heun", N)
    g = device(m, s, pos)
    com0 = np.array([np.mean(g.download(a, N)) for a in range(3)])
    ids0 = g.download(A.F_ID, N)
    g.step(200)
    assert g.n() == N
    com = np.array([np.mean(g.download(a, N)) for a in range(3)])
    # the sum of forces is zero up to rounding: the drift is ~1e-16 L over
    # 200 steps; the bound leaves room for the rounding of np.mean itself
    assert np.all(np.abs(com - com0) < 1e-12 * L), com - com0
    assert np.array_equal(g.download(A.F_ID, N), ids0)
    assert g.rebuilds() >= 2
