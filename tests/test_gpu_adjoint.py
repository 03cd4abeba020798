"""Adjoint of the overdamped-Euler mechanics (cbm_adjoint) on the device:
loss equal to the oracle's, gradient d loss / d(k, r_eq, eps) equal to the
oracle's forward-mode dual-number gradient (the stand-in for ForwardDiff on
the CPU reference) and to central finite differences of the device loss
within 1e-6 relative."""
import os

import numpy as np
import pytest

import cbm._abi as A
import oracle as O
from cbm.api import Context
from helpers import model, sim, uniform_positions

pytestmark = pytest.mark.gpu

THETA_STAR = (1.3, 1.1, 0.08)
THETA0 = np.array([1.0, 1.0, 0.05])


def setup(n=512, steps=40, saves=(10, 20, 30, 40), seed=5):
    L = (n * 4 / 3 * np.pi / 0.3) ** (1 / 3)
    box = [[0.0, L]] * 3
    m = model(3, integrator=A.EULER, neighbors=A.VERLET_DISPLACEMENT, law=A.LAW_REP_ADH, dt=0.01,
              skin=3.0, max_neighbors=64, rebuild_disp=0.15)
    pos = uniform_positions(n, box, seed)
    # target trajectory from theta*
    o = O.Oracle(m, sim(n, box, list(THETA_STAR) + [2.4]))
    for a in range(3):
        o.set(a, pos[a])
    o.init()
    tgt, done = [], 0
    for s in saves:
        o.step(s - done)
        done = s
        tgt.append(np.stack([o.get(a) for a in range(3)]))
    return m, box, pos, np.array(saves), np.stack(tgt), steps


def device(m, box, pos, thetas):
    n = pos[0].size
    g = Context(m, [sim(n, box, list(t) + [2.4]) for t in thetas])
    for q in range(len(thetas)):
        for a in range(3):
            g.upload(a, pos[a], sim=q)
    g.neighbors()
    return g


def test_adjoint_matches_dual_numbers(gpu_available):
    m, box, pos, saves, tgt, steps = setup()
    n = pos[0].size
    o = O.Oracle(m, sim(n, box, list(THETA0) + [2.4]))
    for a in range(3):
        o.set(a, pos[a])
    o.init()
    ref_loss, ref_grad = o.traj_grad_dual(THETA0, steps, saves, tgt.ravel())
    g = device(m, box, pos, [THETA0])
    x0 = g.download(0, n)
    loss, grad = g.adjoint(steps, 20, saves, [tgt])
    assert abs(loss[0] - ref_loss) <= 1e-12 * abs(ref_loss)
    assert np.allclose(grad[0], ref_grad, rtol=1e-9, atol=1e-14 * np.abs(ref_grad).max())
    # the community is left at its initial state, and a second call agrees
    assert np.array_equal(g.download(0, n), x0)
    loss2, grad2 = g.adjoint(steps, 7, saves, [tgt])
    assert np.allclose(grad2, grad, rtol=1e-12)


def test_adjoint_matches_central_differences(gpu_available):
    m, box, pos, saves, tgt, steps = setup(n=1000, steps=30, saves=(15, 30), seed=7)
    thetas = [THETA0.copy()]
    # small h: the law is continuous with slope jumps at 2 r_eq and dc, so
    # central differences carry an O(h) error from pairs straddling a kink
    hs = 1e-7 * np.maximum(np.abs(THETA0), 1e-2)
    for q in range(3):
        for sgn in (1, -1):
            t = THETA0.copy()
            t[q] += sgn * hs[q]
            thetas.append(t)
    g = device(m, box, pos, thetas)
    loss, grad = g.adjoint(steps, 10, saves, [tgt] * len(thetas))
    for q in range(3):
        fd = (loss[1 + 2 * q] - loss[2 + 2 * q]) / (2 * hs[q])
        assert abs(fd - grad[0, q]) <= 1e-6 * abs(grad[0, q]), (q, fd, grad[0, q])


def test_batched_adjoint_parameter_sets(gpu_available):
    m, box, pos, saves, tgt, steps = setup(n=512, steps=20, saves=(10, 20), seed=9)
    rng = np.random.default_rng(1)
    thetas = [THETA0 * rng.uniform(0.9, 1.1, 3) for _ in range(5)]
    g = device(m, box, pos, thetas)
    loss, grad = g.adjoint(steps, 5, saves, [tgt] * len(thetas))
    for q, t in enumerate(thetas):
        o = O.Oracle(m, sim(pos[0].size, box, list(t) + [2.4]))
        for a in range(3):
            o.set(a, pos[a])
        o.init()
        rl, rg = o.traj_grad_dual(t, steps, saves, tgt.ravel())
        assert abs(loss[q] - rl) <= 1e-12 * abs(rl)
        assert np.allclose(grad[q], rg, rtol=1e-9)


def test_adjoint_without_list_snapshots_is_identical(gpu_available):
    """The backward sweep that re-validates and rebuilds lists gives the same
    bits as the one on the forward pass's list snapshots (lists are sorted by
    slot, so the active pairs are summed in the same order)."""
    m, box, pos, saves, tgt, steps = setup(n=512, steps=30, saves=(15, 30), seed=11)
    g = device(m, box, pos, [THETA0])
    loss_s, grad_s = g.adjoint(steps, 8, saves, [tgt])
    os.environ["CBM_ADJ_SNAPSHOT_MB"] = "0"
    try:
        loss_r, grad_r = g.adjoint(steps, 8, saves, [tgt])
    finally:
        del os.environ["CBM_ADJ_SNAPSHOT_MB"]
    assert np.array_equal(loss_s, loss_r)
    assert np.array_equal(grad_s, grad_r)
