"""The oracle's DGADI medium step (Douglas-Gunn ADI, src/integrators.jl:
821-935 with the tridiagonal solvers of :141-428) pinned by closed forms:
a separable sine mode with zero boundaries is an eigenvector of every
sweep, each dividing it by 1 + 4 D r sin^2(pi k / (2 (n - 1))) with
r = dt / (dims dx^2) of the last axis; boundary nodes stay at their initial
values whatever the reaction does.  CPU only."""
import numpy as np

import cbm._abi as A
import oracle as O
from helpers import sim, uniform_positions
from test_oracle_medium import medium_model


def adi_model(dims, n, D, uptake, dt, c0=1.0):
    m = medium_model(dims, n, D, uptake, dt, c0)
    m.medium_alg = A.MEDIUM_DGADI
    return m


def mode(n, k):
    axes = [np.sin(np.pi * k[a] * np.arange(n[a]) / (n[a] - 1)) for a in range(len(n))]
    c = axes[0]
    for a in range(1, len(n)):
        c = np.multiply.outer(axes[a], c)   # x fastest in the flattened field
    return c.ravel()


def factor(n, k, D, dt, box_last, dims):
    dx = box_last / n[-1]
    r = dt / (dims * dx * dx)
    f = 1.0
    for a in range(dims):
        f *= 1 + 4 * D * r * np.sin(np.pi * k[a] / (2 * (n[a] - 1))) ** 2
    return f


def run_mode(n, k, box, D=1.0, dt=0.5, steps=6):
    dims = len(n)
    m = adi_model(dims, n, D, 0.0, dt)
    o = O.Oracle(m, sim(0, [[0.0, b] for b in box]))
    c0 = mode(n, k)
    o.set_medium(c0)
    assert o.init() == 0
    f = factor(n, k, D, dt, box[-1], dims)
    for t in range(1, steps + 1):
        o.step(1)
        got = o.medium(c0.size)
        assert np.max(np.abs(got - c0 / f ** t)) < 1e-13 * np.max(np.abs(c0)) + 1e-15


def test_adi_1d_mode():
    run_mode((33,), (3,), (16.0,))


def test_adi_2d_mode_uses_last_axis_spacing():
    # dx = 0.5 along x, 0.25 along y: both sweeps use r of y
    run_mode((24, 40), (2, 5), (12.0, 10.0))


def test_adi_3d_mode():
    run_mode((12, 10, 9), (1, 2, 3), (6.0, 5.0, 9.0), steps=4)


def test_adi_boundary_held_at_initial_values():
    n = (16, 12)
    box = [[0.0, 16.0], [0.0, 12.0]]
    m = adi_model(2, n, 0.8, 0.5, 0.3)
    N = 300
    pos = uniform_positions(N, box, 5)
    o = O.Oracle(m, sim(N, box))
    o.set(A.F_X, pos[0])
    o.set(A.F_Y, pos[1])
    c0 = np.ones(n[0] * n[1])
    grid = c0.reshape(n[1], n[0])
    grid[0, :] = grid[-1, :] = grid[:, 0] = grid[:, -1] = 5.0
    o.set_medium(c0)
    assert o.init() == 0
    o.step(5)
    got = o.medium(c0.size).reshape(n[1], n[0])
    edge = np.zeros_like(got, dtype=bool)
    edge[0, :] = edge[-1, :] = edge[:, 0] = edge[:, -1] = True
    assert np.all(got[edge] == 5.0)
    assert np.all(np.isfinite(got)) and np.any(got[~edge] != 1.0)
