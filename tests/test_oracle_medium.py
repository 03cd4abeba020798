"""Pins the oracle's medium step (FTCS, zero-flux ghost mirror, agent
uptake) by closed forms and invariants.  CPU only."""
import numpy as np

import cbm._abi as A
import oracle as O
from helpers import model, sim


def medium_model(dims, n, D, uptake, dt, c0=1.0):
    m = model(dims, neighbors=A.FULL, dt=dt)
    for a in range(dims):
        m.medium_n[a] = n[a]
    m.medium_D, m.medium_consume, m.medium_c0 = D, uptake, c0
    return m


def test_fourier_mode_decays_by_exact_factor():
    # cell-centred Neumann grid: cos(pi k (i + 1/2) / n) is an eigenvector of the
    # mirrored 3-point Laplacian with eigenvalue -4 sin^2(pi k / 2n) / dx^2
    nx, ny, L = 48, 32, (24.0, 8.0)
    D, dt = 1.0, 0.01
    m = medium_model(2, (nx, ny), D, 0.0, dt)
    o = O.Oracle(m, sim(0, [[0.0, L[0]], [0.0, L[1]]]))
    kx, ky = 5, 3
    i = np.arange(nx)[None, :]
    j = np.arange(ny)[:, None]
    c = np.cos(np.pi * kx * (i + 0.5) / nx) * np.cos(np.pi * ky * (j + 0.5) / ny)
    o.set_medium(c.ravel())
    o.init()
    dx, dy = L[0] / nx, L[1] / ny
    lam = -4 * np.sin(np.pi * kx / (2 * nx)) ** 2 / dx ** 2 - 4 * np.sin(np.pi * ky / (2 * ny)) ** 2 / dy ** 2
    for k in range(1, 6):
        o.step(1)
        want = c * (1 + dt * D * lam) ** k
        assert np.max(np.abs(o.medium(nx * ny).reshape(ny, nx) - want)) < 1e-13


def test_mass_conserved_without_uptake():
    n = (40, 30, 20)
    m = medium_model(3, n, 0.7, 0.0, 0.002)   # D dt / dx^2 = 0.14, 0.14, 0.14
    o = O.Oracle(m, sim(0, [[0.0, 4.0], [0.0, 3.0], [0.0, 2.0]]))
    rng = np.random.default_rng(2)
    c = rng.uniform(0, 1, np.prod(n))
    o.set_medium(c)
    o.init()
    o.step(20)
    assert abs(o.medium(c.size).sum() - c.sum()) < 1e-12 * c.sum()


def test_uptake_of_agents_in_one_node():
    m = medium_model(2, (10, 10), 0.0, 0.1, 0.2)
    o = O.Oracle(m, sim(3, [[0.0, 10.0], [0.0, 10.0]]))
    o.set(A.F_X, [3.2, 3.7, 8.0])     # two agents in node (3, 5), one in node (8, 5)
    o.set(A.F_Y, [5.5, 5.1, 5.9])
    o.init()
    o.step(1)
    c = o.medium(100).reshape(10, 10)
    assert c[5, 3] == 1.0 + 0.2 * (0.0 - (0.1 * 1.0) * 2.0)
    assert c[5, 8] == 1.0 + 0.2 * (0.0 - (0.1 * 1.0) * 1.0)
    assert np.count_nonzero(c != 1.0) == 2
