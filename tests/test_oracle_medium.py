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


def bordered(m, borders):
    """Sets the model's medium borders from ("mirror" | "copy" | value) per side."""
    for k, b in enumerate(borders):
        if b == "copy":
            m.medium_border[k] = A.BORDER_COPY
        elif b != "mirror":
            m.medium_border[k] = A.BORDER_DIRICHLET
            m.medium_border_value[k] = float(b)
    return m


def test_dirichlet_ends_relax_to_the_linear_profile():
    # fixed border nodes at 1 and 0 and interior nodes integrating the
    # 3-point Laplacian: the fixed point is linear in the node index, and
    # FTCS reaches it (slowest mode factor 1 - 4 r sin^2(pi / 22) per step)
    n, r = 12, 0.2
    m = bordered(medium_model(1, (n,), 1.0, 0.0, r, c0=0.3), [1.0, 0.0])
    o = O.Oracle(m, sim(0, [[0.0, float(n)]]))
    o.init()
    o.step(3000)
    c = o.medium(n)
    assert c[0] == 1.0 and c[-1] == 0.0
    assert np.max(np.abs(c - (1.0 - np.arange(n) / (n - 1)))) < 1e-13


def test_dirichlet_x_and_mirror_y_give_a_profile_constant_in_y():
    nx, ny = 10, 7
    m = bordered(medium_model(2, (nx, ny), 0.5, 0.0, 0.4, c0=0.0), [2.0, -1.0])
    o = O.Oracle(m, sim(0, [[0.0, float(nx)], [0.0, float(ny)]]))
    o.init()
    o.step(4000)
    c = o.medium(nx * ny).reshape(ny, nx)
    want = 2.0 - 3.0 * np.arange(nx) / (nx - 1)
    assert np.max(np.abs(c - want[None, :])) < 1e-12


def test_copy_and_dirichlet_borders_one_step_by_hand():
    # 4 x 4 nodes, c[j, i] = i + 10 j (dx = dy = 1), D = 0.1, dt = 0.5;
    # x- copies the inward neighbor, x+ is Dirichlet 2, y- copies, y+ mirrors.
    # Corners take the x side (the first side in x-, x+, y-, y+).  Inside
    # nodes see a linear field (no change); the top row's inner nodes see
    # lap_y = c[j=2] - c[j=3] = -10 through the mirror: c - 0.5.
    m = bordered(medium_model(2, (4, 4), 0.1, 0.0, 0.5), ["copy", 2.0, "copy", "mirror"])
    o = O.Oracle(m, sim(0, [[0.0, 4.0], [0.0, 4.0]]))
    i = np.arange(4)[None, :]
    j = np.arange(4)[:, None]
    o.set_medium((i + 10.0 * j).ravel())
    o.init()
    o.step(1)
    want = np.array([[1.0, 11.0, 12.0, 2.0],
                     [11.0, 11.0, 12.0, 2.0],
                     [21.0, 21.0, 22.0, 2.0],
                     [31.0, 30.5, 31.5, 2.0]])
    assert np.array_equal(o.medium(16).reshape(4, 4), want)


def test_user_assignments_follow_the_reference_idiom():
    # the docs' mediumODE: inside nodes diffuse, x- reflects (p = p[2]),
    # x+ absorbs (p = 0); here on a 1-D line, one step by hand from c = i^2:
    # inside lap = 2 everywhere, so c' = i^2 + dt D 2 = i^2 + 0.2
    m = medium_model(1, (6,), 1.0, 0.0, 0.1)
    o = O.Oracle(m, sim(0, [[0.0, 6.0]]))
    o.set_medium(np.arange(6.0) ** 2)

    def assign(nd):
        if nd.border[0] == -1:
            return nd.nb[0][1]
        if nd.border[0] == 1:
            return 0.0
        return None

    o.set_medium_body(lambda nd: 1.0 * nd.laplacian(), assign)
    o.init()
    o.step(1)
    assert np.array_equal(o.medium(6), [1.0, 1.2, 4.2, 9.2, 16.2, 0.0])
