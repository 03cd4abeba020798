"""Pins of the oracle's rod model (src/models/rods.jl:21-121, the geometry of
src/metrics.jl:138-246) by closed forms of the geometry and mechanics, not by
restating its formulas: contact points of crossing, T-shaped, parallel,
collinear and skew rods in every branch of rodIntersection; the contact force
of touching rods from Hertz-like overlap, normal damping and both branches of
the capped tangential friction; Newton's third law; the negative-overlap
error; free rods under Euler, Heun and RK4 (the linear decay factors of each
integrator); walk independence of the neighbor algorithms."""
import math

import numpy as np
import pytest

import cbm._abi as A
import oracle as O
from helpers import model, sim

SEP = 0.99          # rodIntersection's separation


def close(a, b, tol=1e-13):
    a, b = np.asarray(a, dtype=float), np.asarray(b, dtype=float)
    assert np.all(np.abs(a - b) <= tol * (1.0 + np.abs(b))), (a, b)


# ------------------------------------------------------------ contact points
def test_crossing_rods_meet_inside_both():
    # axes cross at (0.3, 0): each contact sphere sits 0.99 of the way from
    # its centre to the crossing
    q = O.rod_intersection([0.0, 0.0, 2.0, 0.0], [0.3, 0.2, 2.0, math.pi / 2])
    close(q, [SEP * 0.3, 0.0, 0.3, 0.2 - SEP * 0.2])


def test_t_shape_projects_the_clamped_end():
    # rod 2 stands on rod 1's right half: its end (pulled in to 0.99 l/2)
    # and that end's foot on rod 1
    q = O.rod_intersection([0.0, 0.0, 2.0, 0.0], [0.5, 1.2, 2.0, math.pi / 2])
    close(q, [0.5, 0.0, 0.5, 1.2 - SEP * 1.0])
    # the same pair seen from rod 2 (the mirrored branch)
    q = O.rod_intersection([0.5, 1.2, 2.0, math.pi / 2], [0.0, 0.0, 2.0, 0.0])
    close(q, [0.5, 1.2 - SEP * 1.0, 0.5, 0.0])


def test_parallel_rods_use_the_midpoint():
    # side by side 0.8 apart and shifted 0.5 along the axis (angle 0.3): the
    # midpoint of the centres projects to +-0.25 along each axis
    t = 0.3
    u = np.array([math.cos(t), math.sin(t)])
    n = np.array([-math.sin(t), math.cos(t)])
    c1 = np.array([1.0, -2.0])
    c2 = c1 + 0.8 * n + 0.5 * u
    q = O.rod_intersection([*c1, 2.0, t], [*c2, 2.0, t])
    close(q[:2], c1 + SEP * 0.25 * u, 1e-12)
    close(q[2:], c2 - SEP * 0.25 * u, 1e-12)


def test_collinear_rods_touch_at_their_ends():
    q = O.rod_intersection([0.0, 0.0, 2.0, 0.0], [2.5, 0.0, 2.0, 0.0])
    close(q, [SEP * 1.0, 0.0, 2.5 - SEP * 1.0, 0.0])


def test_skew_rods_outside_both_keep_the_clamped_far_point():
    # the axes meet at (3, 0), outside both rods: rod 2's clamped point and,
    # since its foot lies beyond rod 1, rod 1's clamped point; rod 1's point
    # projects outside rod 2, so rod 2 keeps its own
    q = O.rod_intersection([0.0, 0.0, 2.0, 0.0], [3.0, 2.0, 2.0, math.pi / 2])
    close(q, [SEP * 1.0, 0.0, 3.0, 2.0 - SEP * 1.0])


def test_skew_rods_outside_both_take_the_foot_inside_rod_two():
    # rod 2 at 135 degrees, centred 1.5 from the meeting point (3, 0): rod 1's
    # clamped end (0.99, 0) has its foot on rod 2 within l/2 of its centre,
    # and that foot replaces rod 2's point
    r = 1.5 / math.sqrt(2.0)
    q = O.rod_intersection([0.0, 0.0, 2.0, 0.0], [3.0 - r, r, 2.0, 3 * math.pi / 4])
    close(q, [0.99, 0.0, 3.0 - 2.01 / 2, 2.01 / 2], 1e-12)


def test_contact_points_move_with_the_rods():
    # translating and rotating both rods moves the contact points with them
    rng = np.random.default_rng(7)
    for _ in range(200):
        r1 = [rng.uniform(-2, 2), rng.uniform(-2, 2), rng.uniform(0.5, 3), rng.uniform(-3, 3)]
        r2 = [rng.uniform(-2, 2), rng.uniform(-2, 2), rng.uniform(0.5, 3), rng.uniform(-3, 3)]
        q = O.rod_intersection(r1, r2).reshape(2, 2)
        phi, sh = rng.uniform(-3, 3), rng.uniform(-5, 5, 2)
        R = np.array([[math.cos(phi), -math.sin(phi)], [math.sin(phi), math.cos(phi)]])
        m1 = [*(R @ r1[:2] + sh), r1[2], r1[3] + phi]
        m2 = [*(R @ r2[:2] + sh), r2[2], r2[3] + phi]
        qm = O.rod_intersection(m1, m2).reshape(2, 2)
        close(qm, (R @ q.T).T + sh, 1e-8)


# -------------------------------------------------------------------- forces
K = [50.0, 2.0, 3.0, 0.4]       # kn, gamma_n, gamma_t, mu_cc


def rod(x, y, th=0.0, vx=0.0, vy=0.0, d=1.0, l=2.0, m=1.0):
    return [x, y, d, l, th, vx, vy, m]


def test_side_by_side_rods_push_apart_along_the_normal():
    # parallel and centred 0.8 apart: overlap 0.2, force kn 0.2^1.5 straight
    # down on the lower rod, no torque
    f, rc = O.rod_forces(rod(0, 0), rod(0, 0.8), K)
    assert rc == 0
    close(f, [0.0, -K[0] * 0.2 ** 1.5, 0.0])


def test_offset_contact_gives_the_lever_arm_torque():
    # rod 2 shifted 0.5 along x: the contact spheres at x = 0.2475 and
    # 0.2525; the force (kn delta^1.5 along the normal) acts 0.2475 right of
    # rod 1's centre: W = 0.2475 Fy < 0 (clockwise)
    f, rc = O.rod_forces(rod(0, 0), rod(0.5, 0.8), K)
    a, b = np.array([0.2475, 0.0]), np.array([0.2525, 0.8])
    r = np.linalg.norm(a - b)
    nrm = (a - b) / r
    F = K[0] * (1.0 - r) ** 1.5 * nrm
    close(f, [F[0], F[1], 0.2475 * F[1]], 1e-12)
    assert f[2] < 0


def test_normal_damping_adds_to_the_overlap_force():
    # approaching at 0.3 along the normal (tangential part zero)
    m = 2.0
    f, _ = O.rod_forces(rod(0, 0, vy=0.3, m=m), rod(0, 0.8), K)
    close(f, [0.0, -(K[0] * 0.2 ** 1.5 + K[1] * (m / 2) * 0.2 * 0.3), 0.0], 1e-13)


@pytest.mark.parametrize("capped", [False, True])
def test_tangential_friction_takes_the_smaller_bound(capped):
    # sliding at 0.5 along the axis: friction -min(gt m/2 sqrt(delta), mu Fn)
    # times the tangential velocity
    k = [K[0], K[1], 0.5 if not capped else 400.0, K[3]]
    f, _ = O.rod_forces(rod(0, 0, vx=0.5), rod(0, 0.8), k)
    fn = k[0] * 0.2 ** 1.5
    ft = -(k[2] * 0.5 * math.sqrt(0.2) if not capped else k[3] * fn)
    close(f, [ft * 0.5, -fn, 0.0], 1e-13)


def test_newton_third_law_for_equal_rods():
    rng = np.random.default_rng(3)
    hits = 0
    for _ in range(300):
        a = rod(0.0, 0.0, rng.uniform(-3, 3), rng.normal(), rng.normal(), l=rng.uniform(1, 3))
        b = rod(rng.uniform(-1.5, 1.5), rng.uniform(-1.5, 1.5), rng.uniform(-3, 3), rng.normal(), rng.normal(),
                l=rng.uniform(1, 3))
        q = O.rod_intersection([a[0], a[1], a[3], a[4]], [b[0], b[1], b[3], b[4]])
        # the third law holds where both rods see the same contact pair
        qb = O.rod_intersection([b[0], b[1], b[3], b[4]], [a[0], a[1], a[3], a[4]])
        if np.max(np.abs(q - np.r_[qb[2:], qb[:2]])) > 1e-9:
            continue
        fa, _ = O.rod_forces(a, b, K)
        fb, _ = O.rod_forces(b, a, K)
        close(fa[:2], -fb[:2], 1e-9)
        hits += fa[0] != 0.0
    assert hits > 20


def test_no_contact_and_coincident_spheres_add_nothing():
    f, rc = O.rod_forces(rod(0, 0), rod(0, 1.0), K)            # r = (d + d)/2
    assert rc == 0 and np.all(f == 0.0)
    f, rc = O.rod_forces(rod(0, 0), rod(0, 0), K)              # the rod itself
    assert rc == 0 and np.all(f == 0.0)


def test_negative_overlap_is_an_error():
    # d_i = 0.5 < r = 0.8 < (0.5 + 1.5)/2: the reference's (d_i - r)^1.5 fails
    f, rc = O.rod_forces(rod(0, 0, d=0.5), rod(0, 0.8, d=1.5), K)
    assert rc == -1


# ------------------------------------------------------------------ dynamics
P = [50.0, 2.0, 3.0, 0.4, 0.0, 0.7, 1.3]    # kn, gn, gt, mucc, mucw, beta, beta_w


def free_rod(integ, steps, dt=0.01):
    o = O.Oracle(model(2, integrator=integ, neighbors=A.FULL, law=A.LAW_RODS, dt=dt), sim(1, [[0, 10]] * 2, P))
    o.set(A.F_X, [1.0]); o.set(A.F_Y, [2.0])
    o.set(A.F_VX, [0.4]); o.set(A.F_VY, [-0.2])
    o.set(A.F_THETA, [0.1]); o.set(A.F_OMEGA, [0.5])
    assert o.init() == 0
    o.step(steps)
    assert o.error() == 0
    return o


@pytest.mark.parametrize("integ", [A.EULER, A.HEUN, A.RK4])
def test_free_rod_decays_by_the_integrators_factor(integ):
    # dv/dt = -beta v, domega/dt = -beta_w omega are linear: one step
    # multiplies them by the integrator's polynomial in h = beta dt; x and
    # theta integrate them with the matching weights
    n, dt = 40, 0.01
    o = free_rod(integ, n, dt)

    def factor(h):
        return {A.EULER: 1 - h, A.HEUN: 1 - h + h * h / 2,
                A.RK4: 1 - h + h * h / 2 - h ** 3 / 6 + h ** 4 / 24}[integ]

    def gain(h):       # x_{k+1} - x_k = gain * v_k dt for the same scheme
        return {A.EULER: 1.0, A.HEUN: 1 - h / 2, A.RK4: 1 - h / 2 + h * h / 6 - h ** 3 / 24}[integ]

    for v0, rate, fv, fx, x0 in [(0.4, P[5], A.F_VX, A.F_X, 1.0), (-0.2, P[5], A.F_VY, A.F_Y, 2.0),
                                 (0.5, P[6], A.F_OMEGA, A.F_THETA, 0.1)]:
        h = rate * dt
        g = factor(h)
        close(o.get(fv), [v0 * g ** n], 1e-12)
        close(o.get(fx), [x0 + gain(h) * dt * v0 * (1 - g ** n) / (1 - g)], 1e-12)
    assert o.get(A.F_FX)[0] == 0.0 and o.get(A.F_PRESSURE)[0] == 0.0


def rods_box(neigh, n=300, L=18.0, seed=4, **kw):
    rng = np.random.default_rng(seed)
    m = model(2, integrator=A.HEUN, neighbors=neigh, law=A.LAW_RODS, dt=0.002, **kw)
    o = O.Oracle(m, sim(n, [[0, L]] * 2, P))
    o.set(A.F_X, rng.uniform(1, L - 1, n)); o.set(A.F_Y, rng.uniform(1, L - 1, n))
    o.set(A.F_VX, rng.normal(0, 0.2, n)); o.set(A.F_VY, rng.normal(0, 0.2, n))
    o.set(A.F_THETA, rng.uniform(-3, 3, n)); o.set(A.F_OMEGA, rng.normal(0, 0.3, n))
    o.set(A.F_LENGTH, rng.uniform(1.0, 2.5, n)); o.set(A.F_MASS, rng.uniform(0.5, 2.0, n))
    assert o.init() == 0
    return o


def test_neighbor_algorithms_give_the_same_motion():
    # one cell holding every rod, and Verlet lists wider than the box, give
    # every rod all the others in ascending order, as Full does: the same
    # sums bit for bit.  (Narrower cells or lists can differ: a foot that
    # rodIntersection does not clamp, src/metrics.jl:229-236, can put a
    # contact point outside its rod, so contacts are not bounded by l + d.)
    full = rods_box(A.FULL, n=150)
    cl = rods_box(A.CELL_LINKED, n=150, cell_edge=20.0)
    vl = rods_box(A.VERLET_DISPLACEMENT, n=150, skin=30.0, max_neighbors=256)
    for o in (full, cl, vl):
        o.step(30)
        assert o.error() == 0
    assert np.any(full.get(A.F_PRESSURE) > 0)
    for f in (A.F_X, A.F_Y, A.F_VX, A.F_THETA, A.F_OMEGA, A.F_TORQUE, A.F_PRESSURE):
        assert np.array_equal(cl.get(f), full.get(f)), f
        assert np.array_equal(vl.get(f), full.get(f)), f


def test_phantom_contact_beyond_the_rod():
    # rod 2 at a shallow angle whose axis crosses rod 1's: its clamped end is
    # projected onto rod 1's axis without a clamp, 9 units from rod 1's
    # centre, and the pair touches there (the reference's geometry)
    t = math.atan2(0.5, 10.0)
    q = O.rod_intersection([0.0, 0.0, 2.0, 0.0], [10.0, 0.5, 2.0, t])
    u = np.array([math.cos(t), math.sin(t)])
    b = np.array([10.0, 0.5]) - SEP * 1.0 * u
    close(q, [b[0], 0.0, b[0], b[1]], 1e-12)
    f, rc = O.rod_forces(rod(0, 0), rod(10.0, 0.5, th=t), K)
    assert rc == 0 and f[1] < 0


def test_momentum_balance_of_equal_rods():
    # equal diameters and masses, no drag: contacts move momentum between
    # rods only where the two see the same contact pair; two parallel rods
    # overlapping side by side exchange exactly opposite forces
    p = list(P)
    p[5] = 0.0
    o = O.Oracle(model(2, integrator=A.EULER, neighbors=A.FULL, law=A.LAW_RODS, dt=0.01), sim(2, [[0, 10]] * 2, p))
    o.set(A.F_X, [5.0, 5.0]); o.set(A.F_Y, [5.0, 5.8])
    o.set(A.F_VX, [0.1, -0.1]); o.set(A.F_VY, [0.05, 0.0])
    assert o.init() == 0
    o.step(5)
    close(o.get(A.F_VX).sum(), 0.0, 1e-15)
    close(o.get(A.F_VY).sum(), 0.05, 1e-14)
    close(o.get(A.F_FY), -o.get(A.F_FY)[::-1], 1e-15)


def test_rod_model_refusals():
    for kw in ({"dims": 3}, {"periodic": 1}, {"integrator": A.EM}):
        dims = kw.pop("dims", 2)
        o = O.Oracle(model(dims, law=A.LAW_RODS, **kw), sim(2, [[0, 10]] * dims, P))
        assert o.error() != 0


# ------------------------------------------------------------ rods2dGrowth
G = [0.4, 0.25, 3.0, 0.8]     # growth, sigma_lTarget, lMax, alpha


def growing(n, integ=A.EULER, neigh=A.FULL, dt=0.01, g=G, **kw):
    m = model(2, integrator=integ, neighbors=neigh, law=A.LAW_RODS, dt=dt, rule=A.RULE_RODS_GROWTH, rule_p=g, **kw)
    return O.Oracle(m, sim(n, [[0, 40]] * 2, P, n_max=4 * n + 8))


def test_free_rod_grows_linearly_and_divides_past_its_target():
    # no contacts, no pressure: dt(l) = growth = 0.5, dt = 1/8, so l gains
    # exactly 1/16 a step; the rule divides only when l > lTarget
    g = [0.5, 0.25, 3.0, 0.8]
    o = growing(1, dt=0.125, g=g)
    o.set(A.F_X, [20.0]); o.set(A.F_Y, [20.0]); o.set(A.F_LENGTH, [1.0]); o.set(A.F_LTARGET, [3.0])
    assert o.init() == 0
    o.step(32)
    assert o.n == 1 and o.error() == 0
    assert o.get(A.F_LENGTH)[0] == 3.0
    o.step(1)                                  # l == lTarget: no division
    assert o.n == 1 and o.get(A.F_LENGTH)[0] == 3.0625
    o.step(1)                                  # l > lTarget
    assert o.n == 2


def test_division_places_two_daughters_along_the_axis():
    th, l, d, om = 0.7, 3.2, 0.9, 0.3
    o = growing(1)
    o.set(A.F_X, [20.0]); o.set(A.F_Y, [10.0]); o.set(A.F_THETA, [th]); o.set(A.F_OMEGA, [om])
    o.set(A.F_LENGTH, [l]); o.set(A.F_DIAMETER, [d]); o.set(A.F_LTARGET, [3.0]); o.set(A.F_VX, [0.2])
    assert o.init() == 0
    o.step(1)
    assert o.n == 2 and sorted(o.get(A.F_ID, 2)) == [2, 3]
    ids = list(o.get(A.F_ID, 2))
    a, b = ids.index(2), ids.index(3)
    q = (l + d) / 4
    x, y = o.get(A.F_X), o.get(A.F_Y)
    close([x[a], y[a]], [20.0 + q * math.cos(th), 10.0 + q * math.sin(th)], 1e-15)
    close([x[b], y[b]], [20.0 - q * math.cos(th), 10.0 - q * math.sin(th)], 1e-15)
    assert np.all(o.get(A.F_LENGTH) == (l - d) / 2)
    w = o.get(A.F_OMEGA)
    assert abs((w[a] - om) + (w[b] - om)) < 1e-16 and abs(w[a] - om) <= 0.1
    lt = o.get(A.F_LTARGET)
    assert np.all((lt >= G[2] - G[1]) & (lt < G[2] + G[1])) and lt[a] != lt[b]
    assert np.all(o.get(A.F_VX) == 0.2) and np.all(o.get(A.F_THETA) == th)


def test_division_draws_are_uniform():
    # ww uniform on [-0.1, 0.1) and lTarget on [lMax - sigma, lMax + sigma)
    from scipy import stats
    n = 3000
    o = growing(n, dt=0.01)
    rng = np.random.default_rng(9)
    o.set(A.F_X, rng.uniform(0, 40, n) * 1e3); o.set(A.F_Y, np.zeros(n))   # far apart: no contacts
    o.set(A.F_LENGTH, np.full(n, 2.0)); o.set(A.F_LTARGET, np.full(n, 1.0)); o.set(A.F_OMEGA, np.zeros(n))
    assert o.init() == 0
    o.step(1)
    assert o.n == 2 * n
    w = o.get(A.F_OMEGA)
    assert np.all(w[0::2] + w[1::2] == 0.0) or np.allclose(np.sort(w), -np.sort(w)[::-1])
    assert stats.kstest(np.abs(w), stats.uniform(0, 0.1).cdf).pvalue > 1e-3
    lt = o.get(A.F_LTARGET)
    assert stats.kstest(lt, stats.uniform(G[2] - G[1], 2 * G[1]).cdf).pvalue > 1e-3


def test_pressure_slows_growth():
    # two overlapping parallel rods: dt(l) = growth/(1 + alpha p) with p = |F|
    o = growing(2, dt=0.01)
    o.set(A.F_X, [5.0, 5.0]); o.set(A.F_Y, [5.0, 5.8]); o.set(A.F_LTARGET, [9.0, 9.0])
    assert o.init() == 0
    o.step(1)
    p = P[0] * 0.2 ** 1.5                      # kn delta^1.5 at the start of the step
    close(o.get(A.F_LENGTH), [1.0 + 0.01 * G[0] / (1 + G[3] * p)] * 2, 1e-14)
