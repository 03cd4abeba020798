"""Pins the CPU oracle (oracle/cbm_oracle.cpp) before it is trusted as the
parity target: reference known-answer tests and printed outputs (golden
fixtures), published generator vectors, closed forms, invariants and brute
force.  CPU only."""
from fractions import Fraction

import numpy as np
import pytest

import cbm._abi as A
import oracle as O
from helpers import golden, lattice, model, sim, uniform_positions


# ------------------------------------------------------------------ Philox
def test_philox_known_answers():
    for v in golden("philox_kat.json")["vectors"]:
        got = ["%08x" % x for x in O.philox(v["ctr"], v["key"])]
        assert got == v["out"]


def test_normals_moments_and_ks():
    from scipy import stats
    z = np.array([O.normals(7, s, i, 0, 0x100) for s in range(20) for i in range(1, 1001)]).ravel()
    assert abs(z.mean()) < 4.0 / np.sqrt(z.size)
    assert abs(z.var() - 1.0) < 4.0 * np.sqrt(2.0 / z.size)
    assert stats.kstest(z, "norm").pvalue > 1e-3


# ------------------------------------------------------------ exact floor
def test_floor_div_matches_julia_div():
    # Julia: 1.0 ÷ 0.1 == 9.0 and 0.3 ÷ 0.1 == 2.0 (exact quotient of the binary values)
    assert O.floor_div(1.0, 0.1) == 9
    assert O.floor_div(0.3, 0.1) == 2
    rng = np.random.default_rng(3)
    for _ in range(3000):
        b = float(rng.uniform(0.05, 3.0))
        k = int(rng.integers(0, 200))
        a = float(np.nextafter(k * b, (k + 1) * b * float(rng.integers(0, 2))))
        exact = (Fraction(a) / Fraction(b)).__floor__()
        assert O.floor_div(a, b) == exact


# ------------------------------------------------------- neighbor lattice KAT
ALG = {"Full": A.FULL, "VerletTime": A.VERLET_TIME, "VerletDisplacement": A.VERLET_DISPLACEMENT,
       "CellLinked": A.CELL_LINKED, "CLVD": A.CLVD}


def lattice_case(dims, alg):
    kw = dict(neighbors=ALG[alg["name"]])
    if "skin" in alg:
        kw.update(skin=alg["skin"], max_neighbors=alg["nMaxNeighbors"])
    if "cellEdge" in alg:
        kw.update(cell_edge=alg["cellEdge"])
    if "dtNeighborRecompute" in alg:
        kw.update(dt_recompute=alg["dtNeighborRecompute"])
    m = model(dims, **kw)
    s = sim(3 ** dims, [[0.5, 1.5]] * dims)
    return m, s


@pytest.mark.parametrize("dims", [1, 2, 3])
def test_lattice_neighbor_counts_all_algorithms(dims):
    g = golden("ref_neighbors_lattice.json")
    for alg in g["algorithms"]:
        m, s = lattice_case(dims, alg)
        o = O.Oracle(m, s)
        for a, v in enumerate(lattice(dims)):
            o.set(a, v)
        assert o.init() == 0
        assert o.count_neighbors(g["radius"]).tolist() == g["expected"][str(dims)], alg["name"]


# ------------------------------------------------- CL / Verlet vs brute force
def brute_sets(pos, radius, box=None):
    X = np.stack(pos, 1)
    d = X[:, None, :] - X[None, :, :]
    if box is not None:
        L = box[:, 1] - box[:, 0]
        d = np.where(d > 0.5 * L, d - L, np.where(d < -0.5 * L, d + L, d))
    r = np.sqrt((d ** 2).sum(-1))
    np.fill_diagonal(r, np.inf)
    return [sorted(np.nonzero(r[i] < radius)[0].tolist()) for i in range(len(X))]


@pytest.mark.parametrize("periodic,dims", [(1, 2), (0, 2), (0, 3), (1, 3)])
def test_structures_equal_brute_force(periodic, dims):
    box = np.array([[0.0, 20.0]] * dims)
    n = 300 if dims == 2 else 600
    pos = uniform_positions(n, box, 11 + dims)
    want = brute_sets(pos, 2.0, box if periodic else None)
    for nb, kw in [(A.CELL_LINKED, dict(cell_edge=2.0)), (A.FULL, {}),
                   (A.VERLET_DISPLACEMENT, dict(skin=2.0, max_neighbors=64)),
                   (A.CLVD, dict(skin=2.0, max_neighbors=64, cell_edge=2.5))]:
        o = O.Oracle(model(dims, neighbors=nb, periodic=periodic, **kw), sim(n, box))
        for a in range(dims):
            o.set(a, pos[a])
        assert o.init() == 0
        assert o.neighbor_sets(2.0) == want, nb


# --------------------------------------------------- Aggregation golden run
def agg_oracle(case, g):
    mp = g["model"]
    m = model(2, integrator=A.EM, neighbors=A.FULL, law=A.LAW_AGGREGATION, dt=mp["dt"], lagged=1, noise=mp["D"])
    s = sim(len(case["x0"]), case["simBox"], [mp["rRep"], mp["fRep"], mp["rAtr"], mp["fAtr"], 1.0])
    o = O.Oracle(m, s)
    o.set(A.F_X, case["x0"])
    o.set(A.F_Y, case["y0"])
    assert o.init() == 0
    return o


def test_aggregation_two_agents_lag_golden():
    g = golden("ref_aggregation_lag.json")
    c = g["two_agents"]
    o = agg_oracle(c, g)
    for k in range(4):
        o.step(1)
        assert np.allclose(o.get(A.F_X), c["x"][k], atol=1e-12, rtol=0)
        assert np.allclose(o.get(A.F_FX), c["fx"][k], atol=1e-12, rtol=0)
        assert abs(o.t - c["t"][k]) < 1e-12
    o.step(1)
    assert abs(o.get(A.F_X)[0] - c["x_step5_agent0"]) < 1e-12
    assert abs(o.get(A.F_FX)[1] - c["fx_step5_agent1"]) < 1e-12


def test_aggregation_wall_golden():
    g = golden("ref_aggregation_lag.json")
    c = g["wall"]
    o = agg_oracle(c, g)
    for k in range(len(c["y"])):
        o.step(1)
        assert abs(o.get(A.F_Y)[0] - c["y"][k]) < 1e-12
        assert o.get(A.F_FY)[0] == c["fy"][k]
        assert o.get(A.F_FX)[0] == c["fx"][k]


# ---------------------------------------------------------- add/remove KAT
def test_add_remove_kat():
    g = golden("ref_add_remove.json")
    N = g["N"]
    m = model(1, neighbors=A.FULL, rule=A.RULE_PARITY)
    o = O.Oracle(m, sim(N, [[0.0, 1.0]], n_max=N + g["preallocateAgents"]))
    o.set(A.F_X, np.zeros(N))
    assert o.init() == 0
    assert o.step(1) == 0
    assert o.n == g["after_update"]["N"]
    assert o.id_max == g["after_update"]["idMax_"]
    lf = set(o.get(A.F_LF).tolist())
    assert set(g["after_update"]["lf_contains"]) <= lf
    ids = o.get(A.F_ID)
    assert len(set(ids.tolist())) == N                     # ids unique
    assert set(ids.tolist()) == set(range(2, 101, 2)) | set(range(101, 151))


def test_canonical_compaction_order():
    """Hand-built case: holes fill from the tail in ascending-hole order
    (update.jl:15-119 with the single-thread ascending hole list)."""
    m = model(1, neighbors=A.FULL, rule=A.RULE_PARITY)
    N = 6
    o = O.Oracle(m, sim(N, [[0.0, 1.0]], n_max=9))
    o.set(A.F_X, np.arange(N, dtype=float) * 10)
    o.init()
    o.step(1)
    # ids 1,3,5 removed (slots 0,2,4); ids 2,4,6 add daughters 7,8,9 at slots 6,7,8
    # survivors from the tail: slot 8 (id 9) -> hole 0, slot 7 (id 8) -> hole 2, slot 6 (id 7) -> hole 4
    assert o.get(A.F_ID).tolist() == [9, 2, 8, 4, 7, 6]
    assert o.get(A.F_LF).tolist() == [6, 0, 4, 0, 2, 0]
    # parents moved by x += 1, daughters keep the parent's old x
    assert o.get(A.F_X).tolist() == [50.0, 11.0, 30.0, 31.0, 10.0, 51.0]


# ------------------------------------------------------ integrator closed forms
@pytest.mark.parametrize("integ,factor", [
    (A.EULER, 1 - 0.05),
    (A.HEUN, 1 - 0.05 + 0.05 ** 2 / 2),
    (A.RK4, 1 - 0.05 + 0.05 ** 2 / 2 - 0.05 ** 3 / 6 + 0.05 ** 4 / 24),
])
def test_linear_ode_amplification_and_kat(integ, factor):
    g = golden("ref_ode_convergence.json")
    m = model(3, integrator=integ, neighbors=A.FULL, dt=g["dt"], drift=g["rate"])
    o = O.Oracle(m, sim(g["N"], g["simBox"]))
    for a in range(3):
        o.set(a, np.full(g["N"], g["x0"]))
    o.init()
    o.step(1)
    assert np.allclose(o.get(A.F_X), factor, rtol=0, atol=1e-15)
    for k in range(2, g["steps"] + 1):
        o.step(1)
        assert np.all(np.abs(o.get(A.F_X) - np.exp(-g["rate"] * o.t)) < g["tol"])
    assert np.allclose(o.get(A.F_X), factor ** g["steps"], rtol=1e-12)


@pytest.mark.parametrize("integ", [A.EM, A.EULER_HEUN])
def test_sde_variance_kat(integ):
    """test/testCommunity.jl:469-512 runs the same check for EM and EulerHeun."""
    g = golden("ref_sde_variance.json")
    m = model(3, integrator=integ, neighbors=A.FULL, dt=g["dt"], noise=g["g"], seed=5)
    o = O.Oracle(m, sim(g["N"], g["simBox"]))
    for a in range(3):
        o.set(a, np.zeros(g["N"]))
    o.init()
    for k in range(g["steps"]):
        o.step(1)
        x = o.get(A.F_X)
        assert abs(o.t - x.var()) < g["tol"]
    # tighter: Var[x(t)] = g^2 t within 5 sigma of the sample-variance error
    t = o.t
    for a in range(3):
        v = o.get(a).var()
        assert abs(v - t) < 5 * t * np.sqrt(2.0 / g["N"])


def test_euler_heun_deterministic_part_is_second_order():
    """Without noise the EulerHeun step of dx = -a x is the trapezoidal
    predictor-corrector: x1 = x0 (1 - a dt + (a dt)^2 / 2)."""
    a, dt, x0 = 0.7, 0.1, 1.5
    m = model(1, integrator=A.EULER_HEUN, neighbors=A.FULL, dt=dt, drift=a, seed=3)
    o = O.Oracle(m, sim(4, [[-10.0, 10.0]]))
    o.set(A.F_X, np.full(4, x0))
    o.init()
    o.step(1)
    want = x0 * (1 - a * dt + (a * dt) ** 2 / 2)
    assert np.allclose(o.get(A.F_X), want, rtol=1e-15, atol=0)


def test_euler_heun_corrector_draws_fresh_noise():
    """Reference defect D8 (src/integrators.jl:798-807): the corrector uses a
    new normal draw, not the predictor's.  For dx = -a x dt + g dW from x = 0
    one step has Var = g^2 dt (1 + (a dt)^2 / 4); reusing the predictor's
    draw would give g^2 dt (1 - a dt / 2)^2.  With a dt = 1: 1.25 vs 0.25."""
    n, dt, a, g = 20000, 0.5, 2.0, 1.0
    m = model(1, integrator=A.EULER_HEUN, neighbors=A.FULL, dt=dt, drift=a, noise=g, seed=11)
    o = O.Oracle(m, sim(n, [[-100.0, 100.0]]))
    o.set(A.F_X, np.zeros(n))
    o.init()
    o.step(1)
    v = o.get(A.F_X).var() / (g * g * dt)
    assert abs(v - 1.25) < 5 * 1.25 * np.sqrt(2.0 / n), v


def _one_step_free(integ, g0, slope, x0, dt, seed=13):
    m = model(1, integrator=integ, neighbors=A.FULL, dt=dt, noise=g0, seed=seed)
    o = O.Oracle(m, sim(len(x0), [[-100.0, 100.0]]))
    o.set_noise_slope(slope)
    o.set(A.F_X, x0)
    o.init()
    o.step(1)
    return o.get(A.F_X)


def test_state_dependent_noise_closed_forms():
    """g(x) = g0 + s x evaluated as the reference's g (src/integrators.jl:709-727,
    775-803): EM at u; EulerHeun's predictor at u and its corrector averaging
    g(u) and g(h) with a fresh normal.  The normals are read back from runs
    with g = 1: z1 = EM increment / sqrt(dt), z2 = EulerHeun increment /
    sqrt(dt).  Evaluating g at the new state, or one g for both EulerHeun
    stages, fails these."""
    dt, g0, s = 0.04, 0.3, 0.5
    x0 = np.linspace(-2.0, 2.0, 9)
    sq = np.sqrt(dt)
    z1 = (_one_step_free(A.EM, 1.0, 0.0, np.zeros(9), dt) - 0.0) / sq
    z2 = _one_step_free(A.EULER_HEUN, 1.0, 0.0, np.zeros(9), dt) / sq
    em = _one_step_free(A.EM, g0, s, x0, dt)
    assert np.allclose(em, x0 + (g0 + s * x0) * z1 * sq, rtol=1e-14, atol=1e-15)
    eh = _one_step_free(A.EULER_HEUN, g0, s, x0, dt)
    h = x0 + (g0 + s * x0) * z1 * sq
    assert np.allclose(eh, x0 + ((g0 + s * x0) + (g0 + s * h)) * z2 * sq / 2, rtol=1e-14, atol=1e-15)
    # slope 0: the constant noise
    assert np.allclose(_one_step_free(A.EM, g0, 0.0, x0, dt), x0 + g0 * z1 * sq, rtol=1e-14, atol=1e-15)


# --------------------------------------------------------- force invariants
def test_newton_third_law_and_two_body_closed_form():
    box = np.array([[0.0, 30.0]] * 3)
    n = 800
    pos = uniform_positions(n, box, 21)
    dt = 1e-3
    m = model(3, integrator=A.EULER, neighbors=A.CELL_LINKED, law=A.LAW_REP_ADH, dt=dt, cell_edge=2.4)
    o = O.Oracle(m, sim(n, box, [1.0, 1.0, 0.1, 2.4]))
    for a in range(3):
        o.set(a, pos[a])
    o.init()
    o.step(1)
    F = np.stack([(o.get(a) - pos[a]) / dt for a in range(3)], 1)
    assert np.all(np.abs(F.sum(0)) < 1e-9 * np.abs(F).sum())
    # two agents, linear law: d_{n+1} = d_n + 2 k (c - d_n) dt exactly in the continuum step
    k, c, d0 = 1.0, 2.0, 1.2
    m2 = model(1, integrator=A.EULER, neighbors=A.FULL, law=A.LAW_LINEAR, dt=0.01)
    o2 = O.Oracle(m2, sim(2, [[-10.0, 10.0]], [k, c]))
    o2.set(A.F_X, [-d0 / 2, d0 / 2])
    o2.init()
    d = d0
    for _ in range(50):
        o2.step(1)
        d = d + 2 * k * (c - d) * 0.01
        x = o2.get(A.F_X)
        assert abs((x[1] - x[0]) - d) < 1e-12
    # adhesion well: the force at the well centre has magnitude eps
    o3 = O.Oracle(model(1, integrator=A.EULER, neighbors=A.FULL, law=A.LAW_REP_ADH, dt=1.0),
                  sim(2, [[-10.0, 10.0]], [1.0, 1.0, 0.1, 2.4]))
    o3.set(A.F_X, [0.0, 2.2])
    o3.init()
    o3.step(1)
    x = o3.get(A.F_X)
    assert abs(x[0] - 0.1) < 1e-15 and abs(x[1] - (2.2 - 0.1)) < 1e-15


def test_verlet_list_valid_between_rebuilds():
    box = np.array([[0.0, 25.0]] * 3)
    n = 1200
    pos = uniform_positions(n, box, 31)
    m = model(3, integrator=A.HEUN, neighbors=A.VERLET_DISPLACEMENT, law=A.LAW_REP_ADH, dt=0.005,
              skin=2.7, max_neighbors=48, rebuild_disp=0.15)
    o = O.Oracle(m, sim(n, box, [1.0, 1.0, 0.1, 2.4]))
    for a in range(3):
        o.set(a, pos[a])
    o.init()
    for _ in range(40):
        o.step(1)
        cur = [o.get(a) for a in range(3)]
        within = brute_sets(cur, 2.4)
        lists = o.neighbor_sets(2.4)
        for i in range(0, n, 37):
            assert set(within[i]) <= set(lists[i])
    assert o.rebuilds >= 2


# ------------------------------------------------------------- dual numbers
def test_dual_gradient_matches_central_differences():
    box = np.array([[0.0, 8.0]] * 3)
    n = 64
    pos = uniform_positions(n, box, 41)
    m = model(3, integrator=A.EULER, neighbors=A.VERLET_DISPLACEMENT, law=A.LAW_REP_ADH, dt=0.01,
              skin=3.2, max_neighbors=64, rebuild_disp=0.2)
    o = O.Oracle(m, sim(n, box, [1.0, 1.0, 0.05, 2.4]))
    for a in range(3):
        o.set(a, pos[a])
    o.init()
    steps, saves = 20, [5, 10, 15, 20]
    target = np.concatenate([np.concatenate(uniform_positions(n, box, 50 + s)) for s in saves])
    theta = np.array([1.0, 1.0, 0.05])
    loss, g = o.traj_grad_dual(theta, steps, saves, target)
    assert abs(loss - o.traj_loss(theta, steps, saves, target)) < 1e-12 * abs(loss)
    for q in range(3):
        h = 1e-6 * max(abs(theta[q]), 1e-2)
        tp, tm = theta.copy(), theta.copy()
        tp[q] += h
        tm[q] -= h
        fd = (o.traj_loss(tp, steps, saves, target) - o.traj_loss(tm, steps, saves, target)) / (2 * h)
        assert abs(fd - g[q]) <= 1e-6 * max(abs(g[q]), 1e-8), (q, fd, g[q])
