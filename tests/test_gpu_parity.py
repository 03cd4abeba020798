"""Parity of the HIP path (through the C ABI) with the CPU oracle on the same
seeded inputs, plus the reference known-answer tests run on the device.
Bar: bit-exact neighbor sets and compaction; deterministic states within
1e-12 relative per step and 1e-9 after 100 steps (relative to max(|x|,1));
the built-in laws are expected to be bit-identical (same IEEE expressions,
no contraction, same summation order)."""
import numpy as np
import pytest

import cbm._abi as A
import oracle as O
from cbm.api import Context
from helpers import golden, lattice, model, rel_err, sim, uniform_positions

pytestmark = pytest.mark.gpu

STEP_TOL = 1e-12
LONG_TOL = 1e-9


def pair(m, s, pos, extra=None):
    """Same model/simulation on the oracle and on the device."""
    o = O.Oracle(m, s)
    g = Context(m, [s])
    for a, v in enumerate(pos):
        o.set(a, v)
        g.upload(a, v)
    for f, v in (extra or {}).items():
        o.set(f, v)
        g.upload(f, v)
    assert o.init() == 0
    g.neighbors()
    return o, g


def states(obj, dims, n=None):
    if isinstance(obj, O.Oracle):
        return [obj.get(a) for a in range(dims)]
    return [obj.download(a, obj.n() if n is None else n) for a in range(dims)]


def assert_close(o, g, dims, tol, what=""):
    so, sg = states(o, dims), states(g, dims)
    for a in range(dims):
        e = rel_err(sg[a], so[a])
        assert e <= tol, f"{what} axis {a}: rel err {e}"


# ----------------------------------------------------------- reference KATs
@pytest.mark.parametrize("dims", [1, 2, 3])
def test_lattice_kat_all_algorithms(gpu_available, dims):
    from test_oracle import lattice_case
    g = golden("ref_neighbors_lattice.json")
    for alg in g["algorithms"]:
        m, s = lattice_case(dims, alg)
        ctx = Context(m, [s])
        for a, v in enumerate(lattice(dims)):
            ctx.upload(a, v)
        ctx.neighbors()
        got = ctx.count_neighbors(g["radius"]).tolist()
        assert got == g["expected"][str(dims)], alg["name"]


def test_aggregation_lag_golden_on_device(gpu_available):
    g = golden("ref_aggregation_lag.json")
    mp = g["model"]
    for case in ("two_agents", "wall"):
        c = g[case]
        m = model(2, integrator=A.EM, neighbors=A.FULL, law=A.LAW_AGGREGATION, dt=mp["dt"], lagged=1, noise=mp["D"])
        s = sim(len(c["x0"]), c["simBox"], [mp["rRep"], mp["fRep"], mp["rAtr"], mp["fAtr"], 1.0])
        ctx = Context(m, [s])
        ctx.upload(A.F_X, c["x0"])
        ctx.upload(A.F_Y, c["y0"])
        ctx.neighbors()
        if case == "two_agents":
            for k in range(4):
                ctx.step(1)
                assert np.allclose(ctx.download(A.F_X, 2), c["x"][k], atol=1e-12, rtol=0)
                assert np.allclose(ctx.download(A.F_FX, 2), c["fx"][k], atol=1e-12, rtol=0)
        else:
            for k in range(len(c["y"])):
                ctx.step(1)
                assert abs(ctx.download(A.F_Y, 1)[0] - c["y"][k]) < 1e-12
                assert ctx.download(A.F_FY, 1)[0] == c["fy"][k]


@pytest.mark.parametrize("integ", [A.EULER, A.HEUN, A.RK4])
def test_ode_convergence_kat_on_device(gpu_available, integ):
    g = golden("ref_ode_convergence.json")
    m = model(3, integrator=integ, neighbors=A.FULL, dt=g["dt"], drift=g["rate"])
    s = sim(g["N"], g["simBox"])
    pos = [np.full(g["N"], g["x0"])] * 3
    o, ctx = pair(m, s, pos)
    for k in range(1, g["steps"] + 1):
        ctx.step(1)
        o.step(1)
        x = ctx.download(A.F_X, g["N"])
        assert np.all(np.abs(x - np.exp(-g["rate"] * 0.1 * k)) < g["tol"])
        assert np.array_equal(x, o.get(A.F_X))


@pytest.mark.parametrize("integ", [A.EM, A.EULER_HEUN])
def test_sde_variance_kat_on_device(gpu_available, integ):
    g = golden("ref_sde_variance.json")
    m = model(3, integrator=integ, neighbors=A.FULL, dt=g["dt"], noise=g["g"], seed=5)
    s = sim(g["N"], g["simBox"])
    o, ctx = pair(m, s, [np.zeros(g["N"])] * 3)
    for k in range(1, g["steps"] + 1):
        ctx.step(1)
        x = ctx.download(A.F_X, g["N"])
        assert abs(0.1 * k - x.var()) < g["tol"]
    o.step(g["steps"])
    # same Philox stream as the oracle: Box-Muller transcendentals may differ by an ulp
    assert_close(o, ctx, 3, 1e-12, "SDE")


# --------------------------------------------------------------- C1 parity
def c1(n=1000, L=50.0, seed=1):
    box = [[0.0, L], [0.0, L]]
    m = model(2, integrator=A.EULER, neighbors=A.CELL_LINKED, law=A.LAW_LINEAR, dt=0.01, periodic=1, cell_edge=2.0)
    s = sim(n, box, [1.0, 2.0])
    return m, s, uniform_positions(n, box, seed)


@pytest.mark.parametrize("walk", [0, 1], ids=["skinlist", "sorted-grid"])
def test_c1_cell_linked_periodic_bit_exact(gpu_available, walk):
    m, s, pos = c1()
    m.walk = walk
    o, g = pair(m, s, pos)
    o.step(1)
    g.step(1)
    assert_close(o, g, 2, 0.0, "step 1")
    o.step(99)
    g.step(99)
    assert_close(o, g, 2, 0.0, "step 100")
    # walk 1 sorts the grid every step; walk 0 builds its skin list when the
    # agents have moved, which here (random overlaps relaxing) starts out
    # fast enough for a stretch of builds after every step
    assert (g.rebuilds() >= 100) if walk else (1 <= g.rebuilds() < 100)


def test_c1_neighbor_sets_equal(gpu_available):
    m, s, pos = c1()
    o, g = pair(m, s, pos)
    for _ in range(2):
        counts, lists = g.neighbor_lists(2.0, 64)
        want = o.neighbor_sets(2.0)
        for i in range(s.n):
            assert sorted(lists[i, : counts[i]].tolist()) == want[i]
        o.step(7)
        g.step(7)


# ---------------------------------------------------- C2-shaped 3-D parity
def c2(n=4096, seed=2, integ=A.HEUN, noise=0.0):
    L = (n * (4.0 / 3.0) * np.pi / 0.3) ** (1.0 / 3.0)
    box = [[0.0, L]] * 3
    m = model(3, integrator=integ, neighbors=A.VERLET_DISPLACEMENT, law=A.LAW_REP_ADH, dt=0.005,
              skin=2.7, max_neighbors=32, rebuild_disp=0.15, noise=noise, seed=7)
    s = sim(n, box, [1.0, 1.0, 0.1, 2.4])
    return m, s, uniform_positions(n, box, seed)


def test_c2_verlet_heun_parity(gpu_available):
    m, s, pos = c2()
    o, g = pair(m, s, pos)
    o.step(1)
    g.step(1)
    assert_close(o, g, 3, STEP_TOL, "step 1")
    o.step(99)
    g.step(99)
    assert_close(o, g, 3, LONG_TOL, "step 100")
    assert g.rebuilds() == o.rebuilds


def test_c2_verlet_lists_equal_as_sets(gpu_available):
    m, s, pos = c2(n=2048)
    o, g = pair(m, s, pos)
    for _ in range(3):
        counts, lists = g.neighbor_lists(0.0, 32)
        want = o.neighbor_sets(0.0)
        for i in range(s.n):
            assert sorted(lists[i, : counts[i]].tolist()) == want[i], i
        o.step(20)
        g.step(20)


def test_c2_em_variant_parity(gpu_available):
    m, s, pos = c2(n=2048, integ=A.EM, noise=np.sqrt(2 * 0.01))
    o, g = pair(m, s, pos)
    o.step(1)
    g.step(1)
    assert_close(o, g, 3, STEP_TOL, "EM step 1")
    o.step(49)
    g.step(49)
    assert_close(o, g, 3, LONG_TOL, "EM step 50")


def test_c2_euler_heun_parity(gpu_available):
    m, s, pos = c2(n=2048, integ=A.EULER_HEUN, noise=np.sqrt(2 * 0.01))
    o, g = pair(m, s, pos)
    o.step(1)
    g.step(1)
    assert_close(o, g, 3, STEP_TOL, "EulerHeun step 1")
    o.step(49)
    g.step(49)
    assert_close(o, g, 3, LONG_TOL, "EulerHeun step 50")
    assert g.rebuilds() == o.rebuilds


def test_rk4_cell_linked_3d_parity(gpu_available):
    n = 3000
    box = [[0.0, 30.0]] * 3
    m = model(3, integrator=A.RK4, neighbors=A.CELL_LINKED, law=A.LAW_REP_ADH, dt=0.01, cell_edge=2.4)
    s = sim(n, box, [1.0, 1.0, 0.1, 2.4])
    o, g = pair(m, s, uniform_positions(n, box, 9))
    o.step(1)
    g.step(1)
    assert_close(o, g, 3, 0.0, "RK4 step 1")
    o.step(49)
    g.step(49)
    assert_close(o, g, 3, LONG_TOL, "RK4 step 50")


def test_full_and_clvd_parity(gpu_available):
    n = 500
    box = [[0.0, 15.0]] * 2
    pos = uniform_positions(n, box, 13)
    for nb, kw in [(A.FULL, {}), (A.CLVD, dict(skin=2.6, max_neighbors=96, cell_edge=3.0)),
                   (A.VERLET_TIME, dict(skin=2.6, max_neighbors=96, dt_recompute=0.05))]:
        m = model(2, integrator=A.HEUN, neighbors=nb, law=A.LAW_LINEAR, dt=0.01, **kw)
        s = sim(n, box, [1.0, 2.0])
        o, g = pair(m, s, pos)
        o.step(30)
        g.step(30)
        assert_close(o, g, 2, LONG_TOL, f"nb={nb}")


def test_graph_replay_matches_eager(gpu_available):
    m, s, pos = c1(n=1000)
    _, g1 = pair(m, s, pos)
    _, g2 = pair(m, s, pos)
    g1.step(40)
    g2.step_graph(4, 10)
    for a in range(2):
        assert np.array_equal(g1.download(a, s.n), g2.download(a, s.n))
    m2, s2, pos2 = c2(n=2048)
    _, h1 = pair(m2, s2, pos2)
    _, h2 = pair(m2, s2, pos2)
    h1.step(30)
    h2.step_graph(2, 15)
    for a in range(3):
        assert np.array_equal(h1.download(a, s2.n), h2.download(a, s2.n))
    assert h1.rebuilds() == h2.rebuilds()


def test_list_overflow_is_reported(gpu_available):
    m, s, pos = c2(n=2048)
    m.max_neighbors = 2
    g = Context(m, [s])
    for a in range(3):
        g.upload(a, pos[a])
    g.neighbors()
    with pytest.raises(A.CBMError) as e:
        g.sync()
    assert e.value.code == A.ERR_LIST_OVERFLOW


def test_profile_categories(gpu_available):
    """cbm_profile: stage kinds 0..7, neighbor update 8, whole step 9, EulerHeun 10/11."""
    for integ, cats in ((A.HEUN, (2, 3)), (A.EULER_HEUN, (10, 11)), (A.EM, (1,))):
        m, s, pos = c2(n=2048, integ=integ, noise=0.1 if integ != A.HEUN else 0.0)
        g = Context(m, [s])
        for a in range(3):
            g.upload(a, pos[a])
        g.neighbors()
        ms, nl = g.profile(5)
        for k in cats:
            assert nl[k] == 5 and ms[k] > 0, (integ, k)
        assert nl[8] == 5 and nl[9] == 5 and ms[9] > 0
        assert sum(nl[k] for k in range(12) if k not in cats + (8, 9)) == 0
