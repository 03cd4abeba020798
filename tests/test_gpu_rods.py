"""The 2-D rod model (src/models/rods.jl:21-121) on the device against the
oracle: Full, CellLinked and Verlet walks under Euler, Heun and RK4, with
mixed lengths and masses and dense contacts; the per-rod outputs fx, fy, W
and pressure; graph replays; the negative-overlap error; the Python mirror
(CBMModels.rods2D) end to end.

Tolerance: the contact geometry calls sin, cos and pow, whose device and
host implementations may differ in the last bit, and divides by
sin(theta_j - theta_i), which near-parallel rods make as small as 1e-6 (the
reference's parallel threshold), so last-bit differences can grow a
million-fold in a contact point.  The bound therefore follows the problem's
own conditioning: the oracle is run a second time with every angle moved by
one ulp, and the device may differ from the oracle by at most 100 times that
spread (plus 1e-13 of the field's scale) in every field."""
import numpy as np
import pytest

import cbm._abi as A
import oracle as O
from cbm.api import ABM, CBMIntegrators, CBMModels, CBMNeighbors, Community, Context, evolve_
from helpers import model, sim

pytestmark = pytest.mark.gpu

P = [60.0, 3.0, 4.0, 0.4, 0.0, 0.8, 1.5]     # kn, gamma_n, gamma_t, mu_cc, mu_cw, beta, beta_w
STATE = (A.F_X, A.F_Y, A.F_VX, A.F_VY, A.F_THETA, A.F_OMEGA)
OUTS = (A.F_FX, A.F_FY, A.F_TORQUE, A.F_PRESSURE)


def rods(n, L, seed):
    """Rods of diameter 1 and lengths 1..3 packed at ~0.5 area fraction."""
    rng = np.random.default_rng(seed)
    return {
        A.F_X: rng.uniform(1.0, L - 1.0, n), A.F_Y: rng.uniform(1.0, L - 1.0, n),
        A.F_VX: rng.normal(0, 0.3, n), A.F_VY: rng.normal(0, 0.3, n),
        A.F_THETA: rng.uniform(-np.pi, np.pi, n), A.F_OMEGA: rng.normal(0, 0.5, n),
        A.F_LENGTH: rng.uniform(1.0, 3.0, n), A.F_MASS: rng.uniform(0.5, 2.0, n),
    }


def both(m, s, ag):
    o = O.Oracle(m, s)
    g = Context(m, [s])
    for f, v in ag.items():
        o.set(f, v)
        g.upload(f, v)
    assert o.init() == 0
    g.neighbors()
    return o, g


def nudged(m, s, ag, steps):
    """The oracle after `steps` steps with every angle one ulp larger."""
    o = O.Oracle(m, s)
    for f, v in ag.items():
        o.set(f, np.nextafter(v, np.inf) if f == A.F_THETA else v)
    assert o.init() == 0
    o.step(steps)
    return o


def assert_close(o, g, o2, what, fields=STATE + OUTS):
    n = o.n
    for f in fields:
        ref = o.get(f)
        e_dev = float(np.max(np.abs(g.download(f, n) - ref)))
        e_nat = float(np.max(np.abs(o2.get(f) - ref)))
        scale = max(1.0, float(np.max(np.abs(ref))))
        assert e_dev <= 100.0 * e_nat + 1e-13 * scale, (what, f, e_dev, e_nat)


def box_for(n):
    return float(np.sqrt(n * 2.0 / 0.5))      # mean rod area ~2 (d = 1, l ~ 2)


@pytest.mark.parametrize("integ", [A.EULER, A.HEUN, A.RK4])
def test_rods_full_parity(gpu_available, integ):
    n = 300
    L = box_for(n)
    m = model(2, integrator=integ, neighbors=A.FULL, law=A.LAW_RODS, dt=0.002)
    s, ag = sim(n, [[0, L]] * 2, P), rods(n, L, 11)
    o, g = both(m, s, ag)
    o.step(1)
    g.step(1)
    assert_close(o, g, nudged(m, s, ag, 1), "step 1")
    assert np.count_nonzero(o.get(A.F_PRESSURE)) > n // 4      # dense contacts
    o.step(39)
    g.step(39)
    assert_close(o, g, nudged(m, s, ag, 40), "step 40")


@pytest.mark.parametrize("integ", [A.EULER, A.HEUN, A.RK4])
def test_rods_cell_linked_parity(gpu_available, integ):
    n = 2000
    L = box_for(n)
    m = model(2, integrator=integ, neighbors=A.CELL_LINKED, law=A.LAW_RODS, dt=0.002, cell_edge=4.0)
    s, ag = sim(n, [[0, L]] * 2, P), rods(n, L, 12)
    o, g = both(m, s, ag)
    o.step(1)
    g.step(1)
    assert_close(o, g, nudged(m, s, ag, 1), "step 1")
    o.step(59)
    g.step(59)
    assert_close(o, g, nudged(m, s, ag, 60), "step 60")


def test_rods_verlet_parity(gpu_available):
    n = 2000
    L = box_for(n)
    m = model(2, integrator=A.HEUN, neighbors=A.VERLET_DISPLACEMENT, law=A.LAW_RODS, dt=0.002, skin=4.5,
              max_neighbors=64, rebuild_disp=0.05)
    s, ag = sim(n, [[0, L]] * 2, P), rods(n, L, 13)
    o, g = both(m, s, ag)
    o.step(60)
    g.step(60)
    assert_close(o, g, nudged(m, s, ag, 60), "step 60")
    assert g.rebuilds() == o.rebuilds and o.rebuilds > 1


def test_rods_graph_replays_equal_eager_steps(gpu_available):
    n = 1000
    L = box_for(n)
    m = model(2, integrator=A.HEUN, neighbors=A.CELL_LINKED, law=A.LAW_RODS, dt=0.002, cell_edge=4.0)
    s = sim(n, [[0, L]] * 2, P)
    ag = rods(n, L, 14)
    o, g = both(m, s, ag)
    h = Context(m, [s])
    for f, v in ag.items():
        h.upload(f, v)
    h.neighbors()
    g.step(20)
    h.step_graph(4, 5)
    for f in STATE + OUTS:
        assert np.array_equal(g.download(f, n), h.download(f, n)), f
    o.step(20)
    assert_close(o, g, nudged(m, s, ag, 20), "step 20")


def test_rods_negative_overlap_is_an_error(gpu_available):
    # a thin rod 0.8 from a thick one: d_i < r_ij < (d_i + d_j)/2
    m = model(2, integrator=A.EULER, neighbors=A.FULL, law=A.LAW_RODS, dt=0.001)
    g = Context(m, [sim(2, [[0, 10]] * 2, P)])
    g.upload(A.F_X, np.array([5.0, 5.0]))
    g.upload(A.F_Y, np.array([5.0, 5.8]))
    g.upload(A.F_DIAMETER, np.array([0.5, 1.5]))
    g.neighbors()
    g.step(1)            # steps run asynchronously: the error comes with the sync
    with pytest.raises(Exception, match="overlap"):
        g.sync()


@pytest.mark.parametrize("kw", [{"dims": 3}, {"periodic": 1}, {"integrator": A.EM}, {"lagged": 1}])
def test_rods_refusals(gpu_available, kw):
    kw = dict(kw)
    dims = kw.pop("dims", 2)
    m = model(dims, law=A.LAW_RODS, neighbors=A.FULL, **kw)
    with pytest.raises(Exception):
        Context(m, [sim(2, [[0, 10]] * dims, P)])


def test_rods_python_mirror(gpu_available):
    n = 400
    L = box_for(n)
    ag = rods(n, L, 15)
    abm = ABM(2, agentODE=CBMModels.rods2D(*P), neighborsAlg=CBMNeighbors.CellLinked(cellEdge=4.0),
              agentAlg=CBMIntegrators.Heun())
    com = Community(abm, N=n, dt=0.002, simBox=[[0.0, L], [0.0, L]], x=ag[A.F_X], y=ag[A.F_Y], vx=ag[A.F_VX],
                    vy=ag[A.F_VY], theta=ag[A.F_THETA], **{"ω": ag[A.F_OMEGA]}, l=ag[A.F_LENGTH], m=ag[A.F_MASS])
    evolve_(com, steps=25)
    m = model(2, integrator=A.HEUN, neighbors=A.CELL_LINKED, law=A.LAW_RODS, dt=0.002, cell_edge=4.0)
    o = O.Oracle(m, sim(n, [[0.0, L]] * 2, P))
    for f, v in ag.items():
        o.set(f, v)
    assert o.init() == 0
    o.step(25)
    o2 = nudged(m, sim(n, [[0.0, L]] * 2, P), ag, 25)
    for name, f in (("x", A.F_X), ("theta", A.F_THETA), ("pressure", A.F_PRESSURE), ("W", A.F_TORQUE)):
        e_dev = float(np.max(np.abs(com[name] - o.get(f))))
        e_nat = float(np.max(np.abs(o2.get(f) - o.get(f))))
        assert e_dev <= 100.0 * e_nat + 1e-13 * max(1.0, float(np.max(np.abs(o.get(f))))), (name, e_dev, e_nat)


G0 = [0.5, 0.3, 3.0, 0.0]      # growth, sigma_lTarget, lMax, alpha = 0: division times independent of contacts
G1 = [0.5, 0.3, 3.0, 0.8]


def growth_pair(g, neigh, integ, n0=120, steps_seed=21):
    L = 60.0
    kw = dict(cell_edge=7.0) if neigh == A.CELL_LINKED else {}
    m = model(2, integrator=integ, neighbors=neigh, law=A.LAW_RODS, dt=0.01, rule=A.RULE_RODS_GROWTH, rule_p=g,
              seed=13, **kw)
    s = sim(n0, [[0, L]] * 2, P, n_max=2000)
    rng = np.random.default_rng(steps_seed)
    ag = {A.F_X: rng.uniform(15, 45, n0), A.F_Y: rng.uniform(15, 45, n0), A.F_THETA: rng.uniform(-3, 3, n0),
          A.F_LENGTH: rng.uniform(1.0, 2.9, n0), A.F_LTARGET: rng.uniform(2.7, 3.3, n0)}
    return m, s, ag


@pytest.mark.parametrize("neigh,integ", [(A.FULL, A.EULER), (A.CELL_LINKED, A.HEUN), (A.CELL_LINKED, A.RK4)])
def test_rods_growth_colony(gpu_available, neigh, integ):
    # alpha = 0: the lengths, the division steps, ids and lTargets are exact;
    # positions and angles within the conditioning-scaled bound
    m, s, ag = growth_pair(G0, neigh, integ)
    o, g = both(m, s, ag)
    for chunk in (50, 150):
        o.step(chunk)
        g.step(chunk)
        n = o.n
        assert g.n() == n and o.error() == 0
        for f in (A.F_ID, A.F_LENGTH, A.F_LTARGET):
            assert np.array_equal(g.download(f, n), o.get(f)), f
    assert o.n > s.n + 40                       # about half of the rods divided, some daughters again
    assert_close(o, g, nudged(m, s, ag, 200), "step 200", fields=(A.F_X, A.F_Y, A.F_THETA, A.F_OMEGA, A.F_VX))


def test_rods_growth_with_pressure(gpu_available):
    # alpha > 0: contacts slow growth; the lengths too follow the tolerance
    m, s, ag = growth_pair(G1, A.CELL_LINKED, A.HEUN, steps_seed=22)
    o, g = both(m, s, ag)
    o.step(60)
    g.step(60)
    assert g.n() == o.n and np.array_equal(g.download(A.F_ID, o.n), o.get(A.F_ID))
    assert o.n > s.n
    assert_close(o, g, nudged(m, s, ag, 60), "step 60", fields=STATE + (A.F_LENGTH, A.F_PRESSURE))
