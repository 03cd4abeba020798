"""Division/death rules and the canonical compaction of update! on the
device against the oracle: N, idMax_, and the id/lf of every logical slot
bit-exact; positions within the state tolerance (division axes use device
cos/sin, which may differ from the host's by an ulp)."""
import numpy as np
import pytest

import cbm._abi as A
import oracle as O
from cbm.api import Context
from helpers import golden, model, rel_err, sim, uniform_positions

pytestmark = pytest.mark.gpu


def pair(m, s, fields):
    o = O.Oracle(m, s)
    g = Context(m, [s])
    for f, v in fields.items():
        o.set(f, v)
        g.upload(f, v)
    assert o.init() == 0
    g.neighbors()
    return o, g


def same_slots(o, g, dims, tol):
    n = o.n
    assert g.n() == n
    assert g.id_max() == o.id_max
    assert np.array_equal(g.download(A.F_ID, n), o.get(A.F_ID))
    assert np.array_equal(g.download(A.F_LF, n), o.get(A.F_LF))
    for a in range(dims):
        e = rel_err(g.download(a, n), o.get(a))
        assert e <= tol, (a, e)


def test_add_remove_kat_on_device(gpu_available):
    gk = golden("ref_add_remove.json")
    N = gk["N"]
    m = model(1, neighbors=A.FULL, rule=A.RULE_PARITY)
    s = sim(N, [[0.0, 1.0]], n_max=N + gk["preallocateAgents"])
    o, g = pair(m, s, {A.F_X: np.arange(N, dtype=float) * 0.01})
    o.step(1)
    g.step(1)
    assert g.n() == gk["after_update"]["N"]
    assert g.id_max() == gk["after_update"]["idMax_"]
    assert set(gk["after_update"]["lf_contains"]) <= set(g.download(A.F_LF, N).tolist())
    same_slots(o, g, 1, 0.0)


def test_canonical_compaction_order_on_device(gpu_available):
    m = model(1, neighbors=A.FULL, rule=A.RULE_PARITY)
    o, g = pair(m, sim(6, [[0.0, 1.0]], n_max=9), {A.F_X: np.arange(6, dtype=float) * 10})
    g.step(1)
    assert g.download(A.F_ID, 6).tolist() == [9, 2, 8, 4, 7, 6]
    assert g.download(A.F_LF, 6).tolist() == [6, 0, 4, 0, 2, 0]
    assert g.download(A.F_X, 6).tolist() == [50.0, 11.0, 30.0, 31.0, 10.0, 51.0]


def test_capacity_overflow_is_reported(gpu_available):
    m = model(1, neighbors=A.FULL, rule=A.RULE_PARITY)
    g = Context(m, [sim(10, [[0.0, 1.0]], n_max=12)])
    g.upload(A.F_X, np.zeros(10))
    g.neighbors()
    g.step(1)
    with pytest.raises(A.CBMError) as e:
        g.sync()
    assert e.value.code == A.ERR_CAPACITY


@pytest.mark.parametrize("nb", [A.CELL_LINKED, A.VERLET_DISPLACEMENT, A.FULL])
def test_death_compaction_matches_oracle(gpu_available, nb):
    n = 3000
    box = [[0.0, 60.0], [0.0, 60.0]]
    kw = dict(cell_edge=2.0) if nb == A.CELL_LINKED else (dict(skin=2.6, max_neighbors=48, rebuild_disp=0.3) if nb == A.VERLET_DISPLACEMENT else {})
    m = model(2, integrator=A.EULER, neighbors=nb, law=A.LAW_LINEAR, dt=0.01, rule=A.RULE_DEATH,
              rule_p=(3.0, 0, 0, 0), seed=11, **kw)
    pos = uniform_positions(n, box, 3)
    o, g = pair(m, sim(n, box, [1.0, 2.0]), {A.F_X: pos[0], A.F_Y: pos[1]})
    for k in range(8):
        o.step(5)
        g.step(5)
        same_slots(o, g, 2, 0.0)
    assert o.n < n * 0.5


@pytest.mark.parametrize("dims,nb", [(2, A.CELL_LINKED), (3, A.VERLET_DISPLACEMENT)])
def test_growth_division_matches_oracle(gpu_available, dims, nb):
    n0, cap = 200, 6000
    L = 80.0
    box = [[0.0, L]] * dims
    rng = np.random.default_rng(8)
    pos = [rng.uniform(L / 2 - 6, L / 2 + 6, n0) for _ in range(dims)]
    kw = dict(cell_edge=2.0) if nb == A.CELL_LINKED else dict(skin=2.6, max_neighbors=64, rebuild_disp=0.3)
    m = model(dims, integrator=A.EULER, neighbors=nb, law=A.LAW_LINEAR, dt=0.05,
              rule=A.RULE_GROWTH_DIVISION, rule_p=(0.5, 0.1, 0.5, 2.0), seed=4, **kw)
    g0 = rng.uniform(0.45, 0.55, n0)
    fields = {A.F_X + a: pos[a] for a in range(dims)}
    fields.update({A.F_VOL: np.ones(n0), A.F_VBIRTH: np.ones(n0), A.F_GROWTH: g0})
    o, g = pair(m, sim(n0, box, [1.0, 2.0], n_max=cap), fields)
    for k in range(8):
        o.step(10)
        g.step(10)
        n = o.n
        same_slots(o, g, dims, 1e-9)
        assert np.array_equal(g.download(A.F_VOL, n), o.get(A.F_VOL))
        assert np.array_equal(g.download(A.F_GROWTH, n), o.get(A.F_GROWTH))
    assert o.n > 3 * n0


def test_rule_scan_longer_than_bucket_table(gpu_available):
    """Capacity far above the cell count: the per-slot rule scans need more
    scan tiles than the bucket scan (regression: the scan status array was
    sized for the bucket scan only)."""
    n0, cap = 100, 200_000
    box = [[0.0, 40.0], [0.0, 40.0]]
    rng = np.random.default_rng(2)
    pos = [rng.uniform(15, 25, n0), rng.uniform(15, 25, n0)]
    m = model(2, integrator=A.EULER, neighbors=A.CELL_LINKED, law=A.LAW_LINEAR, dt=0.05, cell_edge=2.0,
              rule=A.RULE_GROWTH_DIVISION, rule_p=(0.5, 0.1, 0.5, 2.0), seed=6)
    fields = {A.F_X: pos[0], A.F_Y: pos[1], A.F_VOL: np.ones(n0), A.F_VBIRTH: np.ones(n0),
              A.F_GROWTH: rng.uniform(0.45, 0.55, n0)}
    o, g = pair(m, sim(n0, box, [1.0, 2.0], n_max=cap), fields)
    o.step(40)
    g.step(40)
    same_slots(o, g, 2, 1e-9)
    assert o.n > n0


@pytest.mark.parametrize("nb", [A.CELL_LINKED, A.VERLET_DISPLACEMENT])
def test_periodic_division_across_the_boundary(gpu_available, nb):
    """Agents dividing next to the faces of a periodic box: a new agent placed
    past a face is wrapped into the box, as the oracle's end-of-step wrap does."""
    n0, L = 300, 30.0
    rng = np.random.default_rng(12)
    pos = [rng.uniform(0.0, L, n0), rng.uniform(0.0, L, n0)]
    pos[0][:100] = rng.uniform(0.0, 0.4, 100)        # against the x faces
    pos[1][100:200] = rng.uniform(L - 0.4, L, 100)   # against the y faces
    kw = dict(cell_edge=2.0) if nb == A.CELL_LINKED else dict(skin=2.6, max_neighbors=256, rebuild_disp=0.3)
    m = model(2, integrator=A.EULER, neighbors=nb, law=A.LAW_LINEAR, dt=0.05, periodic=1,
              rule=A.RULE_GROWTH_DIVISION, rule_p=(0.5, 0.1, 0.5, 2.0), seed=21, **kw)
    fields = {A.F_X: pos[0], A.F_Y: pos[1], A.F_VOL: rng.uniform(1.9, 2.0, n0), A.F_VBIRTH: np.ones(n0),
              A.F_GROWTH: rng.uniform(0.45, 0.55, n0)}
    o, g = pair(m, sim(n0, [[0.0, L], [0.0, L]], [1.0, 2.0], n_max=4000), fields)
    for _ in range(4):
        o.step(5)
        g.step(5)
        n = o.n
        same_slots(o, g, 2, 1e-9)
        for a in range(2):
            x = g.download(a, n)
            assert x.min() >= 0.0 and x.max() < L
    assert o.n > 1.5 * n0
