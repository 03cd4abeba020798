"""Pins of the oracle's built-in growth/division and death rules by what the
mathematics fixes, independently of the restatement itself.

Growth and division (the C3 colony rule; the reference's agentRule idiom of
src/AgentStructure/functionRule.jl:9-111 with the distributions of
src/random.jl:8-22): the volume grows by dvol/dt = growth * vol, integrated
by the model's integrator; when vol >= ratio * vbirth the agent divides into
itself and one new agent displaced -/+ split along a uniform random axis,
each with vol/2, vbirth = vol/2 and a fresh growth rate drawn uniformly on
[g (1 - s), g (1 + s)].  Death: each agent is removed with probability
rate * dt per step.

Checked here, on the oracle alone:
  - a division conserves volume exactly (parent = 2 x half) and sets vbirth;
  - the daughters' midpoint is the parent's position and their separation is
    2 split, both to a few ulp;
  - the axis is uniform: angle KS in 2-D, Archimedes' z and the azimuth KS in
    3-D, a fair sign in 1-D;
  - new growth rates are uniform on [g (1 - s), g (1 + s)] (KS), independent
    between the two daughters;
  - the trigger is vol >= ratio * vbirth, equality included, one ulp below not;
  - the Euler, Heun and RK4 volume growth follows the one-step amplification
    factors of the linear ODE dv/dt = g v;
  - the death fraction per step lies in a 5 sigma binomial band around
    rate * dt, and survival after n steps is (1 - rate dt)^n.
"""
import math

import numpy as np
import pytest
from scipy import stats

import cbm._abi as A
import oracle as O
from helpers import model, sim


def colony(dims, n, vol, vbirth, growth, rule_p, integrator=A.EULER, dt=0.1, seed=7, pos=None, n_max=None,
           rule=A.RULE_GROWTH_DIVISION):
    """An oracle community without forces (law none, no drift), so the
    positions change only by the rule."""
    m = model(dims, integrator=integrator, neighbors=A.FULL, law=A.LAW_NONE, dt=dt, seed=seed, rule=rule,
              rule_p=rule_p)
    box = [[0.0, 100.0]] * dims
    s = sim(n, box, n_max=2 * n if n_max is None else n_max)
    o = O.Oracle(m, s)
    rng = np.random.default_rng(seed)
    if pos is None:
        pos = [rng.uniform(10.0, 90.0, n) for _ in range(dims)]
    for a in range(dims):
        o.set(a, pos[a])
    if rule == A.RULE_GROWTH_DIVISION:
        o.set(A.F_VOL, np.broadcast_to(np.asarray(vol, dtype=float), (n,)))
        o.set(A.F_VBIRTH, np.broadcast_to(np.asarray(vbirth, dtype=float), (n,)))
        o.set(A.F_GROWTH, np.broadcast_to(np.asarray(growth, dtype=float), (n,)))
    o.set(A.F_LF, np.arange(n, dtype=np.int64))
    assert o.init() == 0
    return o, pos


def divide_once(dims, n, split=0.5, gm=0.02, gs=0.1, seed=7):
    """Every agent divides at step 1 (vol = ratio * vbirth exactly, growth 0 so
    the stage leaves vol unchanged).  Returns the oracle, the initial
    positions and the parent/daughter slots."""
    o, pos = colony(dims, n, 2.0, 1.0, 0.0, (gm, gs, split, 2.0), seed=seed)
    o.step(1)
    assert o.n == 2 * n and o.id_max == 2 * n
    par = np.arange(n)
    dau = n + np.arange(n)     # daughters appended in ascending parent slot
    lf = o.get(A.F_LF)
    assert np.array_equal(lf[par], lf[dau])
    return o, pos, par, dau


@pytest.mark.parametrize("dims", [1, 2, 3])
def test_division_conserves_volume_and_sets_vbirth(dims):
    o, _, par, dau = divide_once(dims, 2000)
    vol, vb = o.get(A.F_VOL), o.get(A.F_VBIRTH)
    assert np.all(vol[par] == 1.0) and np.all(vol[dau] == 1.0)
    assert np.all(vol[par] + vol[dau] == 2.0)
    assert np.all(vb[par] == 1.0) and np.all(vb[dau] == 1.0)
    # ids of the daughters: idMax + 1 + (additions before them), slot order
    assert np.array_equal(o.get(A.F_ID)[dau], 2000 + 1 + np.arange(2000))


@pytest.mark.parametrize("dims", [1, 2, 3])
def test_daughters_straddle_the_parent_at_two_split(dims):
    split = 0.5
    o, pos, par, dau = divide_once(dims, 20000, split=split)
    xp = np.stack([o.get(a)[par] for a in range(dims)])
    xd = np.stack([o.get(a)[dau] for a in range(dims)])
    x0 = np.stack(pos)
    mid = (xp + xd) / 2
    assert np.all(np.abs(mid - x0) <= 4 * np.spacing(np.abs(x0) + split))
    sep = np.sqrt(np.sum((xd - xp) ** 2, axis=0))
    assert np.max(np.abs(sep - 2 * split)) <= 1e-13


def axes(dims, n, seed=7):
    o, pos, par, dau = divide_once(dims, n, seed=seed)
    d = np.stack([o.get(a)[dau] - o.get(a)[par] for a in range(dims)])
    return d / np.sqrt(np.sum(d * d, axis=0))


def test_division_axis_is_a_fair_sign_in_1d():
    u = axes(1, 100_000)[0]
    assert set(np.unique(u).tolist()) == {-1.0, 1.0}
    k = int(np.sum(u > 0))
    assert abs(k - 50_000) < 5 * math.sqrt(100_000 * 0.25)


def test_division_axis_is_uniform_in_2d():
    u = axes(2, 100_000)
    th = np.mod(np.arctan2(u[1], u[0]), 2 * np.pi)
    assert stats.kstest(th, stats.uniform(0.0, 2 * np.pi).cdf).pvalue > 1e-3
    # a plausible mistake, an axis on the half circle only, fails the same test
    assert stats.kstest(np.mod(th, np.pi), stats.uniform(0.0, 2 * np.pi).cdf).pvalue < 1e-6


def test_division_axis_is_uniform_on_the_sphere():
    u = axes(3, 100_000)
    # Archimedes: z is uniform on [-1, 1]; the azimuth is uniform and independent of z
    assert stats.kstest(u[2], stats.uniform(-1.0, 2.0).cdf).pvalue > 1e-3
    ph = np.mod(np.arctan2(u[1], u[0]), 2 * np.pi)
    assert stats.kstest(ph, stats.uniform(0.0, 2 * np.pi).cdf).pvalue > 1e-3
    assert abs(np.corrcoef(u[2], ph)[0, 1]) < 0.02
    # a polar angle drawn uniformly instead would crowd the poles and fail
    bad_z = np.cos(np.pi * np.random.default_rng(0).uniform(0, 1, 100_000))
    assert stats.kstest(bad_z, stats.uniform(-1.0, 2.0).cdf).pvalue < 1e-6


def test_new_growth_rates_are_uniform_and_independent():
    gm, gs = 0.02, 0.1
    o, _, par, dau = divide_once(2, 100_000, gm=gm, gs=gs)
    g = o.get(A.F_GROWTH)
    g1, g2 = g[par], g[dau]
    lo, hi = gm * (1 - gs), gm * (1 + gs)
    assert np.all(g >= lo) and np.all(g <= hi)
    dist = stats.uniform(lo, hi - lo).cdf
    assert stats.kstest(g1, dist).pvalue > 1e-3
    assert stats.kstest(g2, dist).pvalue > 1e-3
    assert abs(np.corrcoef(g1, g2)[0, 1]) < 0.02
    assert not np.array_equal(g1, g2)


def test_division_fires_at_ratio_times_vbirth_inclusive():
    ratio, vb = 2.0, 1.3
    thr = ratio * vb
    vols = np.array([thr, np.nextafter(thr, 0.0), np.nextafter(thr, np.inf), 0.5 * thr, 3.0 * thr])
    fires = np.array([True, False, True, False, True])
    o, _ = colony(2, vols.size, vols, vb, 0.0, (0.02, 0.1, 0.5, ratio))
    o.step(1)
    assert o.n == vols.size + fires.sum()
    # the daughters' lf are their parents' slots, in ascending order
    assert o.get(A.F_LF)[vols.size:].tolist() == np.nonzero(fires)[0].tolist()
    vol = o.get(A.F_VOL)
    assert np.array_equal(vol[: vols.size][fires], vols[fires] / 2)
    assert np.array_equal(vol[: vols.size][~fires], vols[~fires])


@pytest.mark.parametrize("integ,amp", [
    (A.EULER, lambda h: 1 + h),
    (A.HEUN, lambda h: 1 + h + h * h / 2),
    (A.RK4, lambda h: 1 + h + h * h / 2 + h ** 3 / 6 + h ** 4 / 24),
])
def test_volume_growth_follows_the_linear_ode_amplification(integ, amp):
    """dv/dt = g v: n steps multiply v by the integrator's one-step factor
    R(g dt) n times (no division: vbirth far above)."""
    n, steps, dt = 500, 200, 0.05
    g = np.linspace(0.005, 0.05, n)
    v0 = np.linspace(0.5, 1.5, n)
    o, _ = colony(2, n, v0, 1e9, g, (0.02, 0.1, 0.5, 2.0), integrator=integ, dt=dt)
    o.step(steps)
    vol = o.get(A.F_VOL)
    expect = v0 * amp(g * dt) ** steps
    assert np.max(np.abs(vol - expect) / expect) < 1e-12
    # the exact exponential differs by the integrator's truncation error
    if integ == A.EULER:
        assert np.max(np.abs(vol - v0 * np.exp(g * dt * steps)) / vol) > 1e-6


def test_death_fraction_is_rate_times_dt():
    n, rate, dt = 400_000, 2.0, 0.05
    o, _ = colony(2, n, 1.0, 1.0, 0.0, (rate, 0, 0, 0), dt=dt, rule=A.RULE_DEATH, n_max=n)
    p = rate * dt
    o.step(1)
    removed = n - o.n
    assert abs(removed - n * p) < 5 * math.sqrt(n * p * (1 - p))
    # the removal is a pure function of (seed, step, id): survivors keep their ids
    ids = o.get(A.F_ID)
    assert ids.size == np.unique(ids).size and ids.max() <= n


def test_survival_after_n_steps():
    n, rate, dt, steps = 200_000, 2.0, 0.05, 10
    o, _ = colony(2, n, 1.0, 1.0, 0.0, (rate, 0, 0, 0), dt=dt, rule=A.RULE_DEATH, n_max=n)
    o.step(steps)
    q = (1 - rate * dt) ** steps
    assert abs(o.n - n * q) < 5 * math.sqrt(n * q * (1 - q))
    # and not the continuous-time survival exp(-rate t) (18 sigma away here)
    qe = math.exp(-rate * dt * steps)
    assert abs(n * qe - n * q) > 5 * math.sqrt(n * q * (1 - q))
