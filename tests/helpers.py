"""Shared builders for the parity tests: one model/simulation description is
fed to both the oracle (CPU restatement) and the HIP library."""
import json
import os

import numpy as np

import cbm._abi as A

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def golden(name):
    with open(os.path.join(GOLDEN, name)) as f:
        return json.load(f)


def model(dims, integrator=A.EULER, neighbors=A.FULL, law=A.LAW_NONE, dt=0.01, periodic=0, lagged=0,
          cell_edge=0.0, skin=0.0, max_neighbors=0, rebuild_disp=0.0, dt_recompute=0.0, drift=0.0,
          noise=0.0, seed=0, rule=A.RULE_NONE, rule_p=(0, 0, 0, 0)):
    m = A.ModelDesc()
    m.dims, m.integrator, m.neighbors, m.law = dims, integrator, neighbors, law
    m.rule = rule
    for q, v in enumerate(rule_p):
        m.rule_p[q] = v
    m.periodic, m.lagged = periodic, lagged
    m.max_neighbors = max_neighbors
    m.dt, m.cell_edge, m.skin, m.rebuild_disp, m.dt_recompute = dt, cell_edge, skin, rebuild_disp, dt_recompute
    m.drift, m.noise, m.seed = drift, noise, seed
    return m


def sim(n, box, p=(), n_max=None):
    s = A.SimDesc()
    s.n = n
    s.n_max = n if n_max is None else n_max
    box = np.asarray(box, dtype=np.float64).reshape(-1, 2)
    for a in range(box.shape[0]):
        s.box_lo[a], s.box_hi[a] = box[a]
    for q, v in enumerate(p):
        s.p[q] = v
    return s


def uniform_positions(n, box, seed):
    """Seeded i.i.d. uniform positions (numpy PCG64; the same arrays go to both sides)."""
    rng = np.random.default_rng(seed)
    box = np.asarray(box, dtype=np.float64).reshape(-1, 2)
    return [rng.uniform(lo, hi, n) for lo, hi in box]


def lattice(dims):
    """The 3^d integer lattice {0,1,2}^d of test/testCommunity.jl:396-401 (x fastest)."""
    v = np.zeros((3, 27))
    v[0] = np.tile([0, 1, 2], 9)
    v[1] = np.tile(np.repeat([0, 1, 2], 3), 3)
    v[2] = np.repeat([0, 1, 2], 9)
    return [v[a, : 3 ** dims].copy() for a in range(dims)]


def rel_err(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    scale = np.maximum(np.abs(b), 1.0)
    return float(np.max(np.abs(a - b) / scale)) if a.size else 0.0
