"""cbm_count_neighbors_noalloc: the neighbor counts of cbm_count_neighbors
through the library's file-scope device array instead of per-call device
allocations.  Checked through the C ABI against cbm_count_neighbors on two
calls in a row, the second with a smaller population, on contexts of
different dimension and neighbor structure; populations beyond the array
are refused."""
import ctypes as C

import numpy as np
import pytest

import cbm._abi as A
from cbm.api import Context
from helpers import model, sim, uniform_positions

pytestmark = pytest.mark.gpu


def counts_noalloc(ctx, radius, n, sim_idx=0):
    out = np.full(max(n, 1), -7, dtype=np.int64)
    rc = A.lib().cbm_count_neighbors_noalloc(ctx.h, sim_idx, float(radius), out.ctypes.data)
    assert rc == A.OK, A.lib().cbm_last_error(ctx.h)
    return out[:n]


def counts_reference(ctx, radius, n, sim_idx=0):
    out = np.full(max(n, 1), -9, dtype=np.int64)
    rc = A.lib().cbm_count_neighbors(ctx.h, sim_idx, float(radius), out.ctypes.data)
    assert rc == A.OK, A.lib().cbm_last_error(ctx.h)
    return out[:n]


def context(dims, n, neighbors, seed, steps):
    box = [[0.0, 30.0]] * dims
    kw = dict(cell_edge=2.0) if neighbors == A.CELL_LINKED else dict(skin=2.5, max_neighbors=64, rebuild_disp=0.2)
    m = model(dims, integrator=A.EULER, neighbors=neighbors, law=A.LAW_LINEAR, dt=0.01, **kw)
    g = Context(m, [sim(n, box, [1.0, 1.0])])
    for c, v in enumerate(uniform_positions(n, box, seed)):
        g.upload(c, v)
    g.neighbors()
    g.step(steps)
    return g


def test_noalloc_counts_equal_the_allocating_entry_point(gpu_available):
    print("no-alloc counts: 3-D cell-linked N=6000, then 2-D Verlet N=700", flush=True)
    big = context(3, 6000, A.CELL_LINKED, 3, 5)
    small = context(2, 700, A.VERLET_DISPLACEMENT, 4, 5)
    a1 = counts_noalloc(big, 2.0, 6000)
    assert np.array_equal(a1, counts_reference(big, 2.0, 6000))
    assert a1.sum() > 0
    a2 = counts_noalloc(small, 1.5, 700)           # second call, smaller population
    assert np.array_equal(a2, counts_reference(small, 1.5, 700))
    l2 = counts_noalloc(small, 0.0, 700)            # Verlet: list lengths
    assert np.array_equal(l2, counts_reference(small, 0.0, 700))
    assert np.array_equal(counts_noalloc(big, 2.0, 6000), a1)   # back to the larger one
    big.close()
    small.close()


def test_noalloc_refuses_populations_beyond_its_array(gpu_available):
    n = (1 << 20) + 1
    m = model(1, integrator=A.EULER, neighbors=A.FULL, law=A.LAW_NONE, dt=0.01)
    g = Context(m, [sim(n, [[0.0, 10.0]])])
    out = np.zeros(n, dtype=np.int64)
    rc = A.lib().cbm_count_neighbors_noalloc(g.h, 0, 1.0, out.ctypes.data)
    assert rc == A.ERR_ARG
    assert b"exceed its scratch array" in A.lib().cbm_last_error(g.h)
    g.close()
