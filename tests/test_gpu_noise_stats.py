"""Statistics of the device noise (Philox4x32-10 + Box-Muller), the checks
the reference applies to its stochastic paths (test/testCommunity.jl:
321-356, 469-512): one Euler-Maruyama step of free agents gives increments
g xi sqrt(dt) with xi ~ N(0, 1) (KS test), independent across axes and
across steps, and the EulerHeun corrector draws fresh normals."""
import numpy as np
import pytest
from scipy import stats

import cbm._abi as A
from cbm.api import Context
from helpers import model, sim

pytestmark = pytest.mark.gpu


def free_agents(integ, n, g, dt, seed):
    m = model(3, integrator=integ, neighbors=A.FULL, dt=dt, noise=g, seed=seed)
    c = Context(m, [sim(n, [[-1e6, 1e6]] * 3)])
    for a in range(3):
        c.upload(a, np.zeros(n))
    c.neighbors()
    return c


@pytest.mark.parametrize("integ", [A.EM, A.EULER_HEUN])
def test_increments_are_standard_normal_and_independent(gpu_available, integ):
    n, g, dt = 200_000, 0.3, 0.04
    c = free_agents(integ, n, g, dt, seed=123)
    c.step(1)
    z1 = np.stack([c.download(a, n) for a in range(3)]) / (g * np.sqrt(dt))
    c.step(1)
    z2 = np.stack([c.download(a, n) for a in range(3)]) / (g * np.sqrt(dt)) - z1
    for z in (z1, z2):
        for a in range(3):
            assert stats.kstest(z[a], "norm").pvalue > 1e-4
        cc = np.corrcoef(z)
        assert np.max(np.abs(cc - np.eye(3))) < 0.02          # axes independent
    step_cc = np.corrcoef(z1[0], z2[0])[0, 1]
    assert abs(step_cc) < 0.02                                  # steps independent
    assert abs(z1.mean()) < 0.01 and abs(z1.var() - 1) < 0.02
