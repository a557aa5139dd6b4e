"""The single-core interval-algebra CPU baseline (oracle/cpu_step.cpp) that
bench.py times as the CPU reference: its adaptive steps must reproduce the
oracle bit for bit (cells with storage indices, values), so the baseline
computes the same thing as the GPU path it is compared with."""
import numpy as np
import pytest

from cpu_step_lib import CpuSim
from initial_conditions import box, disc_2d, seeded_2d, sphere_3d
from oracle_lib import OracleSim


@pytest.mark.parametrize("dim,lmin,lmax,ic,s", [(2, 2, 7, "box", 1), (2, 3, 8, "seeded", 1), (2, 3, 7, "disc", 0), (3, 2, 5, "sphere", 1)])
def test_cpu_baseline_matches_oracle(dim, lmin, lmax, ic, s):
    f = {"box": lambda c: box(c, 2), "seeded": lambda c: seeded_2d(c, 3), "disc": disc_2d, "sphere": sphere_3d}[ic]
    eps = 1e-4 if dim == 2 else 1e-3
    a = np.array([np.cos(np.pi / 5), np.sin(np.pi / 5), 0.3][:dim])
    dt = 0.4 / 3 * 2.0 ** -lmax
    o = OracleSim(dim, lmin, lmax, pred_s=s)
    o.set_poison(False)
    o.set_values(f(o.cells()))
    o.set_dirichlet(0.0)
    c = CpuSim(dim, lmin, lmax, o.cells(), pred_s=s)
    c.set_values(f(c.cells()))
    c.set_dirichlet(0.0)
    for step in range(4):
        assert c.adapt(eps, 2.0) == o.adapt(eps, 2.0), step
        assert np.array_equal(c.cells(), o.cells()), step
        c.upwind_step(a, dt)
        o.upwind_step(a, dt)
        assert np.array_equal(c.values().view(np.uint64), o.values().view(np.uint64)), step
