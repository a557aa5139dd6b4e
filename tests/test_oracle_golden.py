"""The oracle against the reference's own regression fixtures.

demos/FiniteVolume/advection_2d.cpp run with --Tf 0.01
(tests/test_demo_finite_volume.py:55-72): unit square, a=(1,1), L4..10,
stencil radius 1, disc IC, Dirichlet 0, eps=2e-4, regularity 1, dt=2^-11, 21
iterations (the last with dt += Tf - t).  Cell lists must be identical; fields
agree within 1e-12 (the reference's CI compares these files at atol 1e-7,
tests/conftest.py:121-122).
"""
import os

import numpy as np
import pytest

from initial_conditions import disc_2d
from oracle_lib import OracleSim

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
FIELD_TOL = 1e-12


def load(name):
    return np.load(os.path.join(GOLDEN, name))


def demo_sim(s):
    sim = OracleSim(2, 4, 10, pred_s=s)
    sim.set_values(disc_2d(sim.cells()))
    sim.set_dirichlet(0.0)
    sim.adapt(2e-4, 1.0)
    return sim


def check_against(sim, fx):
    c = sim.cells()
    assert c.shape[0] == fx["level"].size
    assert np.array_equal(c[:, 0], fx["level"])
    assert np.array_equal(c[:, 1], fx["i"])
    assert np.array_equal(c[:, 2], fx["j"])
    v = sim.values()
    assert not np.isnan(v).any()
    err = np.max(np.abs(v - fx["u"]))
    assert err <= FIELD_TOL, err
    return err


@pytest.mark.parametrize("s", [0, 1])
def test_advection_2d_init_fixture(s):
    sim = demo_sim(s)
    err = check_against(sim, load(f"advection_2d_pred{s}_init.npz"))
    assert err == 0.0  # 0/1 initial data: every sum is exact


@pytest.mark.slow
@pytest.mark.parametrize("s", [0, 1])
def test_advection_2d_final_fixture(s):
    sim = demo_sim(s)
    dt, Tf, t = 0.5 * (1.0 / (1 << 10)), 0.01, 0.0
    a = np.array([1.0, 1.0])
    n = 0
    while t != Tf:
        sim.adapt(2e-4, 1.0)
        t += dt
        if t > Tf:
            dt += Tf - t
            t = Tf
        sim.upwind_step(a, dt)
        n += 1
    assert n == 21
    check_against(sim, load(f"advection_2d_pred{s}_t0.01.npz"))
