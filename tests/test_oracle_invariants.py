"""Closed-form and invariant checks that pin the oracle's numerics where the
golden fixtures leave them open (3D, prediction/projection exactness, flags)."""
import numpy as np
import pytest

from initial_conditions import centers
from oracle_lib import REFERENCE, OracleSim


def test_projection_prediction_polynomial_exactness_2d():
    # tests/test_projection_prediction_roundtrip.cpp:142-197: cell averages of a
    # separable polynomial of per-axis degree <= 2s are reproduced by prediction
    # and projection.  Here: a linear field is reproduced exactly at every ghost
    # that update_ghost_mr predicts or projects (inside the domain).
    sim = OracleSim(2, 2, 5, pred_s=1)
    ref = sim.cells(REFERENCE)
    c = centers(ref)
    u = 0.25 + 0.5 * c[:, 0] + 0.125 * c[:, 1]     # dyadic coefficients: exact
    st = np.full(sim.nb_cells(REFERENCE), np.nan)
    st[ref[:, 4]] = u
    cells = sim.cells()
    inside = st.copy()
    # keep only the leaves, everything else must be rebuilt by the ghost update
    st2 = np.full_like(st, np.nan)
    st2[cells[:, 4]] = st[cells[:, 4]]
    sim.set_storage(st2)
    sim.update_ghost_mr()
    out = sim.storage()
    # away from the boundary (where Dirichlet/corner values enter the stencils)
    interior = np.all((c[:, :2] > 0.25) & (c[:, :2] < 0.75), axis=1)
    idx = ref[interior, 4]
    assert idx.size > 100
    assert not np.isnan(out[idx]).any()
    assert np.array_equal(out[idx], inside[idx])


def test_upwind_uniform_closed_form():
    # on a uniform mesh with a=(1,1) the upwind step is the textbook
    # u - dt/dx*((u - u_left) + (u - u_down)) up to the flux rounding
    sim = OracleSim(2, 5, 5, pred_s=1)
    cells = sim.cells()
    rng = np.random.default_rng(1)
    u = rng.random(len(cells))
    sim.set_values(u)
    sim.set_dirichlet(0.0)
    dt = 0.5 / 32
    sim.upwind_step(np.array([1.0, 1.0]), dt)
    grid = np.zeros((34, 34))
    grid[1:33, 1:33] = u.reshape(32, 32)
    grid[0, 1:33] = -grid[1, 1:33]
    grid[1:33, 0] = -grid[1:33, 1]
    g = grid[1:33, 1:33]
    expect = g - dt * 32 * ((g - grid[1:33, 0:32]) + (g - grid[0:32, 1:33]))
    assert np.max(np.abs(sim.values().reshape(32, 32) - expect)) < 1e-14


@pytest.mark.parametrize("dim", [2, 3])
def test_adaptation_invariants(dim):
    # is_graduated (algorithm/graduation.hpp:214-243, tests/test_graduation.cpp:11-45)
    # and array().size()==nb_cells() consistency (tests/test_adapt.cpp:132-155)
    from initial_conditions import sphere_3d, disc_2d
    sim = OracleSim(dim, 2, 5 if dim == 2 else 4, pred_s=1)
    cells = sim.cells()
    sim.set_values(disc_2d(cells) if dim == 2 else sphere_3d(cells))
    sim.set_dirichlet(0.0)
    sim.adapt(1e-3, 2.0)
    cells = sim.cells()
    levels = np.unique(cells[:, 0])
    assert len(levels) > 1
    # graduation: a leaf at level L and a leaf at level <= L-2 never touch
    occ = {}
    for row in cells:
        occ[(row[0], row[1], row[2], row[3])] = True
    for row in cells:
        l, x, y, z = row[:4]
        for lb in range(2, 6):
            if l - lb < 0:
                break
            for dz in ((-1, 0, 1) if dim == 3 else (0,)):
                for dy in (-1, 0, 1):
                    for dx in (-1, 0, 1):
                        key = (l - lb, (x + dx) >> lb, (y + dy) >> lb, ((z + dz) >> lb) if dim == 3 else 0)
                        assert key not in occ
    # the leaves tile the unit box
    vol = np.sum(np.ldexp(1.0, -dim * cells[:, 0].astype(np.int64)))
    assert vol == 1.0


@pytest.mark.parametrize("lmin,lmax", [(5, 5), (3, 7)])
def test_hand_loop_equals_upwind_at_unit_velocity(lmin, lmax):
    """The README's hand loop (README.md:135-154) is the upwind scheme of
    stencil_field.hpp at a = (1, 1): F(1, ul, ur) = ul, so both give
    u - dt/dx ((u - u(i-1)) + (u - u(j-1))) up to rounding, on uniform and
    adapted meshes (the upwind oracle is pinned by the golden fixtures)."""
    from initial_conditions import seeded_2d
    a = OracleSim(2, lmin, lmax, pred_s=1)
    a.set_values(seeded_2d(a.cells(), 4))
    a.set_dirichlet(0.0)
    if lmin < lmax:
        a.adapt(1e-3, 2.0)
    b = OracleSim(2, lmin, lmax, pred_s=1, cells=a.cells())
    b.set_values(a.values())
    b.set_dirichlet(0.0)
    dt = 0.5 * 2.0 ** -lmax
    a.upwind_hand_loop(dt)
    b.upwind_step(np.array([1.0, 1.0]), dt)
    assert np.allclose(a.values(), b.values(), rtol=0, atol=1e-13)
