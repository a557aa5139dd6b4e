"""Known answers that pin the oracle's wide ghost layers and its WENO5
operator (meshes of one level, stencil radius 2 or 3).  Each check restates
a result of the reference's own tests or of the mathematics, not the
oracle's formulas:

- corner blocks of width 2 and 3: every corner ghost holds the field at the
  reflection of its centre about the domain corner
  (tests/test_bc_ghost_values.cpp:278-358, run at ghost widths 1..3, :740-760);
- far ghost layers with Dirichlet<1> on a field linear along the normal and
  constant on the face: every layer holds the field at the ghost centre
  (run_further, tests/test_bc_ghost_values.cpp:388-440: the B.C. is exact on
  the first layer, the polynomial extrapolation on the others);
- far ghost layers without a B.C.: Lagrange interpolation through the known
  values gives layer k = f(b - (k - 1) h) for a linear f (derived below);
- the WENO5 flux: exact interface values for linear and quadratic cell
  averages whatever the weights (each candidate of Jiang & Shu, eq. 2.8, is
  exact on quadratics), the ENO property at a jump, and fifth order on smooth
  data (the linear weights 0.1, 0.6, 0.3);
- the WENO5 operator: v . grad u exactly on linear fields at interior cells,
  and the reference's convergence test (tests/test_fv_operators.cpp:898-951):
  observed order above 4 on a sine wave at levels 4, 5, 6.
"""
import math

import numpy as np
import pytest

from initial_conditions import centers
from oracle_lib import CELLS, REFERENCE, OracleSim, weno5_flux
from test_oracle_pinning import cartesian_directions, corner_oracle


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("width", [2, 3])
def test_wide_corner_ghosts_reflect_the_field(dim, width):
    """run_corners at ghost widths 2 and 3 (max_stencil_size 2 * width) on a
    uniform level-3 mesh: each of the width^n ghosts of every corner block
    holds f at the reflection of its centre."""
    f = lambda q: 1.0 + 0.5 * q[:, 0] * q[:, 0] + sum((d + 1) * q[:, d] for d in range(dim))
    sim = OracleSim(dim, 3, 3, pred_s=1, radius=width)
    leaves = sim.cells()
    sim.set_values(f(centers(leaves)))
    sim.update_ghost_mr()
    ref = sim.cells(REFERENCE)
    c = centers(ref)
    nout = np.sum((c[:, :dim] < 0) | (c[:, :dim] > 1), axis=1)
    corner = ref[nout >= 2]
    n = 8
    expected = 4 * width ** 2 if dim == 2 else 8 * width ** 3 + 12 * n * width ** 2
    assert len(corner) == expected
    got = sim.storage()[corner[:, 4]]
    assert np.max(np.abs(got - corner_oracle(centers(corner), dim, f))) <= 1e-11


def normal_ghosts(sim, dim, axis, sgn, width):
    """(layer, ghost row, its boundary leaf) for the ghosts of every boundary
    leaf along the outward normal (axis, sgn), layers 1..width."""
    leaves = sim.cells(CELLS)
    ref = sim.cells(REFERENCE)
    pos = {tuple(int(v) for v in r[:4]): int(r[4]) for r in ref}
    out = []
    for r in leaves:
        l = int(r[0])
        n = 1 << l
        g = [int(r[1]), int(r[2]), int(r[3])]
        if not (g[axis] == (n - 1 if sgn > 0 else 0)):
            continue
        for k in range(1, width + 1):
            q = list(g)
            q[axis] += sgn * k
            key = (l, q[0], q[1], q[2])
            assert key in pos, f"ghost {key} missing from the reference mesh"
            out.append((k, np.array([l, q[0], q[1], q[2], pos[key]]), r))
    return out


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("width", [2, 3])
def test_far_layers_extend_a_linear_field(dim, width):
    """run_further<dim, 1> (tests/test_bc_ghost_values.cpp:388-440) with the
    field f = 1 + 3 x_n for each normal n and the Dirichlet value f on the
    face: the B.C. is exact on layer 1 and the extrapolations of stencil 4
    and 6 on layers 2 and 3, so every ghost along the normal holds f there."""
    level = 4
    checked = 0
    for axis, sgn, _ in cartesian_directions(dim):
        sim = OracleSim(dim, level, level, pred_s=1, radius=width)
        f = lambda rows: 1.0 + 3.0 * centers(rows)[:, axis]
        sim.set_values(f(sim.cells()))
        sim.set_dirichlet(1.0 + 3.0 * (1.0 if sgn > 0 else 0.0))
        sim.update_ghost_mr()
        st = sim.storage()
        for k, g, _ in normal_ghosts(sim, dim, axis, sgn, width):
            assert abs(st[g[4]] - f(g[None, :])[0]) <= 1e-11, (axis, sgn, k, g)
            checked += 1
    assert checked == 2 * dim * width * 16 ** (dim - 1)


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("width", [1, 2, 3])
def test_far_layers_without_bc(dim, width):
    """Without a B.C. every layer is extrapolated (ghost_layers_filled_by_bc
    = 0, bc/apply_field_bc.hpp:504-532): layer 1 copies the boundary cell b,
    and for a linear f the polynomial of degree 2k - 2 through f(b - (k-1)h),
    ..., f(b), then the layers already filled, takes at layer k the value
    f(b - (k - 1) h).  (Lagrange at x = 2 through -1, 0, 1 with values
    f(-1), f0, f0 gives f(-1) * 1 + f0 * (-3 + 3); at x = 3 through -2..2 with
    values f(-2), f(-1), f0, f0, f(-1): f(-2) + f(-1) * (-5 + 5) + f0 * (10 - 10).)"""
    level = 4
    h = 1.0 / (1 << level)
    for axis, sgn, _ in cartesian_directions(dim):
        sim = OracleSim(dim, level, level, pred_s=1, radius=width)
        f = lambda rows: 1.0 + 3.0 * centers(rows)[:, axis] + 0.25 * centers(rows)[:, (axis + 1) % dim]
        sim.set_values(f(sim.cells()))
        sim.update_ghost_mr()
        st = sim.storage()
        for k, g, b in normal_ghosts(sim, dim, axis, sgn, width):
            inner = np.array(b[:4], dtype=np.int64).copy()
            inner[1 + axis] -= sgn * (k - 1)
            assert abs(st[g[4]] - f(inner[None, :])[0]) <= 1e-12, (axis, sgn, k)


def avg_poly(coeffs, i):
    """Average over the cell [i - 1/2, i + 1/2] of sum c_p x^p."""
    a, b = i - 0.5, i + 0.5
    return sum(c * (b ** (p + 1) - a ** (p + 1)) / (p + 1) for p, c in enumerate(coeffs))


@pytest.mark.parametrize("coeffs", [(2.0, -0.75), (1.0, 0.5, 0.25), (-3.0, 0.0, 1.0)])
def test_weno5_flux_exact_on_quadratic_averages(coeffs):
    """Each third-order candidate reconstructs the interface value of a
    polynomial of degree <= 2 from its cell averages, so the weighted sum does
    too: f[i] = average over cell i - 2 (i = 0..4), interface at x = 1/2."""
    f = [avg_poly(coeffs, i - 2) for i in range(5)]
    exact = sum(c * 0.5 ** p for p, c in enumerate(coeffs))
    assert abs(weno5_flux(f) - exact) <= 1e-12


def test_weno5_flux_eno_at_a_jump():
    """A jump between cells 0 and 1 (f = 0, 0, 0, 1, 1): the candidates that
    cross it get weights of order eps^2 / IS^2, the interface value stays at
    the left state."""
    assert abs(weno5_flux([0.0, 0.0, 0.0, 1.0, 1.0])) < 1e-10
    assert abs(weno5_flux([1.0, 1.0, 1.0, 0.0, 0.0]) - 1.0) < 1e-10


def test_weno5_flux_fifth_order_on_smooth_data():
    """On averages of sin over cells of width h the error of the interface
    value shrinks like h^5 (the linear weights give the fifth-order
    reconstruction; other weights stop at third order)."""
    errs = []
    for h in (0.1, 0.05, 0.025):
        x0 = 0.3
        f = [(math.cos(x0 + (i - 2.5) * h) - math.cos(x0 + (i - 1.5) * h)) / h for i in range(5)]
        errs.append(abs(weno5_flux(f) - math.sin(x0 + 0.5 * h)))  # right edge of cell 2
    o1, o2 = math.log2(errs[0] / errs[1]), math.log2(errs[1] / errs[2])
    assert o1 > 4.5 and o2 > 4.5, errs


def interior(rows, dim, width):
    n = 1 << rows[:, 0].astype(np.int64)
    ok = np.ones(len(rows), dtype=bool)
    for d in range(dim):
        x = rows[:, 1 + d]
        ok &= (x >= width) & (x < n - width)
    return ok


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("v", [(1.5, 0.0, 0.0), (-0.75, 1.25, 0.5), (0.5, -2.0, -1.0)])
def test_weno5_operator_linear_field(dim, v):
    """u = 1 + 2x - 3y (+ 0.5z): cell averages are the centre values, every
    interface value is exact, so at cells whose stencils stay inside the
    domain conv(u) = sum_d v_d * du/dx_d."""
    level = 4 if dim == 2 else 3
    sim = OracleSim(dim, level, level, pred_s=1, radius=3)
    rows = sim.cells()
    c = centers(rows)
    g = np.array([2.0, -3.0, 0.5])[:dim]
    sim.set_values(1.0 + c[:, :dim] @ g)
    sim.set_dirichlet(0.0)
    out = sim.convection_weno5(np.array(v))
    inner = interior(rows, dim, 3)
    want = float(np.dot(np.array(v[:dim]), g))
    assert np.max(np.abs(out[rows[inner, 4]] - want)) <= 1e-10


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("vx", [1.5, -0.75])
def test_weno5_operator_exact_to_the_wall(dim, vx):
    """u = a x with Dirichlet 0 vanishes on the wall x = 0, so the three ghost
    layers there extend it exactly (the B.C. on layer 1, the extrapolations
    beyond), and u = a (x - 1) does the same at x = 1: with v = (vx, 0, 0)
    every cell within three cells of that wall, boundary cells included,
    gets conv(u) = vx * a (the transverse fluxes are 0 * u)."""
    level = 4 if dim == 2 else 3
    a = 2.5
    for shift, near in ((0.0, lambda x, n: x < n - 3), (1.0, lambda x, n: x >= 3)):
        sim = OracleSim(dim, level, level, pred_s=1, radius=3)
        rows = sim.cells()
        sim.set_values(a * (centers(rows)[:, 0] - shift))
        sim.set_dirichlet(0.0)
        out = sim.convection_weno5(np.array([vx, 0.0, 0.0]))
        sel = near(rows[:, 1], 1 << level)
        assert np.max(np.abs(out[rows[sel, 4]] - vx * a)) <= 1e-10, shift


def weno5_error(level, dim=2):
    """The reference's weno5_error_1d (tests/test_fv_operators.cpp:901-936) on
    a uniform 2D mesh: u = sin(2 pi x) at the centres, Dirichlet<1> 0,
    velocity (1.5, 0); the largest error against 1.5 * 2 pi cos(2 pi x) on
    cells at least 3 cells from the boundary."""
    sim = OracleSim(dim, level, level, pred_s=1, radius=3)
    rows = sim.cells()
    x = centers(rows)[:, 0]
    k = 2 * math.pi
    sim.set_values(np.sin(k * x))
    sim.set_dirichlet(0.0)
    out = sim.convection_weno5(np.array([1.5, 0.0, 0.0]))
    inner = interior(rows, dim, 3)
    return float(np.max(np.abs(out[rows[inner, 4]] - 1.5 * k * np.cos(k * x[inner]))))


def test_weno5_operator_convergence():
    """weno5_convergence_1d (tests/test_fv_operators.cpp:938-951): observed
    orders between levels 4, 5 and 6 above 4."""
    e4, e5, e6 = weno5_error(4), weno5_error(5), weno5_error(6)
    assert math.log2(e4 / e5) > 4.0 and math.log2(e5 / e6) > 4.0, (e4, e5, e6)


def test_radius_beyond_one_needs_one_level():
    with pytest.raises(Exception):
        OracleSim(2, 2, 4, pred_s=1, radius=3)
