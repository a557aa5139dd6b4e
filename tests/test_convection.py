"""Flux-based linear upwind convection, make_convection_upwind(v)(u)
(schemes/fv/operators/convection_lin.hpp:15-89).

The oracle is pinned by the reference's own known answer
(tests/test_fv_operators.cpp:155-225: on a uniform mesh a linear field
u = sum_d (d+2) x_d with v_d = d+1 gives sum_d v_d (d+2) on interior cells)
and by invariants on adapted meshes: conservation (interior fluxes cancel, so
sum area * out only holds boundary fluxes) and exact zeros for a constant
field away from the boundary, including across level jumps.  The product's
host-built ordered contribution lists are then evaluated here in numpy and
must reproduce the oracle bit for bit (the GPU kernel is checked against the
oracle in test_gpu_parity.py)."""
import numpy as np
import pytest

import samurai_amd as sa
from initial_conditions import centers, sphere_3d
from oracle_lib import REFERENCE, OracleSim


def uniform(dim, level):
    o = OracleSim(dim, level, level, pred_s=1)
    return o


def adapted(dim, lmin, lmax, field, eps=1e-3):
    o = OracleSim(dim, lmin, lmax, pred_s=1)
    o.set_values(field(o.cells()))
    o.set_dirichlet(0.0)
    o.adapt(eps, 2.0)
    return o


def cell_slots(o):
    return o.cells()[:, 4].astype(np.int64)


@pytest.mark.parametrize("dim", [2, 3])
def test_linear_field_exact_on_interior_cells(dim):
    level = 3
    o = uniform(dim, level)
    cells = o.cells()
    b = np.array([d + 2.0 for d in range(dim)])
    v = np.array([d + 1.0 for d in range(dim)])
    o.set_values(centers(cells)[:, :dim] @ b)
    o.set_dirichlet(0.0)
    out = o.convection_upwind(v)[cell_slots(o)]
    n = 1 << level
    interior = np.all((cells[:, 1:1 + dim] >= 1) & (cells[:, 1:1 + dim] <= n - 2), axis=1)
    assert interior.sum() > 0
    assert np.allclose(out[interior], float(v @ b), rtol=0, atol=1e-10)


def field_2d(cells):
    # a compactly supported bump plus a box, both zero near the boundary of the unit square
    c = centers(cells)
    r2 = (c[:, 0] - 0.45) ** 2 + (c[:, 1] - 0.55) ** 2
    bump = np.maximum(0.0, 1.0 - r2 / 0.25 ** 2) ** 2
    box = ((np.abs(c[:, 0] - 0.6) <= 0.1) & (np.abs(c[:, 1] - 0.35) <= 0.08)).astype(np.float64)
    return bump + box


@pytest.mark.parametrize("dim,lmin,lmax,v", [(2, 2, 7, (1.0, 0.5)), (2, 3, 8, (-0.7, 1.3)), (2, 2, 6, (0.0, -1.0)),
                                             (3, 2, 5, (1.0, -0.5, 0.25))])
def test_conservation_on_adapted_meshes(dim, lmin, lmax, v):
    field = field_2d if dim == 2 else (lambda c: 0.5 * sphere_3d(c, radius=0.2))
    o = adapted(dim, lmin, lmax, field)
    cells = o.cells()
    assert len(np.unique(cells[:, 0])) > 1, "the mesh must hold level jumps"
    out = o.convection_upwind(np.array(v))[cell_slots(o)]
    area = np.ldexp(1.0, -dim * cells[:, 0].astype(np.int64))
    total = float(np.sum(area * out))
    scale = float(np.sum(area * np.abs(out)))
    assert scale > 0
    assert abs(total) <= 1e-12 * scale


@pytest.mark.parametrize("dim,lmin,lmax", [(2, 2, 7), (3, 2, 5)])
def test_constant_field_gives_zero_away_from_boundary(dim, lmin, lmax):
    field = field_2d if dim == 2 else (lambda c: 0.5 * sphere_3d(c, radius=0.2))
    o = adapted(dim, lmin, lmax, field)
    o.set_values(np.ones(o.nb_cells()))
    v = np.array([0.8, -0.6, 0.3][:dim])
    out = o.convection_upwind(v)[cell_slots(o)]
    cells = o.cells()
    # boundary ghosts (2*0 - 1) reach the prediction stencils of cells a few cells
    # from the boundary; 4 cells in at the cell's level is clear of them
    n = np.left_shift(1, cells[:, 0].astype(np.int64))
    inner = np.all((cells[:, 1:1 + dim] >= 4) & (cells[:, 1:1 + dim] <= (n - 5)[:, None]), axis=1)
    assert inner.sum() > 0
    if dim == 2:
        # 2D: every cancellation is exact ((-a + a/2) + a/2 at a jump)
        assert np.all(out[inner] == 0.0)
    else:
        # 3D jumps add four quarters, -a + a/4 can round: zero to rounding
        scale = np.abs(v).max() * np.ldexp(1.0, cells[:, 0].astype(np.int64))
        assert np.all(np.abs(out[inner]) <= 1e-13 * scale[inner])


def evaluate_terms(terms, u, nref):
    row, ptr, src, coef = terms
    out = np.zeros(nref)
    for r in range(len(row)):
        acc = 0.0
        for k in range(ptr[r], ptr[r + 1]):
            s0, s1 = src[2 * k], src[2 * k + 1]
            if s1 < 0:
                acc += coef[k] * u[s0]
            else:
                acc += coef[k] * u[s0] + coef[k] * u[s1]
        out[row[r]] = acc
    return out


@pytest.mark.parametrize("dim,lmin,lmax,v", [(2, 2, 7, (1.0, 0.5)), (2, 3, 7, (-0.7, 1.3)), (3, 2, 5, (1.0, -0.5, 0.25)),
                                             (3, 2, 5, (-0.2, 0.9, -1.0))])
def test_host_terms_reproduce_oracle(host_ctx, dim, lmin, lmax, v):
    field = field_2d if dim == 2 else (lambda c: 0.5 * sphere_3d(c, radius=0.2))
    o = adapted(dim, lmin, lmax, field)
    ref = o.convection_upwind(np.array(v))  # runs update_ghost_mr on the oracle's u
    u = o.storage()
    m = sa.make_mesh_from_cells(host_ctx, sa.mesh_config(dim, lmin, lmax, 1), o.cells())
    assert m.nb_cells(sa.REFERENCE) == o.nb_cells(REFERENCE)
    got = evaluate_terms(sa.convection_terms(m, np.array(v)), u, len(u))
    assert np.array_equal(got.view(np.uint64), ref.view(np.uint64))


# ---------------------------------------------------------------------------
# non-linear make_convection_upwind<Field>() (convection_nonlin.hpp:24-79):
# flux f(u) = u * u of the side picked by the sign of the mean value.  The
# reference's tests hold no known answer for it; the oracle is pinned by
# closed forms derived here and by the invariants above.
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("sign", [1.0, -1.0])
def test_nonlinear_closed_form_on_linear_fields(dim, sign):
    """u = sign * (3 + sum_d b_d x_d), never zero.  With u > 0 every flux is
    u_left^2, so on an interior cell the x_d part is (u_i^2 - u_{i-1}^2) / h =
    b_d (2 u_i - b_d h); with u < 0 it is (u_{i+1}^2 - u_i^2) / h =
    b'_d (2 u_i + b'_d h), b' = sign * b."""
    level = 3
    o = uniform(dim, level)
    cells = o.cells()
    b = np.array([d + 2.0 for d in range(dim)])
    c = centers(cells)[:, :dim]
    u = sign * (3.0 + c @ b)
    o.set_values(u)
    o.set_dirichlet(0.0)
    out = o.convection_upwind_nonlinear()[cell_slots(o)]
    h = 1.0 / (1 << level)
    bs = sign * b
    expected = sum(bs[d] * (2 * u - sign * bs[d] * h) for d in range(dim))
    n = 1 << level
    interior = np.all((cells[:, 1:1 + dim] >= 1) & (cells[:, 1:1 + dim] <= n - 2), axis=1)
    assert interior.sum() > 0
    assert np.allclose(out[interior], expected[interior], rtol=0, atol=1e-10)


def test_nonlinear_upwind_side_follows_the_mean():
    """One row of a uniform 2D mesh, u = +1 on x < 1/2 and -3 elsewhere: at
    the jump the mean is -1 < 0, so the flux is (-3)^2 = 9 and the cell left
    of the jump (x index 3 of 8) gets (9 - 1) / h in x; a mean >= 0 (u = +3
    then -1) takes the left value 3^2 = 9 again, giving 0 there."""
    level = 3
    h = 1.0 / (1 << level)
    for left, right, want in ((1.0, -3.0, (9.0 - 1.0) / h), (3.0, -1.0, 0.0)):
        o = uniform(2, level)
        cells = o.cells()
        u = np.where(cells[:, 1] < 4, left, right).astype(np.float64)
        o.set_values(u)
        o.set_dirichlet(0.0)
        out = o.convection_upwind_nonlinear()[cell_slots(o)]
        sel = (cells[:, 1] == 3) & (cells[:, 2] == 4)
        # y fluxes cancel on the interior row (u constant along y there)
        assert out[sel][0] == want


@pytest.mark.parametrize("dim,lmin,lmax", [(2, 2, 7), (2, 3, 8), (3, 2, 5)])
def test_nonlinear_conservation_on_adapted_meshes(dim, lmin, lmax):
    field = field_2d if dim == 2 else (lambda c: 0.5 * sphere_3d(c, radius=0.2))
    o = adapted(dim, lmin, lmax, field)
    cells = o.cells()
    assert len(np.unique(cells[:, 0])) > 1, "the mesh must hold level jumps"
    # mixed signs so that both upwind sides occur
    o.set_values(field(cells) - 0.3 * field(cells[:, [0, 2, 1, 3, 4]]))
    out = o.convection_upwind_nonlinear()[cell_slots(o)]
    area = np.ldexp(1.0, -dim * cells[:, 0].astype(np.int64))
    total = float(np.sum(area * out))
    scale = float(np.sum(area * np.abs(out)))
    assert scale > 0
    assert abs(total) <= 1e-12 * scale


@pytest.mark.parametrize("dim,lmin,lmax", [(2, 2, 7), (3, 2, 5)])
def test_nonlinear_constant_field_gives_zero_away_from_boundary(dim, lmin, lmax):
    field = field_2d if dim == 2 else (lambda c: 0.5 * sphere_3d(c, radius=0.2))
    o = adapted(dim, lmin, lmax, field)
    o.set_values(np.full(o.nb_cells(), -1.0))
    out = o.convection_upwind_nonlinear()[cell_slots(o)]
    cells = o.cells()
    n = np.left_shift(1, cells[:, 0].astype(np.int64))
    inner = np.all((cells[:, 1:1 + dim] >= 4) & (cells[:, 1:1 + dim] <= (n - 5)[:, None]), axis=1)
    assert inner.sum() > 0
    if dim == 2:
        assert np.all(out[inner] == 0.0)
    else:
        scale = np.ldexp(1.0, cells[:, 0].astype(np.int64))
        assert np.all(np.abs(out[inner]) <= 1e-13 * scale[inner])


def evaluate_burgers_terms(terms, u, nref):
    row, ptr, src, coef = terms
    out = np.zeros(nref)
    for r in range(len(row)):
        acc = 0.0
        for k in range(ptr[r], ptr[r + 1]):
            ul, ur = u[src[2 * k]], u[src[2 * k + 1]]
            f = ul * ul if 0.5 * (ul + ur) >= 0 else ur * ur
            acc += f * coef[k]
        out[row[r]] = acc
    return out


@pytest.mark.parametrize("dim,lmin,lmax", [(2, 2, 7), (2, 3, 7), (3, 2, 5)])
def test_host_burgers_terms_reproduce_oracle(host_ctx, dim, lmin, lmax):
    field = field_2d if dim == 2 else (lambda c: 0.5 * sphere_3d(c, radius=0.2))
    o = adapted(dim, lmin, lmax, field)
    cells = o.cells()
    o.set_values(field(cells) - 0.4 * field(cells[:, [0, 2, 1, 3, 4]]))
    ref = o.convection_upwind_nonlinear()
    u = o.storage()
    m = sa.make_mesh_from_cells(host_ctx, sa.mesh_config(dim, lmin, lmax, 1), cells)
    got = evaluate_burgers_terms(sa.convection_terms(m), u, len(u))
    assert np.array_equal(got.view(np.uint64), ref.view(np.uint64))


# ---------------------------------------------------------------------------
# make_convection_upwind<Field>(velocity_field) (convection_lin.hpp:184-264):
# velocity given per cell, as one scalar field per component
# ---------------------------------------------------------------------------

def velocity_storages(o, comps, bcs):
    """Ghost-updated storages of the velocity components (values on o's cells)."""
    return [o.ghost_updated(c, bc) for c, bc in zip(comps, bcs)]


@pytest.mark.parametrize("dim", [2, 3])
def test_velocity_field_known_answer(dim):
    """tests/test_fv_operators.cpp:640-715: uniform level-3 mesh, velocity
    component d = d + 1 everywhere (Dirichlet with the same values), u =
    sum_d (d + 2) x_d with Dirichlet 0: sum_d (d + 1)(d + 2) on interior cells."""
    level = 3
    o = uniform(dim, level)
    cells = o.cells()
    w = [float(d + 1) for d in range(dim)]
    b = np.array([d + 2.0 for d in range(dim)])
    o.set_values(centers(cells)[:, :dim] @ b)
    o.set_dirichlet(0.0)
    vel = velocity_storages(o, [np.full(len(cells), wd) for wd in w], w)
    out = o.convection_upwind_heterogeneous(vel)[cell_slots(o)]
    n = 1 << level
    interior = np.all((cells[:, 1:1 + dim] >= 1) & (cells[:, 1:1 + dim] <= n - 2), axis=1)
    assert interior.sum() > 0
    assert np.allclose(out[interior], float(np.dot(w, b)), rtol=0, atol=1e-10)


@pytest.mark.parametrize("dim,lmin,lmax,v", [(2, 2, 7, (1.0, -0.5)), (2, 3, 8, (-0.7, 1.3)), (3, 2, 5, (1.0, -0.5, 0.25))])
def test_constant_velocity_field_is_the_linear_operator(dim, lmin, lmax, v):
    """With a constant velocity field (and its boundary values) the
    heterogeneous operator is make_convection_upwind(v): same interfaces,
    factors and upwind sides; the reference adds the terms of a cell in another
    order (cell by cell instead of interval by interval), so the two agree to
    rounding."""
    field = field_2d if dim == 2 else (lambda c: 0.5 * sphere_3d(c, radius=0.2))
    o = adapted(dim, lmin, lmax, field)
    n = o.nb_cells()
    vel = velocity_storages(o, [np.full(n, vd) for vd in v], v)
    het = o.convection_upwind_heterogeneous(vel)[cell_slots(o)]
    lin = o.convection_upwind(np.array(v))[cell_slots(o)]
    scale = np.abs(lin).max()
    assert scale > 0
    assert np.max(np.abs(het - lin)) <= 1e-13 * scale


@pytest.mark.parametrize("dim,lmin,lmax", [(2, 2, 7), (3, 2, 5)])
def test_velocity_field_conservation(dim, lmin, lmax):
    """A rotating velocity field: interior fluxes cancel in sum area * out
    (the field u vanishes near the boundary)."""
    field = field_2d if dim == 2 else (lambda c: 0.5 * sphere_3d(c, radius=0.2))
    o = adapted(dim, lmin, lmax, field)
    c = centers(o.cells())
    comps = [-(c[:, 1] - 0.5), c[:, 0] - 0.5] + ([0.3 * np.sin(6 * c[:, 0])] if dim == 3 else [])
    out = o.convection_upwind_heterogeneous(velocity_storages(o, comps, [0.0] * dim))[cell_slots(o)]
    area = np.ldexp(1.0, -dim * o.cells()[:, 0].astype(np.int64))
    total, scale = float(np.sum(area * out)), float(np.sum(area * np.abs(out)))
    assert scale > 0
    assert abs(total) <= 1e-12 * scale


# ---------------------------------------------------------------------------
# upwind_scalar_burgers(k, u) (stencil_field.hpp:181-255), the stencil operator
# of demos/FiniteVolume/scalar_burgers_2d.cpp; no known answer in the
# reference's tests: pinned by closed forms and the flux's four cases
# ---------------------------------------------------------------------------

def test_scalar_burgers_closed_form():
    """u = 3 + 2x + 3y > 0, k = (1, 1): every flux is 0.5 * u_left^2 (the
    left value is the smaller one, same signs), so an interior cell moves by
    -dt * sum_d 0.5 b_d (2 u - b_d h)."""
    level = 4
    o = uniform(2, level)
    cells = o.cells()
    b = np.array([2.0, 3.0])
    u = 3.0 + centers(cells)[:, :2] @ b
    o.set_values(u)
    o.set_dirichlet(0.0)
    dt, h = 1e-3, 1.0 / (1 << level)
    o.upwind_scalar_burgers_step(np.array([1.0, 1.0]), dt)
    got = o.values()
    expected = u - dt * sum(0.5 * b[d] * (2 * u - b[d] * h) for d in range(2))
    n = 1 << level
    interior = np.all((cells[:, 1:3] >= 1) & (cells[:, 1:3] <= n - 2), axis=1)
    assert np.allclose(got[interior], expected[interior], rtol=0, atol=1e-12)


@pytest.mark.parametrize("ul,ur,flux", [(1.0, 2.0, 0.5), (2.0, 1.0, 2.0), (-1.0, 2.0, 0.0), (2.0, -3.0, 4.5), (-2.0, -1.0, 0.5)])
def test_scalar_burgers_flux_cases(ul, ur, flux):
    """k = (1, 1), a row of cells at x < 1/2 holding ul and x >= 1/2 holding
    ur: the interior cell left of the jump sees F_right - F_left = flux -
    0.5 ul^2 on the x axis (the y fluxes cancel on an interior row):
    increasing same sign -> 0.5 min^2, decreasing -> 0.5 max^2, rarefaction
    (ul < 0 < ur) -> 0, shock -> 0.5 max^2."""
    level = 3
    o = uniform(2, level)
    cells = o.cells()
    o.set_values(np.where(cells[:, 1] < 4, ul, ur).astype(np.float64))
    o.set_dirichlet(0.0)
    dt, h = 1e-3, 1.0 / (1 << level)
    o.upwind_scalar_burgers_step(np.array([1.0, 1.0]), dt)
    sel = (cells[:, 1] == 3) & (cells[:, 2] == 4)
    left_flux = 0.5 * ul * ul  # both sides hold ul: 0.5 * ul^2 in every case
    want = ul - dt * ((flux - left_flux) / h)
    assert o.values()[sel][0] == want


def test_scalar_burgers_agrees_with_nonlinear_convection_on_positive_fields():
    """For u > 0 and k > 0 both operators take the left value on every
    interior interface: (u - unp1) / dt = 0.5 * make_convection_upwind<Field>()(u)
    on interior cells of a uniform mesh, to rounding (two independently pinned
    operators, different operation order)."""
    level = 5
    o = uniform(2, level)
    cells = o.cells()
    u = 1.5 + np.sin(3 * centers(cells)[:, 0]) * np.cos(2 * centers(cells)[:, 1]) * 0.5
    o.set_values(u)
    o.set_dirichlet(0.0)
    conv = o.convection_upwind_nonlinear()[cell_slots(o)]
    dt = 1e-4
    o.upwind_scalar_burgers_step(np.array([np.sqrt(2) / 2, np.sqrt(2) / 2]), dt)
    rate = (u - o.values()) / dt
    n = 1 << level
    interior = np.all((cells[:, 1:3] >= 1) & (cells[:, 1:3] <= n - 2), axis=1)
    assert np.allclose(rate[interior], 0.5 * conv[interior], rtol=1e-9, atol=1e-9)
