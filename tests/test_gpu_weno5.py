"""WENO5 convection and the wide ghost layers it needs, on the GPU through the
C ABI, against the oracle bit for bit (same operation order, no contraction):

- update_ghost_mr on one-level meshes of stencil radius 2 and 3 (corner
  blocks and far layers by polynomial extrapolation), with and without a B.C.;
- fields without a B.C. on adapted radius-1 meshes (layer 1 extrapolated from
  the boundary cells);
- make_convection_weno5(v)(u) in 2D and 3D for velocities of every sign;
- the upwind step on a radius-3 mesh.
The closed forms of the WENO5 operator (exact on linear fields, fifth order)
are checked on the GPU output as well."""
import math
import sys

import numpy as np
import pytest

import samurai_amd as sa
from initial_conditions import centers, disc_2d, sphere_3d
from oracle_lib import OracleSim

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def bits(a):
    return np.ascontiguousarray(a).view(np.uint64)


def wide_pair(gpu_ctx, dim, level, r, bc=True):
    """An oracle sim and a product field on the same one-level mesh of stencil
    radius r, identical storage (ghosts zero), Dirichlet 0 when bc."""
    o = OracleSim(dim, level, level, pred_s=1, radius=r)
    o.set_poison(False)
    mesh = sa.make_mesh(gpu_ctx, sa.mesh_config(dim, level, level, 1, r))
    assert np.array_equal(o.cells(sa.REFERENCE), mesh.cells(sa.REFERENCE))
    u = sa.make_scalar_field("u", mesh)
    if bc:
        o.set_dirichlet(0.0)
        sa.make_bc_dirichlet(u, 0.0)
    return o, mesh, u


@pytest.mark.parametrize("dim,level,r", [(2, 5, 2), (2, 5, 3), (3, 3, 2), (3, 4, 3)])
@pytest.mark.parametrize("bc", [True, False])
def test_wide_ghost_update_matches_oracle(gpu_ctx, dim, level, r, bc):
    say("wide ghost update", dim, level, r, bc)
    o, mesh, u = wide_pair(gpu_ctx, dim, level, r, bc)
    st = np.random.default_rng(11).random(mesh.nb_cells(sa.REFERENCE))
    o.set_storage(st)
    u.set_storage(st)
    o.update_ghost_mr()
    sa.update_ghost_mr(u)
    assert np.array_equal(bits(o.storage()), bits(u.storage()))


@pytest.mark.parametrize("dim,lmin,lmax", [(2, 2, 6), (3, 2, 5)])
def test_ghost_update_without_bc_matches_oracle(gpu_ctx, dim, lmin, lmax):
    """A field without a B.C. on an adapted mesh: its layer-1 ghosts beside the
    boundary cells of every level copy those cells, after the B.C. prediction
    of the level has read them (update_outer_ghost.hpp:414-426)."""
    say("ghost update without bc", dim, lmin, lmax)
    o = OracleSim(dim, lmin, lmax, pred_s=1)
    o.set_poison(False)
    cells = o.cells()
    o.set_values(disc_2d(cells) if dim == 2 else sphere_3d(cells))
    o.set_dirichlet(0.0)
    o.adapt(1e-3, 2.0)
    mesh = sa.make_mesh_from_cells(gpu_ctx, sa.mesh_config(dim, lmin, lmax, 1), o.cells())
    o2 = OracleSim(dim, lmin, lmax, pred_s=1, cells=o.cells())
    o2.set_poison(False)
    u = sa.make_scalar_field("u", mesh)
    st = np.random.default_rng(5).random(mesh.nb_cells(sa.REFERENCE))
    o2.set_storage(st)
    u.set_storage(st)
    for _ in range(2):
        o2.update_ghost_mr()
        sa.update_ghost_mr(u)
        assert np.array_equal(bits(o2.storage()), bits(u.storage()))


VELOCITIES = [(1.5, 0.0, 0.0), (-0.75, 1.25, 0.5), (0.5, -2.0, -1.0), (0.0, 0.0, -0.25)]


@pytest.mark.parametrize("dim,level", [(2, 6), (3, 4)])
@pytest.mark.parametrize("v", VELOCITIES)
@pytest.mark.parametrize("path", ["dense", "untiled", "lists"])
def test_weno5_matches_oracle(gpu_ctx, dim, level, v, path, monkeypatch):
    """Every kernel: the list-free one over the dense storage box (the
    default on one-level meshes; in 2D on tiles that share the interface
    fluxes through LDS, SMR_WENO5_UNTILED=1 for one thread per cell without
    sharing) and the ordered-list gather (SMR_WENO5_LISTS=1)."""
    if path == "lists":
        monkeypatch.setenv("SMR_WENO5_LISTS", "1")
    if path == "untiled":
        monkeypatch.setenv("SMR_WENO5_UNTILED", "1")
    say("weno5", dim, level, v, path)
    o, mesh, u = wide_pair(gpu_ctx, dim, level, 3)
    rows = mesh.cells()
    c = centers(rows)
    vals = np.sin(2 * np.pi * c[:, 0]) * np.cos(np.pi * c[:, 1]) + (c[:, 0] > 0.5) * 0.75
    if dim == 3:
        vals = vals + 0.5 * c[:, 2] * c[:, 2]
    o.set_values(vals)
    u.set_cells(vals)
    ref = o.convection_weno5(np.array(v))
    conv = sa.make_convection_weno5(v[:dim])
    got = conv(u).storage()
    assert np.array_equal(bits(got), bits(ref))
    assert np.array_equal(bits(conv(u).storage()), bits(ref))


def test_weno5_closed_forms_on_gpu(gpu_ctx):
    """v . grad u on a linear field at cells whose stencils stay inside, and the
    reference's convergence check (tests/test_fv_operators.cpp:898-951) on
    sin(2 pi x): observed orders above 4 between levels 4, 5 and 6."""
    say("weno5 closed forms")
    o, mesh, u = wide_pair(gpu_ctx, 2, 5, 3)
    rows = mesh.cells()
    c = centers(rows)
    u.set_cells(1.0 + 2.0 * c[:, 0] - 3.0 * c[:, 1])
    out = sa.make_convection_weno5([-0.75, 1.25])(u).cells()
    n = 32
    inner = (rows[:, 1] >= 3) & (rows[:, 1] < n - 3) & (rows[:, 2] >= 3) & (rows[:, 2] < n - 3)
    assert np.max(np.abs(out[inner] - (-0.75 * 2.0 + 1.25 * -3.0))) <= 1e-10
    errs = []
    for level in (4, 5, 6):
        o, mesh, u = wide_pair(gpu_ctx, 2, level, 3)
        rows = mesh.cells()
        x = centers(rows)[:, 0]
        u.set_cells(np.sin(2 * np.pi * x))
        out = sa.make_convection_weno5([1.5, 0.0])(u).cells()
        n = 1 << level
        inner = (rows[:, 1] >= 3) & (rows[:, 1] < n - 3) & (rows[:, 2] >= 3) & (rows[:, 2] < n - 3)
        errs.append(np.max(np.abs(out[inner] - 1.5 * 2 * np.pi * np.cos(2 * np.pi * x[inner]))))
    assert math.log2(errs[0] / errs[1]) > 4.0 and math.log2(errs[1] / errs[2]) > 4.0, errs


def test_weno5_needs_a_wide_one_level_mesh(gpu_ctx):
    mesh = sa.make_mesh(gpu_ctx, sa.mesh_config(2, 4, 4, 1, 1))
    u = sa.make_scalar_field("u", mesh)
    with pytest.raises(sa.SamuraiError):
        sa.make_convection_weno5([1.0, 0.0])(u)


@pytest.mark.parametrize("dim,level", [(2, 6), (3, 4)])
def test_upwind_on_a_radius3_mesh(gpu_ctx, dim, level):
    say("upwind radius 3", dim, level)
    o, mesh, u = wide_pair(gpu_ctx, dim, level, 3)
    v = np.random.default_rng(3).random(mesh.nb_cells())
    o.set_values(v)
    u.set_cells(v)
    unp1 = sa.make_scalar_field("unp1", mesh)
    a = np.array([np.cos(np.pi / 5), -np.sin(np.pi / 5), 0.3][:dim])
    dt = 0.5 / (1 << level)
    o.upwind_step(a, dt)
    sa.update_ghost_mr(u)
    sa.upwind(a, dt, u, unp1)
    assert np.array_equal(bits(o.values()), bits(unp1.cells()))
