"""Restart files with fields on the GPU: dump and load of an adapted mesh with
two scalar fields (the reference's restart_mrmesh with fields,
tests/test_restart.cpp:96-103 and io/restart.hpp:189-237, 341-462), and a run
continued from a restart file, which must match the uninterrupted run bit for
bit (the restart of demos/FiniteVolume/advection_2d.cpp:155-160)."""
import sys

import numpy as np
import pytest

import samurai_amd as sa
from initial_conditions import disc_2d, sphere_3d

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not __import__("os").path.exists(sa.IO_LIB_PATH),
                                                  reason="restart I/O library not built (no HDF5)")]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def bits(a):
    return np.ascontiguousarray(a).view(np.uint64)


def adapted_field(gpu_ctx, dim, lmin, lmax, eps):
    mesh = sa.make_mesh(gpu_ctx, sa.mesh_config(dim, lmin, lmax, 1))
    u = sa.make_scalar_field("u", mesh)
    c = mesh.cells()
    u.set_cells(disc_2d(c) if dim == 2 else sphere_3d(c))
    sa.make_bc_dirichlet(u, 0.0)
    ad = sa.make_MRAdapt(u)
    ad(eps, 1.0)
    return mesh, u, ad


@pytest.mark.parametrize("dim,lmin,lmax", [(2, 3, 8), (3, 2, 5)])
def test_mesh_and_fields_round_trip(tmp_path, gpu_ctx, dim, lmin, lmax):
    say("test_mesh_and_fields_round_trip", dim, lmin, lmax)
    mesh, u, _ = adapted_field(gpu_ctx, dim, lmin, lmax, 1e-3)
    v = sa.make_scalar_field("v", mesh)
    v.set_cells(np.random.default_rng(5).standard_normal(mesh.nb_cells()))
    sa.dump(tmp_path / "restart", mesh, u, v)
    m2, (u2, v2) = sa.load(tmp_path / "restart", gpu_ctx, sa.mesh_config(dim, 1, 2, 1), "u", "v")
    assert (m2.min_level, m2.max_level) == (lmin, lmax)
    for mid in (sa.CELLS, sa.CELLS_AND_GHOSTS, sa.PROJ_CELLS, sa.REFERENCE):
        assert np.array_equal(mesh.cells(mid), m2.cells(mid)), mid
    assert np.array_equal(bits(u.cells()), bits(u2.cells()))
    assert np.array_equal(bits(v.cells()), bits(v2.cells()))
    with pytest.raises(sa.SamuraiError, match="field w does not exist"):
        sa.load(tmp_path / "restart", gpu_ctx, sa.mesh_config(dim, 1, 2, 1), "w")
    # in place, with fields made on the mesh beforehand; a missing field
    # leaves mesh and fields as they were
    m3 = sa.make_mesh(gpu_ctx, sa.mesh_config(dim, 1, 2, 1))
    u3, v3, w3 = (sa.make_scalar_field(n, m3) for n in ("u", "v", "w"))
    with pytest.raises(sa.SamuraiError, match="field w does not exist"):
        sa.load_into(tmp_path / "restart", m3, u3, w3)
    assert (m3.min_level, m3.max_level) == (1, 2) and u3.size() == m3.nb_cells(sa.REFERENCE)
    sa.load_into(tmp_path / "restart", m3, v3, u3)
    assert np.array_equal(mesh.cells(sa.REFERENCE), m3.cells(sa.REFERENCE))
    assert np.array_equal(bits(u.cells()), bits(u3.cells()))
    assert np.array_equal(bits(v.cells()), bits(v3.cells()))


def advance(mesh, u, ad, n, dt, a):
    unp1 = sa.make_scalar_field("unp1", mesh)
    for _ in range(n):
        ad(2e-4, 1.0)
        sa.update_ghost_mr(u)
        unp1.resize()
        sa.upwind(a, dt, u, unp1)
        sa.swap(u, unp1)


@pytest.mark.parametrize("into", [False, True])
def test_restarted_run_matches_the_uninterrupted_run(tmp_path, gpu_ctx, into):
    """Advection of the disc on levels 4..9: 8 steps, dump, 8 more; the run
    loaded from the dump and advanced 8 steps ends on the same leaves with the
    same values, bit for bit.  into: loaded as the reference demo does, into an
    empty mesh whose field was made before the load (demos/FiniteVolume/
    advection_2d.cpp:102-114)."""
    say("test_restarted_run_matches_the_uninterrupted_run", into)
    lmin, lmax = 4, 9
    a = np.array([1.0, 1.0])
    dt = 0.5 / (1 << lmax)
    mesh, u, ad = adapted_field(gpu_ctx, 2, lmin, lmax, 2e-4)
    advance(mesh, u, ad, 8, dt, a)
    sa.dump(tmp_path / "advection_2d_restart", mesh, u)
    advance(mesh, u, ad, 8, dt, a)

    if into:
        m2 = sa.make_mesh_from_cells(gpu_ctx, sa.mesh_config(2, lmin, lmax, 1), np.zeros((0, 4), dtype=np.int32))
        u2 = sa.make_scalar_field("u", m2)
        sa.load_into(tmp_path / "advection_2d_restart", m2, u2)
    else:
        m2, (u2,) = sa.load(tmp_path / "advection_2d_restart", gpu_ctx, sa.mesh_config(2, lmin, lmax, 1), "u")
    sa.make_bc_dirichlet(u2, 0.0)
    ad2 = sa.make_MRAdapt(u2)
    advance(m2, u2, ad2, 8, dt, a)
    assert np.array_equal(mesh.cells(), m2.cells())
    assert np.array_equal(bits(u.cells()), bits(u2.cells()))


@pytest.mark.parametrize("s", [0, 1])
def test_save_of_the_advection_demo_matches_reference_files(tmp_path, gpu_ctx, s):
    """samurai::save of the advection_2d demo (demos/FiniteVolume/advection_2d.cpp:
    48-57, 124, 151) run on the GPU: the initial and final files have the
    reference's points, connectivity and XDMF text, and its values within
    1e-12 (the tolerance of the demo's field parity)."""
    say("test_save_of_the_advection_demo_matches_reference_files", s)
    from test_gpu_parity import run_demo
    from test_save import STEM, check_against_reference, fixture_cells
    saved = []

    def on_init(mesh, u):
        sa.save(tmp_path, STEM.format(s=s, suffix="_init"), mesh, u)
        saved.append("_init")

    (_, _), (_, _), n, (mesh, u) = run_demo(gpu_ctx, s, on_init=on_init, keep=True)
    assert n == 21
    sa.save(tmp_path, STEM.format(s=s, suffix=""), mesh, u)
    for when, name in (("_init", f"advection_2d_pred{s}_init.npz"), ("", f"advection_2d_pred{s}_t0.01.npz")):
        got = check_against_reference(tmp_path, STEM.format(s=s, suffix=when))
        _, u_ref = fixture_cells(name)
        assert np.max(np.abs(got - u_ref)) <= 1e-12, name
    assert saved == ["_init"]
