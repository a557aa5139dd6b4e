"""Host logic of the product (interval algebra, MRMesh ids, storage numbering,
tag-driven rebuild with graduation) against the oracle, and the C ABI surface.
No device is touched: the context is host-only."""
import ctypes
import os
import re

import numpy as np
import pytest

import samurai_amd as sa
from initial_conditions import disc_2d, sphere_3d
from oracle_lib import OracleSim

IDS = [sa.CELLS, sa.CELLS_AND_GHOSTS, sa.PROJ_CELLS, sa.REFERENCE]


def same_mesh(o, m):
    for mid in IDS:
        a, b = o.cells(mid), m.cells(mid)
        assert a.shape == b.shape, (mid, a.shape, b.shape)
        assert np.array_equal(a, b), mid


@pytest.mark.parametrize("dim,lmin,lmax,s", [(2, 2, 5, 1), (2, 4, 7, 0), (2, 3, 3, 1), (3, 2, 4, 1), (3, 1, 3, 0)])
def test_box_mesh_matches_oracle(host_ctx, dim, lmin, lmax, s):
    o = OracleSim(dim, lmin, lmax, pred_s=s)
    m = sa.make_mesh(host_ctx, sa.mesh_config(dim, lmin, lmax, s))
    same_mesh(o, m)


@pytest.mark.parametrize("dim,lmin,lmax,s,eps", [(2, 3, 7, 1, 2e-4), (2, 2, 6, 0, 1e-3), (3, 2, 5, 1, 1e-3)])
def test_rebuild_from_oracle_tags(host_ctx, dim, lmin, lmax, s, eps):
    """Every harten iteration: the product's host rebuild from the oracle's tags
    yields the oracle's new mesh (all ids, storage numbering)."""
    o = OracleSim(dim, lmin, lmax, pred_s=s)
    cells = o.cells()
    o.set_values(disc_2d(cells) if dim == 2 else sphere_3d(cells))
    o.set_dirichlet(0.0)
    m = sa.make_mesh(host_ctx, sa.mesh_config(dim, lmin, lmax, s))
    for ite in range(lmax - lmin):
        fixed = o.harten(ite, eps, 2.0)
        tags, _ = o.last_tags()
        assert m.rebuild_from_tags(tags) == fixed
        same_mesh(o, m)
        if fixed:
            break


def test_mesh_from_cells_matches_oracle(host_ctx):
    o = OracleSim(2, 4, 8, pred_s=1)
    o.set_values(disc_2d(o.cells()))
    o.set_dirichlet(0.0)
    o.adapt(2e-4, 1.0)
    m = sa.make_mesh_from_cells(host_ctx, sa.mesh_config(2, 4, 8, 1), o.cells())
    same_mesh(o, m)


@pytest.mark.parametrize("dim,level,r", [(2, 4, 2), (2, 3, 3), (3, 3, 3)])
def test_wide_ghost_mesh_matches_oracle(host_ctx, dim, level, r):
    """A one-level mesh of stencil radius 2 or 3 (max_stencil_size 4 or 6):
    ghost layers of that width, ids and numbering as the oracle's, programs
    (the far-layer and corner-block recipes among them) in bounds."""
    o = OracleSim(dim, level, level, pred_s=1, radius=r)
    m = sa.make_mesh(host_ctx, sa.mesh_config(dim, level, level, 1, r))
    same_mesh(o, m)
    assert m.validate() == 0


def test_invalid_configs_raise(host_ctx):
    with pytest.raises(sa.SamuraiError):
        sa.make_mesh(host_ctx, sa.mesh_config(2, 6, 4))          # max < min
    with pytest.raises(sa.SamuraiError):
        sa.make_mesh(host_ctx, sa.mesh_config(2, 2, 5, 1, 2))    # stencil radius 2
    with pytest.raises(sa.SamuraiError):
        sa.make_mesh(host_ctx, sa.mesh_config(4, 2, 5))          # dim 4


def test_device_call_on_host_context_raises(host_ctx):
    m = sa.make_mesh(host_ctx, sa.mesh_config(2, 2, 4))
    with pytest.raises(sa.SamuraiError):
        sa.make_scalar_field("u", m)


def test_c_abi_exports_every_declared_symbol():
    header = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "samurai_amd.h")
    text = open(header).read()
    names = sorted(set(re.findall(r"\b(smr_[a-z_0-9]+)\s*\(", text)))
    assert len(names) >= 25
    L = ctypes.CDLL(sa.LIB_PATH)
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


@pytest.mark.parametrize("dim,lmin,lmax,s,eps", [(2, 4, 10, 1, 2e-4), (2, 4, 10, 0, 2e-4), (2, 2, 8, 1, 1e-4),
                                                 (3, 2, 5, 1, 1e-3), (3, 2, 5, 0, 1e-3)])
def test_programs_in_bounds_along_adaptation(host_ctx, dim, lmin, lmax, s, eps):
    """Every mesh met while adapting (including the advection_2d demo meshes)
    gives work lists whose rows exist and whose access windows stay inside one
    reference interval: the condition under which the device kernels run."""
    o = OracleSim(dim, lmin, lmax, pred_s=s)
    cells = o.cells()
    o.set_values(disc_2d(cells) if dim == 2 else sphere_3d(cells))
    o.set_dirichlet(0.0)
    m = sa.make_mesh(host_ctx, sa.mesh_config(dim, lmin, lmax, s))
    assert m.validate() == 0
    for ite in range(lmax - lmin):
        fixed = o.harten(ite, eps, 1.0)
        tags, _ = o.last_tags()
        assert m.rebuild_from_tags(tags) == fixed   # also validates the update_fields lists
        assert m.validate() == 0
        if fixed:
            break


def test_batch_needs_a_gpu_context(host_ctx):
    import samurai_amd as sa
    with pytest.raises(sa.SamuraiError):
        sa.Batch(host_ctx, [])
