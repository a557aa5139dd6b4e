"""The mesh update of a harten call on the device (update_cell_array_from_tag
and the make_graduation fixed point on per-level bitmaps,
algorithm/graduation.hpp:573-843) against the host interval-algebra version
(itself checked against the oracle): the same leaves, storage numbering and
ghost layers from the same tags, for tags a harten call produces and for
synthetic ones that stress the rules (random refinements, whole families
coarsened, arbitrary flag bits)."""
import numpy as np
import pytest

import samurai_amd as sa
from initial_conditions import seeded_2d, sphere_3d
from oracle_lib import OracleSim

pytestmark = pytest.mark.gpu


def adapted_cells(dim, lmin, lmax, seed):
    o = OracleSim(dim, lmin, lmax, pred_s=1)
    o.set_poison(False)
    c = o.cells()
    o.set_values(seeded_2d(c, seed) if dim == 2 else sphere_3d(c, radius=0.2 + 0.02 * seed))
    o.set_dirichlet(0.0)
    o.adapt(1e-3, 2.0)
    return o.cells()


def synthetic_tags(cells, n, kind, rng, dim):
    tags = np.zeros(n, np.uint8)
    idx = cells[:, 4]
    if kind == "refine":
        tags[idx] = np.where(rng.random(len(cells)) < 0.05, 4, 1)
    elif kind == "families":
        tags[idx] = 1
        tags[idx[rng.random(len(cells)) < 0.03]] = 4
        fam = {}
        for r in cells:
            fam.setdefault((r[0], r[1] >> 1, r[2] >> 1, r[3] >> 1), []).append(r[4])
        for members in fam.values():
            if len(members) == 1 << dim and rng.random() < 0.3:
                tags[members] = 2
    else:
        tags[idx] = rng.integers(0, 8, len(cells))
    return tags


@pytest.mark.parametrize("dim,lmin,lmax,seed", [(2, 2, 7, 1), (2, 3, 9, 2), (3, 2, 5, 0), (3, 1, 6, 1)])
@pytest.mark.parametrize("kind", ["refine", "families", "bits"])
def test_device_mesh_update_matches_host(gpu_ctx, dim, lmin, lmax, seed, kind):
    import torch
    print("device mesh update", dim, lmin, lmax, seed, kind, flush=True)
    cells = adapted_cells(dim, lmin, lmax, seed)
    cfg = sa.mesh_config(dim, lmin, lmax, 1)
    mh = sa.make_mesh_from_cells(gpu_ctx, cfg, cells)
    md = sa.make_mesh_from_cells(gpu_ctx, cfg, cells)
    rng = np.random.default_rng(100 * seed + lmax + len(kind))
    for rnd in range(2):
        n = mh.nb_cells(sa.REFERENCE)
        tags = synthetic_tags(mh.cells(), n, kind, rng, dim)
        t = torch.tensor(tags, dtype=torch.uint8, device="cuda")
        torch.cuda.synchronize()
        rh = mh.rebuild_from_tags(tags)
        rd = md.rebuild_from_device_tags(t.data_ptr(), n)
        assert rh == rd, rnd
        for mid in (sa.CELLS, sa.CELLS_AND_GHOSTS, sa.PROJ_CELLS, sa.UNION_CELLS, sa.REFERENCE):
            assert np.array_equal(mh.cells(mid), md.cells(mid)), (rnd, mid)


def test_device_mesh_update_at_the_fixed_point(gpu_ctx):
    """Tags that keep every leaf change nothing (return 1, mesh untouched)."""
    import torch
    cells = adapted_cells(2, 2, 8, 3)
    m = sa.make_mesh_from_cells(gpu_ctx, sa.mesh_config(2, 2, 8, 1), cells)
    n = m.nb_cells(sa.REFERENCE)
    t = torch.ones(n, dtype=torch.uint8, device="cuda")
    torch.cuda.synchronize()
    before = m.cells(sa.REFERENCE)
    assert m.rebuild_from_device_tags(t.data_ptr(), n) == 1
    assert np.array_equal(before, m.cells(sa.REFERENCE))


@pytest.mark.parametrize("dim,lmin,lmax,seed", [(2, 2, 8, 4), (2, 4, 10, 1), (3, 2, 6, 0)])
def test_device_built_work_lists_match_host(gpu_ctx, dim, lmin, lmax, seed):
    """A harten call that changes the mesh builds the new mesh's work lists on
    the device (projection, prediction, detail, criteria, maximum and the
    update_fields lists): each must equal the host's build_programs list of
    the same mesh, run for run, with the same cell prefix, chunk hints and
    per-level offsets, and the time steps must stay those of the oracle."""
    print("device lists", dim, lmin, lmax, seed, flush=True)
    cfg = sa.mesh_config(dim, lmin, lmax, 1)
    mesh = sa.make_mesh(gpu_ctx, cfg)
    u = sa.make_scalar_field("u", mesh)
    c = mesh.cells()
    u.set_cells(seeded_2d(c, seed) if dim == 2 else sphere_3d(c, radius=0.25))
    sa.make_bc_dirichlet(u, 0.0)
    o = OracleSim(dim, lmin, lmax, pred_s=1)
    o.set_poison(False)
    oc = o.cells()
    o.set_values(seeded_2d(oc, seed) if dim == 2 else sphere_3d(oc, radius=0.25))
    o.set_dirichlet(0.0)
    unp1 = sa.make_scalar_field("unp1", mesh)
    adapt = sa.make_MRAdapt(u)
    a = np.array([1.0, 0.5, 0.25][:dim])
    dt = 0.4 / dim / (1 << lmax)
    checked = 0
    for step in range(3):
        assert o.adapt(1e-3, 2.0) == adapt(1e-3, 2.0), step
        rc = mesh.check_device_lists()
        assert rc in (0, 1), rc
        checked += rc == 0
        o.upwind_step(a, dt)
        sa.update_ghost_mr(u)
        unp1.resize()
        sa.upwind(a, dt, u, unp1)
        sa.swap(u, unp1)
        assert np.array_equal(o.cells(), mesh.cells()), step
        assert np.array_equal(o.values().view(np.uint64), u.cells().view(np.uint64)), step
    assert checked > 0
