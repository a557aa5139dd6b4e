"""smr_adapt_tags_noalloc (details in a static per-GPU device array, no
allocation) against the allocating smr_adapt_tags: identical tags and ghost
values, on two calls in a row, the second on a smaller mesh; and its two
argument errors (a storage larger than the array, programs not built)."""
import numpy as np
import pytest

from initial_conditions import seeded_2d

pytestmark = pytest.mark.gpu

EPS, REG = 1e-4, 2.0


def adapted(sa, ctx, lmin, lmax, seed):
    mesh = sa.make_mesh(ctx, sa.mesh_config(2, lmin, lmax, 1))
    u = sa.make_scalar_field("u", mesh)
    u.set_cells(seeded_2d(mesh.cells(), seed))
    sa.make_bc_dirichlet(u, 0.0)
    ad = sa.make_MRAdapt(u)
    ad(EPS, REG)  # builds the mesh version's programs (allocating)
    return mesh, u, ad


def test_noalloc_tags_match_allocating_entry_point(gpu_ctx):
    import samurai_amd as sa
    import torch
    print("no-alloc harten tags vs smr_adapt_tags", flush=True)
    for lmin, lmax, seed in [(3, 9, 4), (2, 6, 5)]:  # the second mesh is smaller
        mesh, u, ad = adapted(sa, gpu_ctx, lmin, lmax, seed)
        n = mesh.nb_cells(sa.REFERENCE)
        for ite in range(2):
            ref = torch.full((n,), 0xEE, dtype=torch.uint8, device="cuda")
            got = torch.full((n,), 0x77, dtype=torch.uint8, device="cuda")
            torch.cuda.synchronize()
            ad.tags_into(ite, EPS, REG, ref.data_ptr())
            gpu_ctx.synchronize()
            st_ref = u.storage()
            ad.tags_into_noalloc(ite, EPS, REG, got.data_ptr())
            gpu_ctx.synchronize()
            assert np.array_equal(ref.cpu().numpy(), got.cpu().numpy()), (lmax, ite)
            assert np.array_equal(st_ref.view(np.uint64), u.storage().view(np.uint64)), (lmax, ite)
            leaves = mesh.cells()[:, 4]
            assert bool((got.cpu().numpy()[leaves] != 0).all()), "every leaf carries a flag"


def test_noalloc_argument_errors(gpu_ctx):
    import samurai_amd as sa
    import torch
    # storage beyond the 2^21 slots of the static array
    big = sa.make_mesh(gpu_ctx, sa.mesh_config(2, 10, 11, 1))
    assert big.nb_cells(sa.REFERENCE) > (1 << 21)
    u = sa.make_scalar_field("u", big)
    ad = sa.make_MRAdapt(u)
    t = torch.zeros(big.nb_cells(sa.REFERENCE), dtype=torch.uint8, device="cuda")
    with pytest.raises(sa.SamuraiError, match="exceed"):
        ad.tags_into_noalloc(0, EPS, REG, t.data_ptr())
    # a mesh version whose programs were never built
    small = sa.make_mesh(gpu_ctx, sa.mesh_config(2, 2, 5, 1))
    v = sa.make_scalar_field("v", small)
    ad2 = sa.make_MRAdapt(v)
    t2 = torch.zeros(small.nb_cells(sa.REFERENCE), dtype=torch.uint8, device="cuda")
    with pytest.raises(sa.SamuraiError, match="not built"):
        ad2.tags_into_noalloc(0, EPS, REG, t2.data_ptr())
