"""Batched form: many independent adaptive instances advanced by one stage
program per phase must give bit-identical meshes and fields to the same
instances advanced one by one through the plain entry points (the advection
demo loop body, demos/FiniteVolume/advection_2d.cpp:129-153)."""
import numpy as np
import pytest

from initial_conditions import batch_field, batch_instance

pytestmark = pytest.mark.gpu

EPS, REG = 1e-4, 2.0


def make_instances(sa, ctx, seeds, levels, min_level=2):
    out = []
    for seed in seeds:
        max_level, a, nb, ng = batch_instance(seed, levels)
        mesh = sa.make_mesh(ctx, sa.mesh_config(2, min_level, max_level, 1))
        u = sa.make_scalar_field("u", mesh)
        u.set_cells(batch_field(mesh.cells(), seed, nb, ng))
        sa.make_bc_dirichlet(u, 0.0)
        unp1 = sa.make_scalar_field("unp1", mesh)
        out.append({"mesh": mesh, "u": u, "unp1": unp1, "a": np.array(a), "dt": 0.5 * 2.0 ** -max_level,
                    "adapt": sa.make_MRAdapt(u)})
    return out


def plain_step(sa, inst, adapt=True):
    if adapt:
        inst["adapt"](EPS, REG)
    sa.update_ghost_mr(inst["u"])
    inst["unp1"].resize()
    sa.upwind(inst["a"], inst["dt"], inst["u"], inst["unp1"])
    sa.swap(inst["u"], inst["unp1"])


def assert_same(p, b, what):
    import samurai_amd as sa
    for k, (x, y) in enumerate(zip(p, b)):
        for mid in (sa.CELLS, sa.CELLS_AND_GHOSTS, sa.PROJ_CELLS, sa.UNION_CELLS, sa.REFERENCE):
            cx, cy = x["mesh"].cells(mid), y["mesh"].cells(mid)
            assert np.array_equal(cx, cy), f"{what}: instance {k} mesh id {mid} differs"
        for name in ("u", "unp1"):
            sx, sy = x[name].storage(), y[name].storage()
            assert sx.shape == sy.shape, f"{what}: instance {k} {name} storage size"
            assert np.array_equal(sx.view(np.uint64), sy.view(np.uint64)), f"{what}: instance {k} {name} storage differs"


def setup_pair(sa, ctx, seeds, levels):
    plain = make_instances(sa, ctx, seeds, levels)
    bat = make_instances(sa, ctx, seeds, levels)
    for x in plain + bat:
        x["adapt"](EPS, REG)
    batch = sa.Batch(ctx, [(x["u"], x["unp1"], x["a"], x["dt"]) for x in bat])
    return plain, bat, batch


# (cells per thread, mode); cells per thread 0: precomputed wave items
@pytest.mark.parametrize("cpt,mode", [(1, 0), (4, 1), (1, 2), (0, 1), (0, 2)])
def test_batch_steps_match_plain(gpu_ctx, cpt, mode):
    import samurai_amd as sa
    seeds = [11, 12, 13, 14, 15, 16]
    plain, bat, batch = setup_pair(sa, gpu_ctx, seeds, (4, 7))
    batch.set_tuning(256, cpt, mode)
    for step in range(4):
        for x in plain:
            plain_step(sa, x)
        batch.step(EPS, REG)
        gpu_ctx.synchronize()
        assert_same(plain, bat, f"mode {mode} step {step}")
    assert batch.nb_cells() == sum(x["mesh"].nb_cells() for x in plain)


def test_batch_mesh_ids_match_host_construction(gpu_ctx):
    """The batched device construction (update_sub_mesh_impl of every changed
    instance per launch) against the host's construction of the same leaves:
    all five ids and the storage numbering, after each of four steps; and the
    work lists built on the device for every changed instance against the
    host's programs of the same mesh."""
    import samurai_amd as sa
    plain, bat, batch = setup_pair(sa, gpu_ctx, [61, 62, 63, 64, 65], (4, 8))
    checked = 0
    for step in range(4):
        batch.step(EPS, REG)
        gpu_ctx.synchronize()
        for k, y in enumerate(bat):
            m = y["mesh"]
            host = sa.make_mesh_from_cells(gpu_ctx, m.cfg, m.cells())
            for mid in (sa.CELLS, sa.CELLS_AND_GHOSTS, sa.PROJ_CELLS, sa.UNION_CELLS, sa.REFERENCE):
                assert np.array_equal(host.cells(mid), m.cells(mid)), (step, k, mid)
            rc = m.check_device_lists()
            assert rc in (0, 1), (step, k, rc)
            checked += rc == 0
    assert checked > 0


@pytest.mark.parametrize("block,cpt", [(64, 1), (128, 8), (1024, 2)])
def test_batch_launch_shapes(gpu_ctx, block, cpt):
    import samurai_amd as sa
    plain, bat, batch = setup_pair(sa, gpu_ctx, [21, 22, 23], (4, 6))
    for mode in (0, 1):
        batch.set_tuning(block, cpt, mode)
        for x in plain:
            plain_step(sa, x)
        batch.step(EPS, REG)
        gpu_ctx.synchronize()
        assert_same(plain, bat, f"block {block} cpt {cpt} mode {mode}")


def test_batch_graph_replays(gpu_ctx):
    """On frozen meshes the advance program is unchanged for each u/unp1
    parity, so mode 2 replays its hipGraph: every replay is checked against
    the eager plain calls made between replays."""
    import samurai_amd as sa
    plain, bat, batch = setup_pair(sa, gpu_ctx, [31, 32, 33, 34], (4, 6))
    batch.set_tuning(256, 2, 2)
    for rep in range(6):
        for x in plain:
            plain_step(sa, x, adapt=False)
        batch.advance()
        gpu_ctx.synchronize()
        assert_same(plain, bat, f"replay {rep}")


def test_batch_autotune_persists(gpu_ctx, tmp_path):
    import samurai_amd as sa
    plain, bat, batch = setup_pair(sa, gpu_ctx, [41, 42], (4, 6))
    f = str(tmp_path / "tune.txt")
    t = batch.autotune(reps=1, cache_file=f)
    assert t["block"] % 64 == 0 and t["mode"] in (0, 1, 2)
    batch.set_tuning(64, 1, 1)
    assert batch.load_tuning(f)
    assert batch.tuning() == t
    for x in plain:
        plain_step(sa, x)
    batch.step(EPS, REG)
    gpu_ctx.synchronize()
    assert_same(plain, bat, "after autotune")


def test_batch_rejects_shared_mesh(gpu_ctx):
    import samurai_amd as sa
    x = make_instances(sa, gpu_ctx, [51], (4, 5))[0]
    w = sa.make_scalar_field("w", x["mesh"])
    with pytest.raises(sa.SamuraiError):
        sa.Batch(gpu_ctx, [(x["u"], x["unp1"], x["a"], x["dt"]), (w, x["unp1"], x["a"], x["dt"])])


def test_batch_3d_and_mixed_dimensions(gpu_ctx):
    """3D instances (and a 2D one beside them) in one batch."""
    import samurai_amd as sa
    from initial_conditions import sphere_3d

    def make(dim, lmax, seed):
        mesh = sa.make_mesh(gpu_ctx, sa.mesh_config(dim, 2, lmax, 1))
        u = sa.make_scalar_field("u", mesh)
        cells = mesh.cells()
        if dim == 3:
            u.set_cells(sphere_3d(cells, center=(0.45 + 0.02 * seed, 0.5, 0.55), radius=0.2))
        else:
            u.set_cells(batch_field(cells, seed, 1, 1))
        sa.make_bc_dirichlet(u, 0.0)
        a = np.array([1.0, -0.5, 0.25][:dim])
        return {"mesh": mesh, "u": u, "unp1": sa.make_scalar_field("unp1", mesh), "a": a,
                "dt": 0.4 / 3 * 2.0 ** -lmax, "adapt": sa.make_MRAdapt(u)}

    specs = [(3, 4, 1), (3, 5, 2), (2, 6, 3)]
    plain = [make(*sp) for sp in specs]
    bat = [make(*sp) for sp in specs]
    for x in plain + bat:
        x["adapt"](1e-3, REG)
    batch = sa.Batch(gpu_ctx, [(x["u"], x["unp1"], x["a"], x["dt"]) for x in bat])
    batch.set_tuning(256, 2, 1)
    for step in range(3):
        for x in plain:
            x["adapt"](1e-3, REG)
            sa.update_ghost_mr(x["u"])
            x["unp1"].resize()
            sa.upwind(x["a"], x["dt"], x["u"], x["unp1"])
            sa.swap(x["u"], x["unp1"])
        batch.step(1e-3, REG)
        gpu_ctx.synchronize()
        assert_same(plain, bat, f"3d step {step}")
