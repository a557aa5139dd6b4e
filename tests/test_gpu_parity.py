"""Parity of the HIP path (through the C ABI) with the oracle and with the
reference's golden fixtures.  Integer/flag data must be identical; FP64 fields
are compared bit for bit against the oracle (same operation order, no
contraction) and within 1e-12 against the fixtures."""
import os
import sys

import numpy as np
import pytest

import samurai_amd as sa
from initial_conditions import disc_2d, seeded_2d, sphere_3d
from oracle_lib import REFERENCE, OracleSim

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def pair(gpu_ctx, dim, lmin, lmax, s, ic="disc", adapt_eps=None, seed=0):
    """An oracle sim and a product field on the same (optionally adapted) mesh,
    with identical full storage (ghosts zero) and a Dirichlet-0 BC."""
    o = OracleSim(dim, lmin, lmax, pred_s=s)
    o.set_poison(False)
    cells = o.cells()
    if ic == "disc":
        v = disc_2d(cells) if dim == 2 else sphere_3d(cells)
    else:
        v = seeded_2d(cells, seed)
    o.set_values(v)
    o.set_dirichlet(0.0)
    if adapt_eps is not None:
        o.adapt(adapt_eps, 2.0)
    mesh = sa.make_mesh_from_cells(gpu_ctx, sa.mesh_config(dim, lmin, lmax, s), o.cells())
    u = sa.make_scalar_field("u", mesh)
    sa.make_bc_dirichlet(u, 0.0)
    st = o.storage()
    st = np.where(np.isnan(st), 0.0, st)
    o.set_storage(st)
    u.set_storage(st)
    return o, mesh, u


def bits(a):
    return np.ascontiguousarray(a).view(np.uint64)


@pytest.mark.parametrize("dim,level", [(2, 6), (3, 4)])
def test_uniform_upwind(gpu_ctx, dim, level):
    say("test_uniform_upwind", dim, level)
    o, mesh, u = pair(gpu_ctx, dim, level, level, 1)
    rng = np.random.default_rng(3)
    v = rng.random(mesh.nb_cells())
    o.set_values(v)
    u.set_cells(v)
    unp1 = sa.make_scalar_field("unp1", mesh)
    a = np.array([np.cos(np.pi / 5), np.sin(np.pi / 5), 0.3][:dim])
    dt = 0.5 / (1 << level)
    o.upwind_step(a, dt)
    sa.update_ghost_mr(u)
    sa.upwind(a, dt, u, unp1)
    assert np.array_equal(bits(o.values()), bits(unp1.cells()))


@pytest.mark.parametrize("dim,lmin,lmax,s", [(2, 2, 6, 1), (2, 3, 7, 0), (3, 2, 5, 1), (3, 2, 4, 0)])
def test_ghost_update_matches_oracle(gpu_ctx, dim, lmin, lmax, s):
    say("test_ghost_update_matches_oracle", dim, lmin, lmax, s)
    o, mesh, u = pair(gpu_ctx, dim, lmin, lmax, s, adapt_eps=1e-3)
    rng = np.random.default_rng(7)
    st = rng.random(mesh.nb_cells(sa.REFERENCE))
    o.set_storage(st)
    u.set_storage(st)
    o.update_ghost_mr()
    sa.update_ghost_mr(u)
    assert np.array_equal(bits(o.storage()), bits(u.storage()))


@pytest.mark.parametrize("dim,lmin,lmax,s", [(2, 2, 6, 1), (3, 2, 5, 1)])
def test_upwind_adapted_matches_oracle(gpu_ctx, dim, lmin, lmax, s):
    say("test_upwind_adapted_matches_oracle", dim, lmin, lmax, s)
    o, mesh, u = pair(gpu_ctx, dim, lmin, lmax, s, adapt_eps=1e-3)
    unp1 = sa.make_scalar_field("unp1", mesh)
    a = np.array([1.0, 0.5, -0.25][:dim])
    dt = 0.4 / (1 << lmax)
    o.upwind_step(a, dt)
    sa.update_ghost_mr(u)
    sa.upwind(a, dt, u, unp1)
    assert np.array_equal(bits(o.values()), bits(unp1.cells()))


@pytest.mark.parametrize("dim,lmin,lmax,s,ic", [(2, 3, 7, 1, "disc"), (2, 2, 7, 1, "seeded"), (2, 3, 6, 0, "disc"),
                                                (3, 2, 5, 1, "disc")])
def test_harten_matches_oracle(gpu_ctx, dim, lmin, lmax, s, ic):
    """Flags, details, new mesh and transferred field of every harten call."""
    say("test_harten_matches_oracle", dim, lmin, lmax, s, ic)
    o, mesh, u = pair(gpu_ctx, dim, lmin, lmax, s, ic=ic)
    import torch
    ad = sa.make_MRAdapt(u)
    eps = 1e-4 if dim == 2 else 1e-3
    for ite in range(lmax - lmin):
        # the device half alone (smr_adapt_tags) first, into a caller's array
        dev_tags = torch.full((mesh.nb_cells(sa.REFERENCE),), 0xEE, dtype=torch.uint8, device="cuda")
        torch.cuda.synchronize()
        ad.tags_into(ite, eps, 2.0, dev_tags.data_ptr())
        gpu_ctx.synchronize()
        fo = o.harten(ite, eps, 2.0)
        fp = ad.harten(ite, eps, 2.0)
        to, do = o.last_tags()
        tp, dp = ad.last_tags()
        assert np.array_equal(to, dev_tags.cpu().numpy()), ite
        assert fo == fp
        assert np.array_equal(to, tp), ite
        assert np.array_equal(bits(do), bits(dp)), ite
        assert np.array_equal(o.cells(), mesh.cells()), ite
        assert np.array_equal(o.cells(REFERENCE), mesh.cells(sa.REFERENCE)), ite
        assert np.array_equal(bits(o.storage()), bits(u.storage())), ite
        if fo:
            break


def run_demo(gpu_ctx, s, n_steps=None, on_init=None, keep=False):
    """demos/FiniteVolume/advection_2d.cpp to Tf = 0.01; on_init(mesh, u) runs
    where the demo saves its initial state; keep adds (mesh, u) to the result."""
    cfg = sa.mesh_config(2, 4, 10, s)
    mesh = sa.make_mesh(gpu_ctx, cfg)
    u = sa.make_scalar_field("u", mesh)
    u.set_cells(disc_2d(mesh.cells()))
    sa.make_bc_dirichlet(u, 0.0)
    unp1 = sa.make_scalar_field("unp1", mesh)
    adapt = sa.make_MRAdapt(u)
    adapt(2e-4, 1.0)
    init = (mesh.cells(), u.cells())
    if on_init is not None:
        on_init(mesh, u)
    dt, Tf, t, n = 0.5 * (1.0 / (1 << 10)), 0.01, 0.0, 0
    a = np.array([1.0, 1.0])
    while t != Tf and (n_steps is None or n < n_steps):
        adapt(2e-4, 1.0)
        t += dt
        if t > Tf:
            dt += Tf - t
            t = Tf
        sa.update_ghost_mr(u)
        unp1.resize()
        sa.upwind(a, dt, u, unp1)
        sa.swap(u, unp1)
        n += 1
    if keep:
        return init, (mesh.cells(), u.cells()), n, (mesh, u)
    return init, (mesh.cells(), u.cells()), n


@pytest.mark.parametrize("s", [0, 1])
def test_advection_2d_demo_against_fixtures(gpu_ctx, s):
    say("test_advection_2d_demo_against_fixtures", s)
    (c0, v0), (c1, v1), n = run_demo(gpu_ctx, s)
    assert n == 21
    for (c, v), name in [((c0, v0), f"advection_2d_pred{s}_init.npz"), ((c1, v1), f"advection_2d_pred{s}_t0.01.npz")]:
        fx = np.load(os.path.join(GOLDEN, name))
        assert c.shape[0] == fx["level"].size, name
        assert np.array_equal(c[:, 0], fx["level"]) and np.array_equal(c[:, 1], fx["i"]) and np.array_equal(c[:, 2], fx["j"])
        assert np.max(np.abs(v - fx["u"])) <= 1e-12, name


@pytest.mark.parametrize("dim,lmin,lmax", [(2, 3, 7), (3, 2, 5)])
def test_time_steps_match_oracle(gpu_ctx, dim, lmin, lmax):
    """Adapt + ghost update + upwind for several steps, bit for bit."""
    say("test_time_steps_match_oracle", dim, lmin, lmax)
    o, mesh, u = pair(gpu_ctx, dim, lmin, lmax, 1)
    unp1 = sa.make_scalar_field("unp1", mesh)
    ad = sa.make_MRAdapt(u)
    a = np.array([1.0, 1.0, 1.0][:dim])
    dt = (0.5 if dim == 2 else 0.4 / 3) / (1 << lmax)
    eps = 1e-4 if dim == 2 else 1e-3
    for step in range(4):
        no = o.adapt(eps, 2.0)
        npr = ad(eps, 2.0)
        assert no == npr
        o.upwind_step(a, dt)
        sa.update_ghost_mr(u)
        unp1.resize()
        sa.upwind(a, dt, u, unp1)
        sa.swap(u, unp1)
        assert np.array_equal(o.cells(), mesh.cells()), step
        assert np.array_equal(bits(o.values()), bits(u.cells())), step


@pytest.mark.parametrize("dim,lmin,lmax,v", [(2, 2, 7, (1.0, 0.5)), (2, 3, 7, (-0.7, 1.3)), (3, 2, 5, (1.0, -0.5, 0.25)),
                                             (3, 2, 5, (-0.2, 0.9, -1.0))])
def test_convection_upwind_matches_oracle(gpu_ctx, dim, lmin, lmax, v):
    """make_convection_upwind(v)(u) on adapted meshes (level jumps in every direction)."""
    say("convection", dim, lmin, lmax, v)
    o, mesh, u = pair(gpu_ctx, dim, lmin, lmax, 1, adapt_eps=1e-3)
    ref = o.convection_upwind(np.array(v))
    out = sa.make_convection_upwind(np.array(v))(u)
    got = out.storage()
    assert np.array_equal(bits(got), bits(ref))
    # the cached contribution lists are reused for a second field on the mesh
    out2 = sa.make_convection_upwind(np.array(v))(u)
    assert np.array_equal(bits(out2.storage()), bits(ref))


def mixed_sign(o, dim):
    """A field with both signs on the oracle's cells: the disc (sphere) minus
    a shifted copy, so that both upwind sides of the Burgers flux occur."""
    c = o.cells()
    f = disc_2d if dim == 2 else sphere_3d
    shifted = c.copy()
    shifted[:, 1] = c[:, 2]
    shifted[:, 2] = c[:, 1]
    return f(c) - 0.6 * f(shifted)


@pytest.mark.parametrize("dim,lmin,lmax", [(2, 2, 7), (2, 3, 8), (3, 2, 5)])
def test_convection_nonlinear_matches_oracle(gpu_ctx, dim, lmin, lmax):
    """make_convection_upwind<Field>()(u) (scalar, flux u * u) on adapted meshes."""
    say("convection nonlinear", dim, lmin, lmax)
    o, mesh, u = pair(gpu_ctx, dim, lmin, lmax, 1, adapt_eps=1e-3)
    v = mixed_sign(o, dim)
    o.set_values(v)
    u.set_cells(v)
    ref = o.convection_upwind_nonlinear()
    conv = sa.make_convection_upwind()
    got = conv(u).storage()
    assert np.array_equal(bits(got), bits(ref))
    assert np.array_equal(bits(conv(u).storage()), bits(ref))


@pytest.mark.parametrize("dim,lmin,lmax", [(2, 3, 8), (3, 2, 5)])
def test_burgers_steps_match_oracle(gpu_ctx, dim, lmin, lmax):
    """Adapt + unp1 = u - dt * make_convection_upwind<Field>()(u), four steps,
    bit for bit (the time loop of demos/FiniteVolume/burgers.cpp, scalar)."""
    say("burgers steps", dim, lmin, lmax)
    o, mesh, u = pair(gpu_ctx, dim, lmin, lmax, 1)
    v = mixed_sign(o, dim)
    o.set_values(v)
    u.set_cells(v)
    ad = sa.make_MRAdapt(u)
    conv = sa.make_convection_upwind()
    f = sa.make_scalar_field("f", mesh)
    unp1 = sa.make_scalar_field("unp1", mesh)
    dt = 0.25 / (1 << lmax)
    for step in range(4):
        assert o.adapt(1e-3, 2.0) == ad(1e-3, 2.0)
        assert np.array_equal(o.cells(), mesh.cells()), step
        slots = o.cells()[:, 4].astype(np.int64)
        o.set_values(o.values() - dt * o.convection_upwind_nonlinear()[slots])
        f.resize()
        unp1.resize()
        conv(u, out=f)
        sa.sub_scaled(u, dt, f, unp1)
        sa.swap(u, unp1)
        assert np.array_equal(bits(o.values()), bits(u.cells())), step


def rotating_velocity(cells, dim):
    from initial_conditions import centers
    c = centers(cells)
    comps = [-(c[:, 1] - 0.5), c[:, 0] - 0.5]
    if dim == 3:
        comps.append(0.3 * np.sin(6 * c[:, 0]) - 0.1)
    return comps


@pytest.mark.parametrize("dim,lmin,lmax", [(2, 2, 7), (2, 3, 8), (3, 2, 5)])
def test_convection_velocity_field_matches_oracle(gpu_ctx, dim, lmin, lmax):
    """make_convection_upwind<Field>(velocity_field)(u) with a rotating velocity
    field (components with Dirichlet values) on adapted meshes."""
    say("convection velocity field", dim, lmin, lmax)
    o, mesh, u = pair(gpu_ctx, dim, lmin, lmax, 1, adapt_eps=1e-3)
    comps = rotating_velocity(o.cells(), dim)
    bcs = [0.25, -0.5, 0.125][:dim]
    vel = []
    for d in range(dim):
        f = sa.make_scalar_field(f"v{d}", mesh)
        f.set_cells(comps[d])
        sa.make_bc_dirichlet(f, bcs[d])
        vel.append(f)
    ref = o.convection_upwind_heterogeneous([o.ghost_updated(c, b) for c, b in zip(comps, bcs)])
    conv = sa.make_convection_upwind(vel)
    assert np.array_equal(bits(conv(u).storage()), bits(ref))
    # the velocity components' ghosts were updated in place, as the reference does
    for d in range(dim):
        assert np.array_equal(bits(vel[d].storage()), bits(o.ghost_updated(comps[d], bcs[d])))
    assert np.array_equal(bits(conv(u).storage()), bits(ref))


def two_discs(cells):
    """demos/FiniteVolume/scalar_burgers_2d.cpp:20-50: +1 in the disc of radius
    0.1 at (0.5, 0.5), -1 in the one at (0.2, 0.2), 0 elsewhere."""
    from initial_conditions import centers
    c = centers(cells)
    u = np.where((c[:, 0] - 0.5) ** 2 + (c[:, 1] - 0.5) ** 2 <= 0.01, 1.0, 0.0)
    return np.where((c[:, 0] - 0.2) ** 2 + (c[:, 1] - 0.2) ** 2 <= 0.01, -1.0, u)


@pytest.mark.parametrize("lmin,lmax", [(3, 8), (4, 9)])
def test_scalar_burgers_steps_match_oracle(gpu_ctx, lmin, lmax):
    """The scalar_burgers_2d time loop: adapt, update_ghost_mr, unp1 = u - dt *
    upwind_scalar_burgers(k, u), swap; five steps, bit for bit."""
    say("scalar burgers steps", lmin, lmax)
    o = OracleSim(2, lmin, lmax, pred_s=1)
    o.set_poison(False)
    o.set_values(two_discs(o.cells()))
    o.set_dirichlet(0.0)
    mesh = sa.make_mesh(gpu_ctx, sa.mesh_config(2, lmin, lmax, 1))
    u = sa.make_scalar_field("u", mesh)
    u.set_cells(two_discs(mesh.cells()))
    sa.make_bc_dirichlet(u, 0.0)
    unp1 = sa.make_scalar_field("unp1", mesh)
    ad = sa.make_MRAdapt(u)
    k = np.array([np.sqrt(2.0) / 2.0, np.sqrt(2.0) / 2.0])
    dt = 0.05 / (1 << lmax)
    for step in range(5):
        assert o.adapt(2e-4, 1.0) == ad(2e-4, 1.0)
        o.upwind_scalar_burgers_step(k, dt)
        sa.update_ghost_mr(u)
        unp1.resize()
        sa.upwind_scalar_burgers(k, dt, u, unp1)
        sa.swap(u, unp1)
        assert np.array_equal(o.cells(), mesh.cells()), step
        assert np.array_equal(bits(o.values()), bits(u.cells())), step


@pytest.mark.parametrize("lmin,lmax", [(5, 5), (2, 7)])
def test_hand_loop_matches_oracle(gpu_ctx, lmin, lmax):
    """The README's for_each_interval upwind loop."""
    say("hand loop", lmin, lmax)
    o, mesh, u = pair(gpu_ctx, 2, lmin, lmax, 1, ic="seeded", adapt_eps=1e-3 if lmin < lmax else None, seed=2)
    dt = 0.5 * 2.0 ** -lmax
    unp1 = sa.make_scalar_field("unp1", mesh)
    sa.update_ghost_mr(u)
    sa.upwind_hand_loop(dt, u, unp1)
    o.upwind_hand_loop(dt)
    assert np.array_equal(bits(unp1.cells()), bits(o.values()))


@pytest.mark.parametrize("dim,lmin,lmax", [(2, 7, 7), (2, 3, 8), (3, 4, 4), (3, 2, 5)])
def test_every_upwind_launch_shape_matches_oracle(gpu_ctx, dim, lmin, lmax):
    """All upwind launch shapes (cells per thread, XCD order, wave tiles, column
    walks, non-temporal stores) give the oracle's values bit for bit, on
    uniform meshes (long runs, tall stacks) and adapted ones (short runs)."""
    say("upwind launch shapes", dim, lmin, lmax)
    o, mesh, u = pair(gpu_ctx, dim, lmin, lmax, 1, adapt_eps=1e-3 if lmin < lmax else None)
    a = np.array([-0.8, 0.6, 0.3][:dim])
    dt = 0.4 / (1 << lmax)
    sa.update_ghost_mr(u)
    o.upwind_step(a, dt)
    ref = bits(o.values())
    try:
        for variant in range(sa.upwind_variant_count()):
            unp1 = sa.make_scalar_field("unp1", mesh)
            sa.set_upwind_variant(variant)
            sa.upwind(a, dt, u, unp1)
            assert np.array_equal(ref, bits(unp1.cells())), f"variant {variant}"
    finally:
        sa.set_upwind_variant(-1)


@pytest.mark.parametrize("dim,lmin,lmax", [(2, 2, 7), (3, 2, 5)])
def test_several_fields_adapt_like_the_oracle(gpu_ctx, dim, lmin, lmax):
    """make_MRAdapt(u, v) (mr/adapt.hpp:138-146, 391-397): the details of both
    fields drive one mesh, and both fields are transferred to every new mesh;
    u is advected, v rides along.  Cells and both fields bit for bit against
    the oracle, at every step."""
    say("test_several_fields_adapt_like_the_oracle", dim, lmin, lmax)
    fu = disc_2d if dim == 2 else sphere_3d
    fv = (lambda c: seeded_2d(c, 4)) if dim == 2 else (lambda c: sphere_3d(c, center=(0.3, 0.6, 0.4), radius=0.12))
    eps = 1e-4 if dim == 2 else 1e-3
    a = np.array([1.0, 0.5, 0.25][:dim])
    dt = 0.4 / 3 * 2.0 ** -lmax
    o = OracleSim(dim, lmin, lmax, pred_s=1)
    o.set_poison(False)
    o.set_values(fu(o.cells()))
    o.set_dirichlet(0.0)
    o.add_field(fv(o.cells()))
    mesh = sa.make_mesh(gpu_ctx, sa.mesh_config(dim, lmin, lmax, 1))
    u = sa.make_scalar_field("u", mesh)
    v = sa.make_scalar_field("v", mesh)
    u.set_cells(fu(mesh.cells()))
    v.set_cells(fv(mesh.cells()))
    sa.make_bc_dirichlet(u, 0.0)
    sa.make_bc_dirichlet(v, 0.0)
    unp1 = sa.make_scalar_field("unp1", mesh)
    adapt = sa.make_MRAdapt(u, v)
    for step in range(4):
        assert adapt(eps, 2.0) == o.adapt(eps, 2.0), step
        assert np.array_equal(o.cells(), mesh.cells()), step
        assert np.array_equal(bits(o.values()), bits(u.cells())), step
        assert np.array_equal(bits(o.extra_values(0)), bits(v.cells())), step
        sa.update_ghost_mr(u)
        unp1.resize()
        sa.upwind(a, dt, u, unp1)
        sa.swap(u, unp1)
        o.upwind_step(a, dt)
        assert np.array_equal(bits(o.values()), bits(u.cells())), step


@pytest.mark.parametrize("dim,lmin,lmax,nf", [(2, 2, 7, 1), (2, 3, 8, 2), (3, 2, 5, 1)])
def test_relative_detail_adapts_like_the_oracle(gpu_ctx, dim, lmin, lmax, nf):
    """MRadaptation(mra_config().relative_detail(true)) (mr/adapt.hpp:319-322,
    mr/rel_detail.hpp:74-112): each field's details divided by its max |value|
    over the leaves; scaled fields (maxima 3.7 and 0.6) so that the inverse is
    not a power of two.  Cells and values bit for bit against the oracle over
    three adapted steps; the allocation-free tag entry point refuses it."""
    say("test_relative_detail_adapts_like_the_oracle", dim, lmin, lmax, nf)
    fu = (lambda c: 3.7 * disc_2d(c)) if dim == 2 else (lambda c: 3.7 * sphere_3d(c))
    fv = lambda c: 0.6 * seeded_2d(c, 4)
    eps = 1e-4 if dim == 2 else 1e-3
    a = np.array([1.0, 0.5, 0.25][:dim])
    dt = 0.4 / 3 * 2.0 ** -lmax
    o = OracleSim(dim, lmin, lmax, pred_s=1)
    o.set_poison(False)
    o.set_values(fu(o.cells()))
    o.set_dirichlet(0.0)
    o.set_relative_detail(True)
    mesh = sa.make_mesh(gpu_ctx, sa.mesh_config(dim, lmin, lmax, 1))
    u = sa.make_scalar_field("u", mesh)
    u.set_cells(fu(mesh.cells()))
    sa.make_bc_dirichlet(u, 0.0)
    fields = [u]
    if nf == 2:
        o.add_field(fv(o.cells()))
        v = sa.make_scalar_field("v", mesh)
        v.set_cells(fv(mesh.cells()))
        sa.make_bc_dirichlet(v, 0.0)
        fields.append(v)
    unp1 = sa.make_scalar_field("unp1", mesh)
    adapt = sa.make_MRAdapt(*fields)
    for step in range(3):
        assert adapt(eps, 2.0, relative_detail=True) == o.adapt(eps, 2.0), step
        assert np.array_equal(o.cells(), mesh.cells()), step
        assert np.array_equal(bits(o.values()), bits(u.cells())), step
        sa.update_ghost_mr(u)
        unp1.resize()
        sa.upwind(a, dt, u, unp1)
        sa.swap(u, unp1)
        o.upwind_step(a, dt)
    import torch
    tags = torch.zeros(mesh.nb_cells(sa.REFERENCE), dtype=torch.uint8, device="cuda")
    with pytest.raises(sa.SamuraiError, match="relative detail"):
        adapt.tags_into_noalloc(0, eps, 2.0, tags.data_ptr())


@pytest.mark.parametrize("dim,lmin,lmax", [(2, 2, 7), (3, 2, 5)])
def test_refine_boundary_adapts_like_the_oracle(gpu_ctx, dim, lmin, lmax):
    """--refine-boundary (keep_boundary_refined, mr/adapt.hpp:245-272,
    341-346): the finest boundary leaves are kept; cells and values bit for
    bit against the oracle over three adapted steps."""
    say("test_refine_boundary_adapts_like_the_oracle", dim, lmin, lmax)
    fu = disc_2d if dim == 2 else sphere_3d
    eps = 1e-3
    a = np.array([1.0, 0.5, 0.25][:dim])
    dt = 0.4 / 3 * 2.0 ** -lmax
    o = OracleSim(dim, lmin, lmax, pred_s=1)
    o.set_poison(False)
    o.set_values(fu(o.cells()))
    o.set_dirichlet(0.0)
    o.set_refine_boundary(True)
    mesh = sa.make_mesh(gpu_ctx, sa.mesh_config(dim, lmin, lmax, 1))
    u = sa.make_scalar_field("u", mesh)
    u.set_cells(fu(mesh.cells()))
    sa.make_bc_dirichlet(u, 0.0)
    unp1 = sa.make_scalar_field("unp1", mesh)
    adapt = sa.make_MRAdapt(u)
    for step in range(3):
        assert adapt(eps, 2.0, refine_boundary=True) == o.adapt(eps, 2.0), step
        assert np.array_equal(o.cells(), mesh.cells()), step
        assert np.array_equal(bits(o.values()), bits(u.cells())), step
        sa.update_ghost_mr(u)
        unp1.resize()
        sa.upwind(a, dt, u, unp1)
        sa.swap(u, unp1)
        o.upwind_step(a, dt)
    c = mesh.cells()
    assert np.any((c[:, 0] == lmax) & (c[:, 1] == 0))   # the boundary stayed at the finest level
