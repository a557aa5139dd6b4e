"""Each benchmarked workload at its own size, through the C ABI, against the
oracle: cells (with their storage indices) and FP64 cell values bit for bit.

- 2D advection L2..8, box [0.4, 0.6]^2, a = (1, 1), eps 1e-4, regularity 2,
  dt = 0.5 * 2^-8, 50 steps: compared with the oracle after every step.
- 2D advection L4..12 (seeded 3 boxes + 2 Gaussians, a at pi/5): the first
  adaptation from the full level-12 mesh and three steps, compared with
  digests of the oracle's output (tests/golden/config_digests.json; the
  oracle needs minutes here); and the upwind kernel on the uniform level-12
  mesh against the textbook donor-cell update computed in numpy.
- 3D advection L2..8 (sphere r = 0.15, a = (1, 1, 1), eps 1e-3): the first
  adaptation and two steps against digests, and L2..6 against the oracle
  after every step.
- 256 batched instances (max level U{7..10}): two batched steps against the
  same instances advanced by the plain entry points, and eight of them
  against the oracle.
- The 32 frozen adjoint instances (L2..10): the adapted meshes and the
  forward solution after 100 steps against digests of the oracle's.
"""
import hashlib
import json
import math
import os
import sys

import numpy as np
import pytest

from initial_conditions import batch_field, batch_instance, box, seeded_2d, sphere_3d
from oracle_lib import OracleSim

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "config_digests.json")
REG = 2.0


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def bits(a):
    return np.ascontiguousarray(a).view(np.uint64)


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def record(mesh, u):
    cells = mesh.cells().astype(np.int64)
    return {"leaves": int(len(cells)), "cells_sha256": sha(cells), "values_sha256": sha(u.cells().astype(np.float64))}


def digests():
    with open(GOLDEN) as f:
        return json.load(f)


def advance(sa, u, unp1, a, dt):
    sa.update_ghost_mr(u)
    unp1.resize()
    sa.upwind(a, dt, u, unp1)
    sa.swap(u, unp1)


def test_advection_2d_box_L2_8_50_steps(gpu_ctx):
    import samurai_amd as sa
    say("advection 2D L2..8 box, 50 steps")
    lmin, lmax, eps = 2, 8, 1e-4
    a, dt = np.array([1.0, 1.0]), 0.5 * 2.0 ** -lmax
    o = OracleSim(2, lmin, lmax, pred_s=1)
    o.set_poison(False)
    o.set_values(box(o.cells(), 2))
    o.set_dirichlet(0.0)
    mesh = sa.make_mesh(gpu_ctx, sa.mesh_config(2, lmin, lmax, 1))
    u = sa.make_scalar_field("u", mesh)
    u.set_cells(box(mesh.cells(), 2))
    sa.make_bc_dirichlet(u, 0.0)
    unp1 = sa.make_scalar_field("unp1", mesh)
    adapt = sa.make_MRAdapt(u)
    for step in range(50):
        assert adapt(eps, REG) == o.adapt(eps, REG), step
        assert np.array_equal(o.cells(), mesh.cells()), f"cells after the adaptation of step {step}"
        advance(sa, u, unp1, a, dt)
        o.upwind_step(a, dt)
        assert np.array_equal(bits(o.values()), bits(u.cells())), f"values after step {step}"
    assert mesh.nb_cells() > 1000


def donor_cell_2d(v, a, dt, level):
    """unp1 = u - dt * ((((-F_L) + F_R) + (-F_D)) + F_U) / dx on a uniform mesh
    with Dirichlet-0 ghosts (g = 2 * 0 - u_in), F(a, l, r) = (0.5 a)(l + r) +
    (0.5 |a|)(l - r): the textbook first-order upwind update written in the
    reference's operation order (stencil_field.hpp:37-53, 92-97); v[y, x]."""
    p = np.zeros((v.shape[0] + 2, v.shape[1] + 2))
    p[1:-1, 1:-1] = v
    p[0, 1:-1], p[-1, 1:-1] = -v[0, :], -v[-1, :]
    p[1:-1, 0], p[1:-1, -1] = -v[:, 0], -v[:, -1]
    c = p[1:-1, 1:-1]

    def flux(ak, l, r):
        return (0.5 * ak) * (l + r) + (0.5 * abs(ak)) * (l - r)

    fl = flux(a[0], p[1:-1, :-2], c)
    fr = flux(a[0], c, p[1:-1, 2:])
    fd = flux(a[1], p[:-2, 1:-1], c)
    fu = flux(a[1], c, p[2:, 1:-1])
    s = (((-fl) + fr) + (-fd)) + fu
    return c - dt * (s * float(1 << level))


def test_uniform_L12_upwind_matches_textbook(gpu_ctx):
    """The default upwind kernel (column walks over pairs of 64-column strips)
    and a sample of the other launch shapes on the full uniform level-12 mesh
    (16,777,216 leaves)."""
    import samurai_amd as sa
    say("uniform L12 upwind")
    level = 12
    mesh = sa.make_mesh(gpu_ctx, sa.mesh_config(2, level, level, 1))
    cells = mesh.cells()
    assert len(cells) == 1 << (2 * level)
    u = sa.make_scalar_field("u", mesh)
    v = seeded_2d(cells, 0)
    u.set_cells(v)
    sa.make_bc_dirichlet(u, 0.0)
    unp1 = sa.make_scalar_field("unp1", mesh)
    a = np.array([math.cos(math.pi / 5), math.sin(math.pi / 5)])
    dt = 0.5 * 2.0 ** -level
    grid = np.zeros((1 << level, 1 << level))
    grid[cells[:, 2], cells[:, 1]] = v
    want = donor_cell_2d(grid, a, dt, level)[cells[:, 2], cells[:, 1]]
    sa.update_ghost_mr(u)
    nvar = sa.upwind_variant_count()
    try:
        for variant in [-1, 0, nvar // 3, nvar // 2, nvar - 1]:
            sa.set_upwind_variant(variant)
            unp1.set_cells(np.full(len(cells), np.nan))
            sa.upwind(a, dt, u, unp1)
            assert np.array_equal(bits(unp1.cells()), bits(want)), f"variant {variant}"
    finally:
        sa.set_upwind_variant(-1)


def run_digest_workload(sa, gpu_ctx, name, ic):
    d = digests()[name]
    mesh = sa.make_mesh(gpu_ctx, sa.mesh_config(d["dim"], d["min_level"], d["max_level"], 1))
    u = sa.make_scalar_field("u", mesh)
    v0 = ic(mesh.cells())
    assert sha(v0) == d["initial_values_sha256"], "initial field differs from the one the digests were made from"
    u.set_cells(v0)
    sa.make_bc_dirichlet(u, 0.0)
    unp1 = sa.make_scalar_field("unp1", mesh)
    adapt = sa.make_MRAdapt(u)
    a = np.array(d["velocity"])
    for k, want in enumerate(d["steps"]):
        calls = adapt(d["eps"], d["regularity"])
        got = record(mesh, u)
        say(f"{name}: step {k}, {got['leaves']} leaves")
        assert calls == want["harten_calls"], k
        assert got == want["adapted"], f"step {k}: after the adaptation"
        advance(sa, u, unp1, a, d["dt"])
        assert record(mesh, u) == want["stepped"], f"step {k}: after the upwind step"


def test_advection_2d_L4_12_matches_oracle_digests(gpu_ctx):
    import samurai_amd as sa
    run_digest_workload(sa, gpu_ctx, "advection_2d_L4_12_seed0", lambda c: seeded_2d(c, 0))


def test_advection_3d_L2_8_matches_oracle_digests(gpu_ctx):
    import samurai_amd as sa
    run_digest_workload(sa, gpu_ctx, "advection_3d_L2_8_sphere", lambda c: sphere_3d(c))


def test_advection_3d_L2_6_steps_match_oracle(gpu_ctx):
    import samurai_amd as sa
    say("advection 3D L2..6 sphere")
    lmin, lmax, eps = 2, 6, 1e-3
    a, dt = np.array([1.0, 1.0, 1.0]), 0.4 * 2.0 ** -lmax / 3
    o = OracleSim(3, lmin, lmax, pred_s=1)
    o.set_poison(False)
    o.set_values(sphere_3d(o.cells()))
    o.set_dirichlet(0.0)
    mesh = sa.make_mesh(gpu_ctx, sa.mesh_config(3, lmin, lmax, 1))
    u = sa.make_scalar_field("u", mesh)
    u.set_cells(sphere_3d(mesh.cells()))
    sa.make_bc_dirichlet(u, 0.0)
    unp1 = sa.make_scalar_field("unp1", mesh)
    adapt = sa.make_MRAdapt(u)
    for step in range(4):
        assert adapt(eps, REG) == o.adapt(eps, REG), step
        assert np.array_equal(o.cells(), mesh.cells()), step
        advance(sa, u, unp1, a, dt)
        o.upwind_step(a, dt)
        assert np.array_equal(bits(o.values()), bits(u.cells())), step


def make_batch_set(sa, ctx, n, levels, seed0):
    out = []
    for k in range(n):
        seed = seed0 + k
        max_level, a, nb, ng = batch_instance(seed, levels)
        mesh = sa.make_mesh(ctx, sa.mesh_config(2, 2, max_level, 1))
        u = sa.make_scalar_field("u", mesh)
        u.set_cells(batch_field(mesh.cells(), seed, nb, ng))
        sa.make_bc_dirichlet(u, 0.0)
        out.append({"mesh": mesh, "u": u, "unp1": sa.make_scalar_field("unp1", mesh), "a": np.array(a),
                    "dt": 0.5 * 2.0 ** -max_level, "max_level": max_level, "seed": seed, "nb": nb, "ng": ng})
    return out


def test_batch_256_instances_match_plain_and_oracle(gpu_ctx):
    """The batched workload: 256 instances, max level U{7..10}, eps 1e-4."""
    import samurai_amd as sa
    say("batch of 256 instances")
    eps, n, levels, seed0 = 1e-4, 256, (7, 10), 1000
    bset = make_batch_set(sa, gpu_ctx, n, levels, seed0)
    pset = make_batch_set(sa, gpu_ctx, n, levels, seed0)
    batch = sa.Batch(gpu_ctx, [(x["u"], x["unp1"], x["a"], x["dt"]) for x in bset])
    batch.adapt(eps, REG)
    adapts = [sa.make_MRAdapt(x["u"]) for x in pset]
    for ad in adapts:
        ad(eps, REG)
    # eight instances, two of each max level, on the oracle from the same start
    sample = []
    for lv in range(levels[0], levels[1] + 1):
        sample += [k for k, x in enumerate(bset) if x["max_level"] == lv][:2]
    orc = {}
    for k in sample:
        x = bset[k]
        o = OracleSim(2, 2, x["max_level"], pred_s=1)
        o.set_poison(False)
        o.set_values(batch_field(o.cells(), x["seed"], x["nb"], x["ng"]))
        o.set_dirichlet(0.0)
        o.adapt(eps, REG)
        orc[k] = o
    say(f"batch: oracle sample {sample}")

    def compare(what):
        for k, (p, b) in enumerate(zip(pset, bset)):
            assert np.array_equal(p["mesh"].cells(), b["mesh"].cells()), f"{what}: instance {k} cells"
            assert np.array_equal(bits(p["u"].storage()), bits(b["u"].storage())), f"{what}: instance {k} storage"
        for k, o in orc.items():
            b = bset[k]
            assert np.array_equal(o.cells(), b["mesh"].cells()), f"{what}: instance {k} cells vs oracle"
            assert np.array_equal(bits(o.values()), bits(b["u"].cells())), f"{what}: instance {k} values vs oracle"

    gpu_ctx.synchronize()
    compare("initial adaptation")
    for step in range(2):
        batch.step(eps, REG)
        for x, ad in zip(pset, adapts):
            ad(eps, REG)
            advance(sa, x["u"], x["unp1"], x["a"], x["dt"])
        for k, o in orc.items():
            o.adapt(eps, REG)
            o.upwind_step(bset[k]["a"], bset[k]["dt"])
        gpu_ctx.synchronize()
        compare(f"step {step}")
        say(f"batch: step {step} ok ({batch.nb_cells()} leaves)")


def test_batch_tag_change_flags_of_instances_sharing_a_wave(gpu_ctx):
    """Many small instances of one depth: their tag-change segments are packed
    back to back in one stage, so most waves hold cells of several instances,
    each with its own mesh-change flag.  Every instance must adapt as it does
    alone, at every step."""
    import samurai_amd as sa
    say("batch: small instances sharing waves")
    eps, n = 1e-4, 96

    def make():
        out = []
        for k in range(n):
            mesh = sa.make_mesh(gpu_ctx, sa.mesh_config(2, 2, 4, 1))
            u = sa.make_scalar_field("u", mesh)
            lo = 0.15 + 0.5 * ((k * 37) % 97) / 97.0
            u.set_cells(box(mesh.cells(), 2, lo, lo + 0.125))
            sa.make_bc_dirichlet(u, 0.0)
            th = 2.0 * math.pi * k / n
            out.append({"mesh": mesh, "u": u, "unp1": sa.make_scalar_field("unp1", mesh),
                        "a": np.array([math.cos(th), math.sin(th)]), "dt": 0.5 * 2.0 ** -4})
        return out

    bset, pset = make(), make()
    batch = sa.Batch(gpu_ctx, [(x["u"], x["unp1"], x["a"], x["dt"]) for x in bset])
    batch.set_tuning(256, 1, 1)
    adapts = [sa.make_MRAdapt(x["u"]) for x in pset]
    for step in range(8):
        batch.step(eps, REG)
        for x, ad in zip(pset, adapts):
            ad(eps, REG)
            advance(sa, x["u"], x["unp1"], x["a"], x["dt"])
        gpu_ctx.synchronize()
        for k, (p, b) in enumerate(zip(pset, bset)):
            assert np.array_equal(p["mesh"].cells(), b["mesh"].cells()), f"step {step}: instance {k} cells"
            assert np.array_equal(bits(p["u"].storage()), bits(b["u"].storage())), f"step {step}: instance {k}"


def test_adjoint_32_instances_forward_matches_oracle_digests(gpu_ctx):
    """The adjoint workload's meshes (adapted once by the batched form) and
    its forward pass of 100 steps with checkpoints every 10, per instance."""
    import samurai_amd as sa
    import torch
    d = digests()["adjoint_forward_32x_L2_10"]
    n, lmax = d["instances"], d["max_level"]
    say(f"adjoint: {n} instances, {d['steps']} steps")
    inst = []
    for k in range(n):
        mesh = sa.make_mesh(gpu_ctx, sa.mesh_config(2, d["min_level"], lmax, 1))
        u = sa.make_scalar_field("u", mesh)
        u.set_cells(seeded_2d(mesh.cells(), d["seed0"] + k))
        sa.make_bc_dirichlet(u, 0.0)
        inst.append({"mesh": mesh, "u": u, "unp1": sa.make_scalar_field("unp1", mesh)})
    a = np.array(d["velocity"])
    dt = 0.5 * 2.0 ** -lmax
    b = sa.Batch(gpu_ctx, [(x["u"], x["unp1"], a, dt) for x in inst])
    b.adapt(d["eps"], d["regularity"])
    del b
    for k, x in enumerate(inst):
        assert record(x["mesh"], x["u"]) == d["per_instance"][k]["adapted"], f"instance {k}: adapted mesh"
    prob = sa.AdjointProblem(gpu_ctx, [x["mesh"] for x in inst], [dt] * n, bc=[0.0] * n, steps=d["steps"],
                             checkpoint_every=10)
    prob.set_velocity(np.array([a] * n))
    u0 = torch.tensor(np.concatenate([x["u"].cells() for x in inst]), dtype=torch.float64, device="cuda")
    uT = torch.empty_like(u0)
    torch.cuda.synchronize()
    prob.forward_ptr(u0.data_ptr(), uT.data_ptr())
    got = uT.cpu().numpy()
    off = prob.offsets()
    for k in range(n):
        assert sha(got[off[k]:off[k + 1]]) == d["per_instance"][k]["forward"]["values_sha256"], f"instance {k}: forward"
