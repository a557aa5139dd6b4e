"""Adjoint of frozen-mesh upwind steps (update_ghost_mr + upwind + swap,
algorithm/update_ghost_mr.hpp:194-237, stencil_field.hpp:16-97): the forward
pass must reproduce the plain path bit for bit, and the gradients of
L = 0.5 * sum area * (u_T - target)^2 with respect to the initial cells and
the velocity must agree with central finite differences (h = 1e-6) of the
GPU forward and of the oracle to 1e-7 relative."""
import numpy as np
import pytest

from initial_conditions import batch_field, batch_instance
from oracle_lib import OracleSim

pytestmark = pytest.mark.gpu

EPS, REG, H = 1e-4, 2.0, 1e-6


def make_frozen(sa, ctx, seeds, levels):
    out = []
    for seed in seeds:
        max_level, a, nb, ng = batch_instance(seed, levels)
        mesh = sa.make_mesh(ctx, sa.mesh_config(2, 2, max_level, 1))
        u = sa.make_scalar_field("u", mesh)
        u.set_cells(batch_field(mesh.cells(), seed, nb, ng))
        sa.make_bc_dirichlet(u, 0.0)
        sa.make_MRAdapt(u)(EPS, REG)
        cells = mesh.cells()
        out.append({"mesh": mesh, "u": u, "a": np.array(a), "dt": 0.5 * 2.0 ** -max_level, "u0": u.cells(),
                    "cells": cells, "area": np.ldexp(1.0, -2 * cells[:, 0].astype(np.int64)), "max_level": max_level})
    return out


def problem(sa, ctx, inst, steps, every):
    p = sa.AdjointProblem(ctx, [x["mesh"] for x in inst], [x["dt"] for x in inst], bc=[0.0] * len(inst), steps=steps,
                          checkpoint_every=every)
    p.set_velocity(np.array([x["a"] for x in inst]))
    return p


def run_forward(torch, p, u0):
    u0 = torch.tensor(u0, dtype=torch.float64, device="cuda")
    uT = torch.empty_like(u0)
    torch.cuda.synchronize()
    p.forward_ptr(u0.data_ptr(), uT.data_ptr())
    return uT.cpu().numpy()


def loss(uT, area, target):
    r = uT - target
    return 0.5 * float(np.sum(area * r * r))


def setup(sa, ctx, seeds, levels, steps, every):
    inst = make_frozen(sa, ctx, seeds, levels)
    p = problem(sa, ctx, inst, steps, every)
    u0 = np.concatenate([x["u0"] for x in inst])
    area = np.concatenate([x["area"] for x in inst])
    rng = np.random.Generator(np.random.PCG64(7))
    target = u0 + 0.1 * rng.standard_normal(len(u0))
    return inst, p, u0, area, target


def gradients(torch, p, u0, area, target):
    uT = run_forward(torch, p, u0)
    gT = torch.tensor(area * (uT - target), dtype=torch.float64, device="cuda")
    g0 = torch.empty_like(gT)
    torch.cuda.synchronize()
    ga = p.backward_ptr(gT.data_ptr(), g0.data_ptr())
    return uT, g0.cpu().numpy(), ga


def test_adjoint_forward_matches_plain(gpu_ctx):
    import samurai_amd as sa
    import torch
    inst, p, u0, area, target = setup(sa, gpu_ctx, [61, 62, 63], (4, 6), steps=6, every=2)
    uT = run_forward(torch, p, u0)
    off = p.offsets()
    for k, x in enumerate(inst):
        u, unp1 = x["u"], sa.make_scalar_field("unp1", x["mesh"])
        for _ in range(6):
            sa.update_ghost_mr(u)
            unp1.resize()
            sa.upwind(x["a"], x["dt"], u, unp1)
            sa.swap(u, unp1)
        got = uT[off[k]:off[k + 1]]
        assert np.array_equal(got.view(np.uint64), u.cells().view(np.uint64)), f"instance {k}"


def test_adjoint_matches_gpu_central_fd(gpu_ctx):
    import samurai_amd as sa
    import torch
    inst, p, u0, area, target = setup(sa, gpu_ctx, [71, 72, 73, 74], (4, 6), steps=20, every=5)
    _, g0, ga = gradients(torch, p, u0, area, target)
    rng = np.random.Generator(np.random.PCG64(3))
    for trial in range(2):
        v = rng.standard_normal(len(u0))
        fd = (loss(run_forward(torch, p, u0 + H * v), area, target) - loss(run_forward(torch, p, u0 - H * v), area, target)) / (2 * H)
        ad = float(np.dot(g0, v))
        assert abs(fd - ad) <= 1e-7 * abs(ad), (trial, fd, ad)
    a0 = np.array([x["a"] for x in inst])
    for i, d in [(0, 0), (1, 1), (3, 0)]:
        ap, am = a0.copy(), a0.copy()
        ap[i, d] += H
        am[i, d] -= H
        p.set_velocity(ap)
        lp = loss(run_forward(torch, p, u0), area, target)
        p.set_velocity(am)
        lm = loss(run_forward(torch, p, u0), area, target)
        fd = (lp - lm) / (2 * H)
        assert abs(fd - ga[i, d]) <= 1e-7 * abs(ga[i, d]), (i, d, fd, ga[i, d])
    p.set_velocity(a0)


def test_adjoint_matches_oracle_fd(gpu_ctx):
    """Central differences of the loss computed entirely by the oracle."""
    import samurai_amd as sa
    import torch
    steps = 8
    inst, p, u0, area, target = setup(sa, gpu_ctx, [81, 82], (4, 5), steps=steps, every=4)
    _, g0, ga = gradients(torch, p, u0, area, target)
    off = p.offsets()

    def oracle_loss(u0v, avec):
        total = 0.0
        for k, x in enumerate(inst):
            o = OracleSim(2, 2, x["max_level"], pred_s=1, cells=x["cells"])
            o.set_values(u0v[off[k]:off[k + 1]])
            o.set_dirichlet(0.0)
            for _ in range(steps):
                o.upwind_step(avec[k], x["dt"])
            r = o.values() - target[off[k]:off[k + 1]]
            total += 0.5 * float(np.sum(area[off[k]:off[k + 1]] * r * r))
        return total

    a0 = np.array([x["a"] for x in inst])
    rng = np.random.Generator(np.random.PCG64(5))
    v = rng.standard_normal(len(u0))
    fd = (oracle_loss(u0 + H * v, a0) - oracle_loss(u0 - H * v, a0)) / (2 * H)
    ad = float(np.dot(g0, v))
    assert abs(fd - ad) <= 1e-7 * abs(ad), (fd, ad)
    for i, d in [(0, 1), (1, 0)]:
        ap, am = a0.copy(), a0.copy()
        ap[i, d] += H
        am[i, d] -= H
        fd = (oracle_loss(u0, ap) - oracle_loss(u0, am)) / (2 * H)
        assert abs(fd - ga[i, d]) <= 1e-7 * abs(ga[i, d]), (i, d, fd, ga[i, d])


def test_checkpoint_interval_does_not_change_gradients(gpu_ctx):
    import samurai_amd as sa
    import torch
    inst, p1, u0, area, target = setup(sa, gpu_ctx, [91, 92], (4, 6), steps=8, every=2)
    p2 = problem(sa, gpu_ctx, inst, 8, 8)
    _, g1, a1 = gradients(torch, p1, u0, area, target)
    _, g2, a2 = gradients(torch, p2, u0, area, target)
    assert np.array_equal(g1, g2)
    assert np.array_equal(a1, a2)


# (block, cells per thread, mode): the default is (256, 0, 3), one block per
# instance for runs of small stages and wave items for the others
@pytest.mark.parametrize("shape", [(128, 2, 1), (128, 2, 2), (256, 0, 1), (256, 0, 2), (512, 4, 3), (256, 1, 0)])
def test_adjoint_launch_modes_agree(gpu_ctx, shape):
    import samurai_amd as sa
    import torch
    inst, p, u0, area, target = setup(sa, gpu_ctx, [93, 94, 95], (4, 6), steps=6, every=3)
    uT0, g00, ga0 = gradients(torch, p, u0, area, target)
    p.set_tuning(*shape)
    for rep in range(2):  # modes 2 and 3: the second pass replays the captured graphs
        uT, g0, ga = gradients(torch, p, u0, area, target)
        assert np.array_equal(uT, uT0) and np.array_equal(g0, g00) and np.array_equal(ga, ga0), rep


def test_torch_autograd_binding(gpu_ctx):
    import samurai_amd as sa
    import torch
    from samurai_amd.torch_adjoint import advect
    inst, p, u0, area, target = setup(sa, gpu_ctx, [96, 97], (4, 5), steps=4, every=2)
    _, g_ref, ga_ref = gradients(torch, p, u0, area, target)
    u0_t = torch.tensor(u0, dtype=torch.float64, device="cuda", requires_grad=True)
    a_t = torch.tensor(np.array([x["a"] for x in inst]), dtype=torch.float64, requires_grad=True)
    area_t = torch.tensor(area, dtype=torch.float64, device="cuda")
    target_t = torch.tensor(target, dtype=torch.float64, device="cuda")
    uT = advect(u0_t, a_t, p)
    L = 0.5 * (area_t * (uT - target_t) ** 2).sum()
    L.backward()
    assert np.allclose(u0_t.grad.cpu().numpy(), g_ref, rtol=1e-12, atol=0)
    assert np.allclose(a_t.grad.numpy(), ga_ref[:, :2], rtol=1e-12, atol=0)
