"""Benchmark: MR-upwind cell-updates/s of the adaptive 2D advection step on one
MI355X, with the HBM roofline of the upwind kernel and the oracle timed on the
host as a CPU baseline.

Workload (the large 2D case): unit square, a = (cos pi/5, sin pi/5), MRMesh
min_level 4, max_level 12 (4096^2 finest), prediction radius 1, stencil
radius 1, Dirichlet 0, eps = 1e-4, regularity 2, dt = 0.5 * 2^-12, float64.
Initial field: seeded sum of 3 indicator boxes and 2 Gaussians
(tests/initial_conditions.py:seeded_2d, seed 0), adapted once before timing
like the reference demo does before its loop.

One step = MRadaptation + update_ghost_mr + (unp1 = u - dt*upwind(a, u)) + swap,
exactly the loop body of demos/FiniteVolume/advection_2d.cpp:129-153.  The
timed region includes the host mesh rebuilds and every transfer they need.

Also measured: the uniform level-12 mesh with the same field (pure stencil
roofline check) and the oracle on a bounded sample of the same workload.
"""
import argparse
import ctypes
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

HBM_PEAK_GBS = 8000.0

# HBM bytes per launch of the uniform level-12 upwind, from two rocprofv3 --pmc
# runs of this program (FETCH_SIZE, WRITE_SIZE), each scaled by the counter's
# reading for the copy kernel of known size measured in the same run
# (profiles/upwind_walk_uniform_L12.md): FETCH 68,160 KiB x 2.0 + WRITE
# 136,259 KiB x 1.0 for k_upwind_walk_pair<16, 8> on walks of 10 rows.  None
# when not measured for the current kernel.
MEASURED_UPWIND_TRAFFIC_BYTES = 279121488


def progress(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


class HipEvents:
    """hipEvent timing on an explicit stream (the library's own stream)."""

    def __init__(self):
        self.hip = ctypes.CDLL("libamdhip64.so")
        self.hip.hipEventCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
        self.hip.hipEventRecord.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        self.hip.hipEventSynchronize.argtypes = [ctypes.c_void_p]
        self.hip.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), ctypes.c_void_p, ctypes.c_void_p]
        self.hip.hipEventDestroy.argtypes = [ctypes.c_void_p]

    def create(self):
        e = ctypes.c_void_p()
        if self.hip.hipEventCreate(ctypes.byref(e)) != 0:
            raise RuntimeError("hipEventCreate failed")
        return e

    def record(self, e, stream):
        if self.hip.hipEventRecord(e, ctypes.c_void_p(stream)) != 0:
            raise RuntimeError("hipEventRecord failed")

    def elapsed_ms(self, a, b):
        self.hip.hipEventSynchronize(b)
        ms = ctypes.c_float()
        if self.hip.hipEventElapsedTime(ctypes.byref(ms), a, b) != 0:
            raise RuntimeError("hipEventElapsedTime failed")
        return float(ms.value)

    def destroy(self, e):
        self.hip.hipEventDestroy(e)


def make_problem(sa, ctx, min_level, max_level, seed):
    from initial_conditions import seeded_2d
    mesh = sa.make_mesh(ctx, sa.mesh_config(2, min_level, max_level, 1))
    u = sa.make_scalar_field("u", mesh)
    u.set_cells(seeded_2d(mesh.cells(), seed))
    sa.make_bc_dirichlet(u, 0.0)
    unp1 = sa.make_scalar_field("unp1", mesh)
    return mesh, u, unp1


def _cpu_worker(job):
    """One instance on a CPU path for `budget` seconds, in a fresh process with
    one host thread: "interval" is the single-core interval-algebra CPU
    baseline (oracle/cpu_step.cpp: the reference's algorithm on interval
    sets), "dense" the dense-bitmap oracle.  Imports nothing that touches the
    GPU."""
    os.environ["SMR_HOST_THREADS"] = "1"
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    kind, cells, values, min_level, max_level, a, dt, eps, reg, adapt, budget = job
    t0 = time.perf_counter()
    if kind == "interval":
        from cpu_step_lib import CpuSim
        o = CpuSim(2, min_level, max_level, cells)
    else:
        from oracle_lib import OracleSim
        o = OracleSim(2, min_level, max_level, pred_s=1, cells=cells)
    o.set_values(values)
    o.set_dirichlet(0.0)
    setup = time.perf_counter() - t0
    steps, updates, spent = 0, 0, 0.0
    while steps == 0 or spent < budget:
        t = time.perf_counter()
        if adapt:
            o.adapt(eps, reg)
        n = o.nb_cells()
        o.upwind_step(a, dt)
        spent += time.perf_counter() - t
        steps += 1
        updates += n
    return updates, spent, steps, setup


def _run_workers(jobs):
    import multiprocessing as mp
    t0 = time.perf_counter()
    with mp.get_context("spawn").Pool(len(jobs)) as pool:
        res = pool.map(_cpu_worker, jobs)
    return res, time.perf_counter() - t0


def cpu_baseline(cells, values, a, dt, eps, reg, budget_s):
    """The single-core CPU baseline (interval algebra, the reference's
    algorithm) on the GPU run's final adapted mesh and field for about
    budget_s seconds, and one step of the dense-bitmap oracle beside it."""
    job = ("interval", cells, values, 4, 12, a, dt, eps, reg, True, budget_s)
    (res,), _ = _run_workers([job])
    updates, spent, steps, setup = res
    (dres,), _ = _run_workers([("dense",) + job[1:-1] + (0.0,)])
    return {"value": updates / spent, "unit": "cell-updates/s", "cores": 1, "kind": "port",
            "sample": f"{steps} steps (MRadaptation + update_ghost_mr + upwind) of the interval-algebra CPU path "
                      f"(oracle/cpu_step.cpp: the product's host interval algebra and operator bodies on one core, "
                      f"-O3 -ffp-contract=off, bit-identical to the oracle) on the GPU run's final adapted L4..12 mesh "
                      f"({len(cells)} leaves), {spent:.1f} s of CPU time, setup {setup:.1f} s untimed",
            "dense_oracle": {"value": dres[0] / dres[1], "unit": "cell-updates/s", "cores": 1,
                             "sample": f"{dres[2]} step(s) of the dense-bitmap oracle on the same mesh, {dres[1]:.1f} s"}}


def cpu_baseline_parallel(jobs, label):
    """The interval-algebra CPU path on several instances at once, one
    single-threaded process per instance (processes started fresh, not
    forked from this GPU process)."""
    workers = max(1, min(len(jobs), 16, os.cpu_count() or 1))
    jobs = [("interval",) + tuple(j) for j in jobs[:workers]]
    res, wall = _run_workers(jobs)
    updates = sum(r[0] for r in res)
    busy = max(r[1] for r in res)
    return {"value": updates / busy, "unit": "cell-updates/s", "cores": workers, "kind": "port",
            "sample": f"{label}: {workers} instances, one single-threaded interval-algebra CPU process each "
                      f"(oracle/cpu_step.cpp), {sum(r[2] for r in res)} steps in {busy:.1f} s of concurrent CPU time "
                      f"({wall:.1f} s wall with process start)"}


def make_batch_set(sa, ctx, n, levels, seed0):
    """n instances of the batched workload (tests/initial_conditions.py:batch_instance)."""
    from initial_conditions import batch_field, batch_instance
    out = []
    for k in range(n):
        seed = seed0 + k
        max_level, a, nb, ng = batch_instance(seed, levels)
        mesh = sa.make_mesh(ctx, sa.mesh_config(2, 2, max_level, 1))
        u = sa.make_scalar_field("u", mesh)
        u.set_cells(batch_field(mesh.cells(), seed, nb, ng))
        sa.make_bc_dirichlet(u, 0.0)
        unp1 = sa.make_scalar_field("unp1", mesh)
        out.append({"mesh": mesh, "u": u, "unp1": unp1, "a": np.array(a), "dt": 0.5 * 2.0 ** -max_level})
    return out


def bench_batched(sa, ctx, args, eps, reg):
    """The batched workload advanced by one Batch, and the same instances
    advanced one by one through the plain entry points (same start state)."""
    levels = (args.batch_min_max_level, args.batch_max_max_level)
    t_setup = time.perf_counter()
    bset = make_batch_set(sa, ctx, args.batch, levels, 1000 + args.seed)
    pset = make_batch_set(sa, ctx, args.batch, levels, 1000 + args.seed)
    for group in (bset, pset):
        b = sa.Batch(ctx, [(x["u"], x["unp1"], x["a"], x["dt"]) for x in group])
        b.adapt(eps, reg)  # initial adaptation (setup, like the demo before its loop)
        del b
    ctx.synchronize()
    progress(f"batched: {args.batch} instances set up and adapted ({time.perf_counter() - t_setup:.1f} s)")
    batch = sa.Batch(ctx, [(x["u"], x["unp1"], x["a"], x["dt"]) for x in bset])
    tune_file = sa.default_tuning_file()
    tuned = batch.load_tuning(tune_file)
    if not tuned or args.retune:
        batch.autotune(reps=3, cache_file=tune_file)
    setup = time.perf_counter() - t_setup
    progress(f"batched: tuning {batch.tuning()}")

    for _ in range(args.warmup):
        batch.step(eps, reg)
    ctx.synchronize()
    updates = 0
    t0 = time.perf_counter()
    for _ in range(args.batch_steps):
        batch.step(eps, reg)
        updates += batch.nb_cells()
    ctx.synchronize()
    bwall = time.perf_counter() - t0
    progress(f"batched: {args.batch_steps} steps in {bwall:.2f} s")
    # phase breakdown of one more step (the timers synchronize)
    sa.timers_reset()
    sa.timers_enable(True)
    batch.step(eps, reg)
    ctx.synchronize()
    sa.timers_enable(False)
    bphases = {k: {"ms": v["seconds"] * 1e3, "calls": v["calls"]} for k, v in sa.timers_report().items()}

    adapts = [sa.make_MRAdapt(x["u"]) for x in pset]

    def plain_step():
        n = 0
        for x, ad in zip(pset, adapts):
            ad(eps, reg)
            sa.update_ghost_mr(x["u"])
            x["unp1"].resize()
            sa.upwind(x["a"], x["dt"], x["u"], x["unp1"])
            sa.swap(x["u"], x["unp1"])
            n += x["mesh"].nb_cells()
        return n

    for _ in range(args.warmup):
        plain_step()
    ctx.synchronize()
    pupdates = 0
    t0 = time.perf_counter()
    for _ in range(args.batch_steps):
        pupdates += plain_step()
    ctx.synchronize()
    pwall = time.perf_counter() - t0
    progress(f"plain loop: {args.batch_steps} steps in {pwall:.2f} s")
    leaves = [x["mesh"].nb_cells() for x in bset]
    cpu = None
    if args.cpu_budget > 0:
        jobs = [(x["mesh"].cells(), x["u"].cells(), 2, x["mesh"].max_level, x["a"], x["dt"], eps, reg, True,
                 args.cpu_budget / 2) for x in bset]
        cpu = cpu_baseline_parallel(jobs, "MRadaptation + update_ghost_mr + upwind on the batched instances' final meshes")
        progress(f"batched cpu baseline: {cpu['value']:.3g} cell-updates/s on {cpu['cores']} cores")
    return {
        "workload": f"{args.batch} independent 2D advection instances, max_level U{{{levels[0]}..{levels[1]}}}, "
                    f"min_level 2, unit velocity at angle U[0,2pi), 1-5 seeded boxes/Gaussians, eps=1e-4, "
                    f"regularity 2, dt=0.5*dx_finest, Dirichlet 0",
        "value": updates / bwall, "unit": "cell-updates/s", "steps": args.batch_steps,
        "ms_per_step": bwall / args.batch_steps * 1e3,
        "plain_loop_value": pupdates / pwall, "plain_loop_ms_per_step": pwall / args.batch_steps * 1e3,
        "speedup_vs_plain_loop": (updates / bwall) / (pupdates / pwall),
        "leaves_total": int(sum(leaves)), "leaves_min": int(min(leaves)), "leaves_max": int(max(leaves)),
        "tuning": batch.tuning(), "tuning_from_cache": bool(tuned and not args.retune),
        "setup_s": setup, "cpu_baseline": cpu, "phases_one_step": bphases,
    }


def bench_3d(sa, ctx, args):
    """The 3D workload: unit cube, a = (1, 1, 1), L2..8, sphere r = 0.15 at the
    centre, eps 1e-3, regularity 2, dt = 0.4 * 2^-8 / 3, Dirichlet 0."""
    from initial_conditions import sphere_3d
    t0 = time.perf_counter()
    mesh = sa.make_mesh(ctx, sa.mesh_config(3, 2, 8, 1))
    u = sa.make_scalar_field("u", mesh)
    u.set_cells(sphere_3d(mesh.cells()))
    sa.make_bc_dirichlet(u, 0.0)
    unp1 = sa.make_scalar_field("unp1", mesh)
    adapt = sa.make_MRAdapt(u)
    adapt(1e-3, 2.0)
    setup = time.perf_counter() - t0
    a = np.array([1.0, 1.0, 1.0])
    dt = 0.4 * 2.0 ** -8 / 3

    def step():
        adapt(1e-3, 2.0)
        n = mesh.nb_cells()
        sa.update_ghost_mr(u)
        unp1.resize()
        sa.upwind(a, dt, u, unp1)
        sa.swap(u, unp1)
        return n

    for _ in range(args.warmup):
        step()
    ctx.synchronize()
    updates = 0
    t0 = time.perf_counter()
    for _ in range(args.steps_3d):
        updates += step()
    ctx.synchronize()
    wall = time.perf_counter() - t0
    progress(f"3D L2..8: {args.steps_3d} steps in {wall:.2f} s")
    return {"workload": "3D upwind advection, unit cube, a=(1,1,1), MRMesh L2..8, sphere r=0.15 at the centre, "
                        "eps=1e-3, regularity 2, dt=0.4*2^-8/3, Dirichlet 0",
            "value": updates / wall, "unit": "cell-updates/s", "steps": args.steps_3d,
            "ms_per_step": wall / args.steps_3d * 1e3, "mean_leaves": updates / args.steps_3d, "setup_s": setup}


def bench_convection(sa, ctx, mesh, u, a, reps):
    """The flux-based convection operators on the final adapted mesh: time per
    call (ghost update of the inputs + gather of the ordered contribution
    lists), after one untimed call that builds and uploads the lists."""
    from initial_conditions import centers
    c = centers(mesh.cells())
    vel = []
    for d, comp in enumerate([-(c[:, 1] - 0.5), c[:, 0] - 0.5]):
        v = sa.make_scalar_field(f"v{d}", mesh)
        v.set_cells(comp)
        sa.make_bc_dirichlet(v, 0.0)
        vel.append(v)
    f = sa.make_scalar_field("f", mesh)
    ops = {"constant_velocity": (sa.make_convection_upwind(a), a),
           "nonlinear_burgers": (sa.make_convection_upwind(), None),
           "velocity_field": (sa.make_convection_upwind(vel), None)}
    res = {}
    for name, (op, lists_of) in ops.items():
        op(u, out=f)
        ctx.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            op(u, out=f)
        ctx.synchronize()
        res[name] = {"ms_per_call": (time.perf_counter() - t0) / reps * 1e3,
                     "terms": int(len(sa.convection_terms(mesh, lists_of)[3]))}
    res["leaves"] = mesh.nb_cells()
    progress("convection operators: " + ", ".join(f"{k} {v['ms_per_call']:.3f} ms" for k, v in res.items() if isinstance(v, dict)))
    return res


def bench_weno5(sa, ctx, reps, level=10):
    """make_convection_weno5 on a uniform 2D mesh of 4^level cells built with
    max_stencil_size 6 (the WENO5 operator needs one level and three ghost
    layers): time per call (ghost update + gather).  The gather reads per
    term six 8-byte slots, a coefficient and a direction byte (57 B), and the
    six stencil values; list_bytes is the list traffic of one call."""
    from initial_conditions import centers
    mesh = sa.make_mesh(ctx, sa.mesh_config(2, level, level, 1, 3))
    c = centers(mesh.cells())
    u = sa.make_scalar_field("u", mesh)
    u.set_cells(np.sin(2 * np.pi * c[:, 0]) * np.cos(2 * np.pi * c[:, 1]))
    sa.make_bc_dirichlet(u, 0.0)
    f = sa.make_scalar_field("f", mesh)
    op = sa.make_convection_weno5([1.0, -0.5])
    op(u, out=f)
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        op(u, out=f)
    ctx.synchronize()
    ms = (time.perf_counter() - t0) / reps * 1e3
    n = mesh.nb_cells()
    terms = 2 * 2 * n  # two interfaces per cell and direction, each seen from both sides (boundary: one side)
    return {"workload": f"uniform 2D L{level} ({n} cells), stencil radius 3, v=(1,-0.5), Dirichlet 0",
            "ms_per_call": ms, "cells": n, "terms": terms, "list_bytes": 57 * terms + 8 * 2 * n,
            "cell_updates_per_s": n / (ms * 1e-3)}


def bench_adjoint(sa, ctx, args, eps, reg):
    """Gradients of L = 0.5 sum area (u_T - u_target)^2 after 100 frozen-mesh
    steps on 32 instances (L2..10, adapted once), w.r.t. u0 and a; u_target is
    the forward solution from a field of a perturbed seed."""
    import torch
    from initial_conditions import seeded_2d
    n, steps, every, lmax = args.adjoint, args.adjoint_steps, 10, 10
    theta = math.pi / 5
    a = np.array([[math.cos(theta), math.sin(theta)]] * n)
    t_setup = time.perf_counter()
    inst = []
    for k in range(n):
        mesh = sa.make_mesh(ctx, sa.mesh_config(2, 2, lmax, 1))
        u = sa.make_scalar_field("u", mesh)
        u.set_cells(seeded_2d(mesh.cells(), 2000 + k))
        sa.make_bc_dirichlet(u, 0.0)
        inst.append({"mesh": mesh, "u": u, "unp1": sa.make_scalar_field("unp1", mesh)})
    b = sa.Batch(ctx, [(x["u"], x["unp1"], a[0], 0.5 * 2.0 ** -lmax) for x in inst])
    b.adapt(eps, reg)  # adapt once at t=0, then the meshes stay frozen
    del b
    prob = sa.AdjointProblem(ctx, [x["mesh"] for x in inst], [0.5 * 2.0 ** -lmax] * n, bc=[0.0] * n, steps=steps,
                             checkpoint_every=every)
    prob.set_velocity(a)
    cells = [x["mesh"].cells() for x in inst]
    u0 = torch.tensor(np.concatenate([x["u"].cells() for x in inst]), dtype=torch.float64, device="cuda")
    area = torch.tensor(np.concatenate([np.ldexp(1.0, -2 * c[:, 0].astype(np.int64)) for c in cells]),
                        dtype=torch.float64, device="cuda")
    pert = torch.tensor(np.concatenate([seeded_2d(c, 3000 + k) for k, c in enumerate(cells)]), dtype=torch.float64,
                        device="cuda")
    target = torch.empty_like(u0)
    uT = torch.empty_like(u0)
    torch.cuda.synchronize()
    prob.forward_ptr(pert.data_ptr(), target.data_ptr())
    setup = time.perf_counter() - t_setup
    progress(f"adjoint: {n} frozen instances, {prob.nb_cells()} cells ({setup:.1f} s)")

    def loss_of(u):
        torch.cuda.synchronize()
        prob.forward_ptr(u.data_ptr(), uT.data_ptr())
        return 0.5 * float((area * (uT - target) ** 2).sum())

    g0 = torch.empty_like(u0)
    loss_of(u0)  # warm-up (programs built)
    gT = (area * (uT - target)).contiguous()
    torch.cuda.synchronize()
    prob.backward_ptr(gT.data_ptr(), g0.data_ptr())
    reps = args.adjoint_reps
    t0 = time.perf_counter()
    for _ in range(reps):
        prob.forward_ptr(u0.data_ptr(), uT.data_ptr())
    t_fwd = (time.perf_counter() - t0) / reps
    gT = (area * (uT - target)).contiguous()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        ga = prob.backward_ptr(gT.data_ptr(), g0.data_ptr())
    t_bwd = (time.perf_counter() - t0) / reps
    progress(f"adjoint: forward {t_fwd * 1e3:.1f} ms, backward {t_bwd * 1e3:.1f} ms")

    # central differences, h = 1e-6, along a random direction of u0 and along a[0, 0]
    h = 1e-6
    rng = np.random.Generator(np.random.PCG64(11))
    v = torch.tensor(rng.standard_normal(u0.numel()), dtype=torch.float64, device="cuda")
    fd_u = (loss_of(u0 + h * v) - loss_of(u0 - h * v)) / (2 * h)
    ad_u = float((g0 * v).sum())
    ap, am = a.copy(), a.copy()
    ap[0, 0] += h
    am[0, 0] -= h
    prob.set_velocity(ap)
    lp = loss_of(u0)
    prob.set_velocity(am)
    lm = loss_of(u0)
    prob.set_velocity(a)
    fd_a = (lp - lm) / (2 * h)
    ncells = prob.nb_cells()
    cpu = None
    if args.cpu_budget > 0:
        off = prob.offsets()
        u0h = u0.cpu().numpy()
        jobs = [(c, u0h[off[k]:off[k + 1]], 2, lmax, a[k], 0.5 * 2.0 ** -lmax, eps, reg, False, args.cpu_budget / 2)
                for k, c in enumerate(cells)]
        cpu = cpu_baseline_parallel(jobs, "frozen-mesh forward steps (update_ghost_mr + upwind) of the adjoint instances")
        progress(f"adjoint cpu baseline: {cpu['value']:.3g} cell-updates/s on {cpu['cores']} cores")
    return {
        "workload": f"{n} frozen 2D instances L2..{lmax} (seeded 3 boxes + 2 Gaussians, adapted once with eps=1e-4), "
                    f"{steps} upwind steps, a=(cos pi/5, sin pi/5), checkpoint every {every}, "
                    f"loss 0.5*sum area*(u_T-u_target)^2",
        "cells": int(ncells), "steps": steps,
        "forward_ms": t_fwd * 1e3, "backward_ms": t_bwd * 1e3,
        "backward_over_forward": t_bwd / t_fwd,
        "forward_cell_updates_per_s": ncells * steps / t_fwd,
        "backward_cell_updates_per_s": ncells * steps / t_bwd,
        "fd_u0_rel_err": abs(fd_u - ad_u) / abs(ad_u), "fd_a_rel_err": abs(fd_a - ga[0, 0]) / abs(ga[0, 0]),
        "fd_h": h, "tuning": "library default (256, 0, 3): blocks of 256, precomputed 64-cell wave items, runs of small stages as one block per instance, each checkpoint segment replayed as a HIP graph",
        "setup_s": setup, "cpu_baseline_forward": cpu,
    }


def main():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=20)
    p.add_argument("--warmup", type=int, default=2)
    p.add_argument("--seed", type=int, default=0)
    p.add_argument("--uniform-reps", type=int, default=20)
    p.add_argument("--roofline-warmup", type=float, default=0.3, help="seconds of untimed launches before the roofline timings")
    p.add_argument("--phase-steps", type=int, default=3, help="extra steps timed per phase (0 disables)")
    p.add_argument("--cpu-budget", type=float, default=20.0, help="seconds of oracle time (0 disables)")
    p.add_argument("--batch", type=int, default=256, help="instances of the batched workload (0 disables)")
    p.add_argument("--batch-steps", type=int, default=5)
    p.add_argument("--batch-min-max-level", type=int, default=7)
    p.add_argument("--batch-max-max-level", type=int, default=10)
    p.add_argument("--retune", action="store_true", help="rerun the batch autotuner even if a cached shape exists")
    p.add_argument("--steps-3d", type=int, default=10, help="timed steps of the 3D workload (0 disables)")
    p.add_argument("--adjoint", type=int, default=32, help="instances of the adjoint workload (0 disables)")
    p.add_argument("--adjoint-steps", type=int, default=100)
    p.add_argument("--adjoint-reps", type=int, default=2)
    p.add_argument("--convection", type=int, default=1, help="time the flux-based convection operators (0 disables)")
    args = p.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}): this build has no multi-GPU path; "
              f"independent instances would shard across GPUs without communication.", file=sys.stderr)
        sys.exit(2)

    import samurai_amd as sa
    ctx = sa.Context(0)
    stream = ctx.stream()
    ev = HipEvents()

    theta = math.pi / 5
    a = np.array([math.cos(theta), math.sin(theta)])
    dt = 0.5 * (1.0 / (1 << 12))
    eps, reg = 1e-4, 2.0

    # ---------------- adaptive L4..12 workload ----------------
    mesh, u, unp1 = make_problem(sa, ctx, 4, 12, args.seed)
    adapt = sa.make_MRAdapt(u)
    adapt(eps, reg)

    def step(times=None):
        adapt(eps, reg)
        n = mesh.nb_cells()
        unp1.resize()
        if times is not None:
            # the time step proper on the adapted mesh: update_ghost_mr, then
            # the upwind kernel, each bracketed by events on the library's stream
            e0, e1, e2 = times
            ev.record(e0, stream)
            sa.update_ghost_mr(u)
            ev.record(e1, stream)
            sa.upwind(a, dt, u, unp1)
            ev.record(e2, stream)
        else:
            sa.update_ghost_mr(u)
            sa.upwind(a, dt, u, unp1)
        sa.swap(u, unp1)
        return n

    for _ in range(args.warmup):
        step()
    ctx.synchronize()
    updates = 0
    t0 = time.perf_counter()
    for _ in range(args.steps):
        updates += step()
    ctx.synchronize()
    wall = time.perf_counter() - t0
    progress(f"adaptive L4..12: {args.steps} steps in {wall:.2f} s")
    # the device time of update_ghost_mr and upwind and their algorithmic
    # bytes, on extra steps after the timed ones (the byte counts are host set
    # algebra, kept out of the timed region)
    ghost_ms, kms, ghost_b, upw_b = [], [], [], []
    evs = (ev.create(), ev.create(), ev.create())
    for _ in range(max(3, args.steps // 4)):
        step(evs)
        ghost_ms.append(ev.elapsed_ms(evs[0], evs[1]))
        kms.append(ev.elapsed_ms(evs[1], evs[2]))
        ghost_b.append(mesh.ghost_update_bytes())
        upw_b.append(mesh.upwind_bytes())
    for e in evs:
        ev.destroy(e)
    upw_bytes = float(np.mean(upw_b))
    ghost_bytes = float(np.mean(ghost_b))
    upw_ms = float(np.mean(kms))
    ghost_ms = float(np.mean(ghost_ms))
    adaptive_gbs = upw_bytes / (upw_ms * 1e-3) / 1e9
    step_gbs = (ghost_bytes + upw_bytes) / ((ghost_ms + upw_ms) * 1e-3) / 1e9

    # phase breakdown under the reference's timer names (synchronizing, so
    # measured on extra steps outside the timed region)
    phases = None
    if args.phase_steps > 0:
        sa.timers_reset()
        sa.timers_enable(True)
        for _ in range(args.phase_steps):
            step()
        ctx.synchronize()
        sa.timers_enable(False)
        phases = {k: {"ms_per_step": v["seconds"] / args.phase_steps * 1e3, "calls": v["calls"]}
                  for k, v in sa.timers_report().items()}
    final_cells = mesh.nb_cells()
    convection = bench_convection(sa, ctx, mesh, u, a, 20) if args.convection else None

    # ---------------- uniform level-12 roofline check ----------------
    # Four (u, unp1) pairs on the uniform mesh, each ghost-updated once; the
    # timed region is back-to-back upwind launches cycling through the pairs,
    # so every launch reads a u last touched >= 1 GB of traffic earlier (no
    # reuse through the 256 MB last-level cache).
    from initial_conditions import seeded_2d
    umesh = sa.make_mesh(ctx, sa.mesh_config(2, 12, 12, 1))
    uvals = seeded_2d(umesh.cells(), args.seed)
    pairs = []
    for _ in range(4):
        uu = sa.make_scalar_field("u", umesh)
        uu.set_cells(uvals)
        sa.make_bc_dirichlet(uu, 0.0)
        sa.update_ghost_mr(uu)
        pairs.append((uu, sa.make_scalar_field("unp1", umesh)))
    # untimed warm-up launches for about 0.3 s (after the host-bound adaptive
    # run the GPU clocks are not at their sustained level yet), so the timed
    # region measures the sustained rate; the copy reference below gets the same
    t_w = time.perf_counter()
    while time.perf_counter() - t_w < args.roofline_warmup:
        for _ in range(10):
            for uu, vv in pairs:
                sa.upwind(a, dt, uu, vv)
        ctx.synchronize()
    e0, e1 = ev.create(), ev.create()
    ev.record(e0, stream)
    for _ in range(args.uniform_reps):
        for uu, vv in pairs:
            sa.upwind(a, dt, uu, vv)
    ev.record(e1, stream)
    u_ms = ev.elapsed_ms(e0, e1) / (args.uniform_reps * len(pairs))
    ubytes = umesh.upwind_bytes()
    uniform_gbs = ubytes / (u_ms * 1e-3) / 1e9
    uniform_cups = umesh.nb_cells() / (u_ms * 1e-3)
    del pairs

    # measured copy bandwidth: an elementwise kernel reading 134 MB and writing
    # 134 MB per launch (y = 1.0 * x), four buffer pairs as above
    import torch
    bufs = [(torch.ones(1 << 24, dtype=torch.float64, device="cuda"), torch.empty(1 << 24, dtype=torch.float64, device="cuda"))
            for _ in range(4)]
    t_w = time.perf_counter()
    while time.perf_counter() - t_w < args.roofline_warmup:
        for _ in range(10):
            for x, y in bufs:
                torch.mul(x, 1.0, out=y)
        torch.cuda.synchronize()
    c0, c1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    c0.record()
    for _ in range(args.uniform_reps):
        for x, y in bufs:
            torch.mul(x, 1.0, out=y)
    c1.record()
    c1.synchronize()
    copy_ms = c0.elapsed_time(c1) / (args.uniform_reps * len(bufs))
    copy_gbs = 2 * 8 * (1 << 24) / (copy_ms * 1e-3) / 1e9
    del bufs
    progress(f"uniform L12 upwind {u_ms * 1e3:.1f} us ({uniform_gbs:.0f} GB/s), copy {copy_gbs:.0f} GB/s")

    plain_3d = None
    if args.steps_3d > 0:
        plain_3d = bench_3d(sa, ctx, args)

    batched = None
    if args.batch > 0:
        batched = bench_batched(sa, ctx, args, eps, reg)

    adjoint = None
    if args.adjoint > 0:
        adjoint = bench_adjoint(sa, ctx, args, eps, reg)

    # last: its 4 M-term lists (a few hundred MB of host memory while they are
    # built) would otherwise slow the host work of the timings after it
    if convection is not None:
        convection["weno5"] = bench_weno5(sa, ctx, 20)
        progress(f"weno5 operator: {convection['weno5']['ms_per_call']:.3f} ms per call")

    cpu = None
    if args.cpu_budget > 0:
        cpu = cpu_baseline(mesh.cells(), u.cells(), a, dt, eps, reg, args.cpu_budget)

    out = {
        "metric": "MR-upwind cell-updates/s + %HBM peak; batched & adjoint vs plain, same GPU",
        "value": updates / wall,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "2D upwind advection, unit square, a=(cos pi/5, sin pi/5), MRMesh L4..12, "
                               "seeded 3 boxes + 2 Gaussians, eps=1e-4, regularity 2, dt=0.5*2^-12, "
                               "prediction radius 1, Dirichlet 0",
                   "seed": args.seed, "min_level": 4, "max_level": 12,
                   "mean_leaves": updates / args.steps, "final_leaves": final_cells},
        "roofline": {"bound": "hbm", "achieved": uniform_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": uniform_gbs / HBM_PEAK_GBS, "traffic": MEASURED_UPWIND_TRAFFIC_BYTES,
                     "kernel": "upwind (column-walk kernel k_upwind_walk_pair: two adjacent 64-column strips per wave, walks of 10 rows all fetched ahead, x neighbours by lane shifts, non-temporal loads and stores) on the uniform level-12 mesh (16,777,216 leaves)",
                     "bytes_per_launch": ubytes, "avg_launch_ms": u_ms, "launches_timed": args.uniform_reps * 4,
                     "untimed_warmup_s": args.roofline_warmup,
                     "cell_updates_per_s": uniform_cups, "measured_copy_GBs": copy_gbs,
                     "frac_of_measured_copy": uniform_gbs / copy_gbs},
        "roofline_adaptive": {"kernel": "update_ghost_mr + upwind of the timed adaptive steps (the time step on the "
                                        "adapted mesh; the adaptation's device passes are not included)",
                              "bytes_per_step": ghost_bytes + upw_bytes, "ghost_update_bytes": ghost_bytes,
                              "upwind_bytes": upw_bytes, "ghost_update_ms": ghost_ms, "upwind_ms": upw_ms,
                              "achieved": step_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": step_gbs / HBM_PEAK_GBS,
                              "upwind_achieved": adaptive_gbs, "upwind_frac": adaptive_gbs / HBM_PEAK_GBS},
        "cpu_baseline": cpu,
        "phases": phases,
        "plain_3d": plain_3d,
        "convection": convection,
        "batched": batched,
        "adjoint": adjoint,
    }
    print(json.dumps(out), flush=True)
    # every library object released while the HIP runtime is whole
    ctx.close()


if __name__ == "__main__":
    main()
