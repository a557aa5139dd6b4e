"""The C++ drop-in layer (include/samurai_amd/samurai.hpp): the advection
demo written against it (samurai_amd/demos/advection_2d.cpp, the loop of
demos/FiniteVolume/advection_2d.cpp) reproduces the reference's own golden
fixtures of that demo: cells exactly, values within 1e-12."""
import os
import subprocess
import sys

import numpy as np
import pytest

import samurai_amd as sa

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
DEMO = os.path.join(ROOT, "samurai_amd", "demos", "advection_2d")


@pytest.mark.parametrize("s", [0, 1])
def test_cpp_demo_reproduces_reference_fixtures(tmp_path, s):
    print("cpp demo, prediction order", s, flush=True)
    assert os.path.exists(DEMO), "build the library first (python samurai_amd/build.py)"
    prefix = str(tmp_path / f"adv{s}")
    r = subprocess.run([DEMO, "--prediction-order", str(s), "--out", prefix, "--path", str(tmp_path)], capture_output=True, text=True,
                       timeout=120)
    sys.stdout.write(r.stdout)
    sys.stderr.write(r.stderr)
    assert r.returncode == 0
    assert "steps 21 " in r.stdout
    for suffix, name in [("_init.txt", f"advection_2d_pred{s}_init.npz"), ("_final.txt", f"advection_2d_pred{s}_t0.01.npz")]:
        got = np.loadtxt(prefix + suffix)
        fx = np.load(os.path.join(GOLDEN, name))
        assert got.shape[0] == fx["level"].size, name
        assert np.array_equal(got[:, 0].astype(np.int64), fx["level"])
        assert np.array_equal(got[:, 1].astype(np.int64), fx["i"])
        assert np.array_equal(got[:, 2].astype(np.int64), fx["j"])
        assert np.max(np.abs(got[:, 3] - fx["u"])) <= 1e-12, name


@pytest.mark.skipif(not os.path.exists(sa.IO_LIB_PATH), reason="I/O library not built (no HDF5)")
@pytest.mark.parametrize("s", [0, 1])
def test_cpp_demo_writes_the_reference_files_and_restarts(tmp_path, s):
    """The demo's samurai::save output (demos/FiniteVolume/advection_2d.cpp:
    47-59, 124, 151) is the reference's own regression files: same points,
    connectivity and XDMF text, values within 1e-12.  Restarted from its
    restart file of the initial state (--restart-file, :105-113), it ends on
    the same final file."""
    from test_save import check_against_reference, fixture_cells
    name = "test_finite_volume_advection_2d_finite-volume-advection-2d-0.01"
    stem = f"{name}_pred_{s}"
    for restart in (False, True):
        print("cpp demo files, prediction order", s, "restart", restart, flush=True)
        out = tmp_path / ("restarted" if restart else "fresh")
        out.mkdir()
        args = [DEMO, "--prediction-order", str(s), "--path", str(out), "--filename", name]
        if restart:
            args += ["--restart-file", str(tmp_path / "fresh" / f"{stem}_restart_init")]
        r = subprocess.run(args, capture_output=True, text=True, timeout=120)
        sys.stdout.write(r.stdout)
        sys.stderr.write(r.stderr)
        assert r.returncode == 0
        assert "steps 21 " in r.stdout
        for suffix, fx in (("_init", f"advection_2d_pred{s}_init.npz"), ("", f"advection_2d_pred{s}_t0.01.npz")):
            got = check_against_reference(out, stem + suffix)
            _, u_ref = fixture_cells(fx)
            assert np.max(np.abs(got - u_ref)) <= 1e-12, (restart, fx)
        assert os.path.exists(out / f"{stem}_restart.h5")


def test_cpp_scalar_burgers_demo_matches_oracle(tmp_path):
    """The scalar_burgers_2d loop written against the C++ drop-in layer
    (samurai_amd/demos/scalar_burgers_2d.cpp, the loop of
    demos/FiniteVolume/scalar_burgers_2d.cpp) on levels 3..7 for a few steps:
    its final leaves and values equal the oracle's run of the same loop bit for
    bit."""
    import sys as _sys
    _sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_lib import OracleSim
    from test_gpu_parity import two_discs
    demo = os.path.join(ROOT, "samurai_amd", "demos", "scalar_burgers_2d")
    assert os.path.exists(demo), "build the library first (python samurai_amd/build.py)"
    lmin, lmax, Tf, cfl = 3, 7, 0.004, 0.05
    print("cpp scalar burgers demo", lmin, lmax, Tf, flush=True)
    r = subprocess.run([demo, "--min-level", str(lmin), "--max-level", str(lmax), "--Tf", str(Tf), "--cfl", str(cfl),
                        "--path", str(tmp_path), "--out", str(tmp_path / "b")], capture_output=True, text=True, timeout=120)
    sys.stdout.write(r.stdout)
    sys.stderr.write(r.stderr)
    assert r.returncode == 0
    got = np.loadtxt(tmp_path / "b_final.txt")

    o = OracleSim(2, lmin, lmax, pred_s=1)
    o.set_poison(False)
    o.set_values(two_discs(o.cells()))
    o.set_dirichlet(0.0)
    k = np.array([np.sqrt(2.0) / 2.0, np.sqrt(2.0) / 2.0])
    dt, t, n = cfl / (1 << lmax), 0.0, 0
    o.adapt(2e-4, 1.0)
    while t != Tf:
        o.adapt(2e-4, 1.0)
        t += dt
        if t > Tf:
            dt += Tf - t
            t = Tf
        o.upwind_scalar_burgers_step(k, dt)
        n += 1
    assert f"steps {n} " in r.stdout
    cells = o.cells()
    assert got.shape[0] == cells.shape[0]
    assert np.array_equal(got[:, 0].astype(np.int64), cells[:, 0])
    assert np.array_equal(got[:, 1].astype(np.int64), cells[:, 1])
    assert np.array_equal(got[:, 2].astype(np.int64), cells[:, 2])
    assert np.array_equal(got[:, 3].view(np.uint64), o.values().view(np.uint64))


README_DEMO = os.path.join(ROOT, "samurai_amd", "demos", "readme_loop_device")


def test_readme_loop_as_device_lambda_matches_the_hand_loop(gpu_ctx, tmp_path):
    """README.md:86-154 with its for_each_interval loop written as a device
    lambda over include/samurai_amd/device.hpp field views: 20 steps of
    MRadaptation(1e-4, 2), update_ghost_mr, the loop and swap, bit-identical
    to the same steps through smr_upwind_hand_loop (the built-in kernel of
    that loop, itself checked against the oracle)."""
    from initial_conditions import box
    assert os.path.exists(README_DEMO), "build the library first (python samurai_amd/build.py)"
    out = str(tmp_path / "readme.txt")
    r = subprocess.run([README_DEMO, "--steps", "20", "--out", out], capture_output=True, text=True, timeout=120)
    sys.stdout.write(r.stdout)
    sys.stderr.write(r.stderr)
    assert r.returncode == 0
    got = np.loadtxt(out)
    mesh = sa.make_mesh(gpu_ctx, sa.mesh_config(2, 2, 8, 1))
    u = sa.make_scalar_field("u", mesh)
    u.set_cells(box(mesh.cells(), 2))
    sa.make_bc_dirichlet(u, 0.0)
    unp1 = sa.make_scalar_field("u", mesh)
    adapt = sa.make_MRAdapt(u)
    dt = 0.5 * 2.0 ** -8
    for _ in range(20):
        adapt(1e-4, 2.0)
        sa.update_ghost_mr(u)
        unp1.resize()
        sa.upwind_hand_loop(dt, u, unp1)
        sa.swap(u, unp1)
    cells = mesh.cells()
    assert np.array_equal(got[:, :3].astype(np.int64), cells[:, :3])
    assert np.array_equal(got[:, 3].view(np.uint64), u.cells().view(np.uint64))


def test_device_lookup_of_an_absent_cell_raises_out_of_range():
    """u(level, i - 3, j) outside the reference mesh: for_each_interval raises
    std::out_of_range (level_cell_array.hpp:674-680)."""
    r = subprocess.run([README_DEMO, "--steps", "2", "--probe-missing"], capture_output=True, text=True, timeout=120)
    sys.stdout.write(r.stdout)
    assert r.returncode == 3
    assert "out_of_range" in r.stdout
