"""Restart files (io/restart.hpp layout) on the CPU: the reference's round
trips (tests/test_restart.cpp), the LevelCellArray arrays written for known
cell lists, derived by hand from the reference's construction rules
(level_cell_array.hpp:465-522), and the file structure.  Meshes are host-only;
fields need the GPU and are in test_gpu_restart.py."""
import os
import re
import subprocess

import numpy as np
import pytest

import samurai_amd as sa
from oracle_lib import OracleSim
from initial_conditions import disc_2d, sphere_3d

H5DUMP = "/opt/conda/bin/h5dump"

pytestmark = pytest.mark.skipif(not os.path.exists(sa.IO_LIB_PATH), reason="restart I/O library not built (no HDF5)")

def cells_of(cl):
    """CellList-like {level: {(y, z): [(s, e), ...]}} -> rows (level, x, y, z) in for_each_cell order."""
    rows = []
    for level in sorted(cl):
        for (y, z) in sorted(cl[level], key=lambda k: (k[1], k[0])):
            for s, e in cl[level][(y, z)]:
                rows += [(level, x, y, z) for x in range(s, e)]
    return np.array(rows, dtype=np.int32)


def read(h5, dset, nfields=1):
    """A dataset's values, parsed from h5dump's text output: a list of
    numbers, or of tuples of `nfields` numbers for a compound type."""
    out = subprocess.run([H5DUMP, "-d", dset, "-m", "%.17g", h5], check=True, capture_output=True, text=True).stdout
    body = out[out.index("DATA {") + len("DATA {"):]
    body = re.sub(r"\(\d+(,\d+)*\):", " ", body)
    vals = [float(v) if ("." in v or "e" in v) else int(v) for v in re.findall(r"-?\d+(?:\.\d*)?(?:e[-+]?\d+)?", body)]
    if nfields == 1:
        return vals
    return [tuple(vals[i:i + nfields]) for i in range(0, len(vals), nfields)]


# tests/test_restart.cpp:51-84: the CellList of each dimension
CL_2D = {1: {(0, 0): [(2, 8)], (1, 0): [(10, 15)]}, 4: {(6, 0): [(0, 1)]}}
CL_3D = {1: {(0, 0): [(2, 8)], (1, 0): [(10, 15)]}, 4: {(6, 1): [(0, 1)]}}


@pytest.mark.parametrize("dim,cl", [(2, CL_2D), (3, CL_3D)])
def test_cell_array_round_trip(tmp_path, dim, cl):
    cells = cells_of(cl)
    sa.dump_cells(tmp_path / "mesh", dim, cells)
    d, back = sa.load_cells(tmp_path / "mesh")
    assert d == dim
    assert np.array_equal(back, cells)


@pytest.mark.skipif(not os.path.exists(H5DUMP), reason="h5dump not available")
def test_level_cell_array_arrays_2d(tmp_path):
    """Level 1 has rows y = 0 ([2, 8)) and y = 1 ([10, 15)): one y-interval
    [0, 2) whose index maps y to row positions 0, 1 (index 0), offsets
    [0, 1, 2]; the x-intervals number the cells in for_each_cell order
    (CellArray's update_index, cell_array.hpp:484-493): 0 - 2 and 6 - 10.  Level 4 has row y = 6: y-interval [6, 7) with index -6 (row
    position 0), offsets [0, 1]."""
    sa.dump_cells(tmp_path / "mesh", 2, cells_of(CL_2D))
    h5 = str(tmp_path / "mesh.h5")
    x1 = read(h5, "/mesh/level/1/dim/0/intervals/data", 4)
    y1 = read(h5, "/mesh/level/1/dim/1/intervals/data", 4)
    o1 = read(h5, "/mesh/level/1/dim/1/offsets/data")
    assert x1 == [(2, 8, 1, -2), (10, 15, 1, -4)]
    assert y1 == [(0, 2, 1, 0)]
    assert o1 == [0, 1, 2]
    y4 = read(h5, "/mesh/level/4/dim/1/intervals/data", 4)
    assert y4 == [(6, 7, 1, -6)]
    assert read(h5, "/mesh/level/4/dim/1/offsets/data") == [0, 1]
    assert read(h5, "/mesh/level/1/dim/0/intervals/partition") == [0, 2]
    assert read(h5, "/n_process") == [1]
    assert read(h5, "/mesh/dim") == [2]
    assert read(h5, "/mesh/min_level") == [1]
    assert read(h5, "/mesh/max_level") == [4]
    assert read(h5, "/mesh/origin_point") == [0.0, 0.0]
    assert read(h5, "/mesh/scaling_factor") == [1.0]


@pytest.mark.skipif(not os.path.exists(H5DUMP), reason="h5dump not available")
def test_level_cell_array_arrays_3d(tmp_path):
    """3D level 1: plane z = 0 holds rows y = 0, 1 (one y-interval [0, 2),
    index 0), so one z-interval [0, 1) with index 0 and plane offsets [0, 1].
    Level 4: plane z = 1, row y = 6: z-interval [1, 2) index -1, y-interval
    [6, 7) index -6."""
    sa.dump_cells(tmp_path / "mesh", 3, cells_of(CL_3D))
    h5 = str(tmp_path / "mesh.h5")
    z1 = read(h5, "/mesh/level/1/dim/2/intervals/data", 4)
    y1 = read(h5, "/mesh/level/1/dim/1/intervals/data", 4)
    assert z1 == [(0, 1, 1, 0)]
    assert y1 == [(0, 2, 1, 0)]
    assert read(h5, "/mesh/level/1/dim/2/offsets/data") == [0, 1]
    assert read(h5, "/mesh/level/1/dim/1/offsets/data") == [0, 1, 2]
    z4 = read(h5, "/mesh/level/4/dim/2/intervals/data", 4)
    y4 = read(h5, "/mesh/level/4/dim/1/intervals/data", 4)
    assert z4 == [(1, 2, 1, -1)]
    assert y4 == [(6, 7, 1, -6)]


@pytest.mark.skipif(not os.path.exists(H5DUMP), reason="h5dump not available")
def test_interval_index_is_the_storage_offset(tmp_path, host_ctx):
    """The x-intervals of an MRMesh's leaves carry the storage offset of their
    run (value of x at index + x), as the reference's mesh[cells] does after
    renumbering (mesh.hpp:893-911)."""
    m = sa.make_mesh(host_ctx, sa.mesh_config(2, 2, 5, 1))
    sa.dump(tmp_path / "mesh", m)
    cells = m.cells()
    xi = read(str(tmp_path / "mesh.h5"), "/mesh/level/5/dim/0/intervals/data", 4)
    first = {(int(c[2]), int(c[1])): int(c[4]) for c in cells}
    for i in xi:
        row = None
        for (y, x), st in first.items():
            if x == i[0] and st - x == i[3]:
                row = y
                break
        assert row is not None, i


def adapted(ctx, dim, lmin, lmax):
    """A graduated multi-level mesh: the oracle's adaptation of a disc
    (sphere) on the box mesh, rebuilt by the product from its leaves."""
    o = OracleSim(dim, lmin, lmax, pred_s=1)
    c = o.cells()
    o.set_values(disc_2d(c) if dim == 2 else sphere_3d(c))
    o.set_dirichlet(0.0)
    o.adapt(1e-3, 1.0)
    return sa.make_mesh_from_cells(ctx, sa.mesh_config(dim, lmin, lmax, 1), o.cells())


@pytest.mark.parametrize("dim,lmin,lmax", [(2, 2, 5), (2, 3, 7), (3, 2, 4)])
def test_mesh_round_trip(tmp_path, host_ctx, dim, lmin, lmax):
    """restart_mrmesh (tests/test_restart.cpp:96-103) on box and adapted
    meshes: the loaded mesh has the same leaves, ids and storage numbering."""
    cfg = sa.mesh_config(dim, lmin, lmax, 1)
    for m in (sa.make_mesh(host_ctx, cfg), adapted(host_ctx, dim, lmin, lmax)):
        sa.dump(tmp_path / "mesh", m)
        m2 = sa.load_mesh(tmp_path / "mesh", host_ctx, sa.mesh_config(dim, 1, 2, 1))
        assert (m2.min_level, m2.max_level) == (lmin, lmax)
        for mid in (sa.CELLS, sa.CELLS_AND_GHOSTS, sa.PROJ_CELLS, sa.REFERENCE):
            assert np.array_equal(m.cells(mid), m2.cells(mid)), mid


def test_load_errors(tmp_path, host_ctx):
    sa.dump_cells(tmp_path / "mesh", 2, cells_of(CL_2D))
    with pytest.raises(sa.SamuraiError, match="dimension"):
        sa.load_mesh(tmp_path / "mesh", host_ctx, sa.mesh_config(3, 1, 4, 1))
    with pytest.raises(sa.SamuraiError, match="cannot open"):
        sa.load_cells(tmp_path / "absent")


def test_io_abi_exports_every_declared_symbol():
    import ctypes
    import re
    header = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "samurai_amd_io.h")
    names = re.findall(r"\b(smr_\w+)\s*\(", open(header).read())
    assert names
    sa.io_lib()
    L = ctypes.CDLL(sa.IO_LIB_PATH)
    for n in names:
        assert hasattr(L, n), n


def test_load_into_replaces_the_mesh_in_place(tmp_path, host_ctx):
    """samurai::load(file, mesh) into an empty mesh made with other levels
    (the reference demo's make_empty_mesh + load, demos/FiniteVolume/
    advection_2d.cpp:102-113): the mesh object takes the file's leaves and
    levels; a file of another dimension leaves it unchanged."""
    src = adapted(host_ctx, 2, 3, 7)
    sa.dump(tmp_path / "r", src)
    m = sa.make_mesh_from_cells(host_ctx, sa.mesh_config(2, 1, 2, 1), np.zeros((0, 4), dtype=np.int32))
    assert m.nb_cells() == 0
    sa.load_into(tmp_path / "r", m)
    assert (m.min_level, m.max_level) == (3, 7)
    for mid in (sa.CELLS, sa.CELLS_AND_GHOSTS, sa.PROJ_CELLS, sa.UNION_CELLS, sa.REFERENCE):
        assert np.array_equal(src.cells(mid), m.cells(mid)), mid
    sa.dump_cells(tmp_path / "c3", 3, cells_of(CL_3D))
    before = m.cells(sa.REFERENCE)
    with pytest.raises(sa.SamuraiError, match="dimension"):
        sa.load_into(tmp_path / "c3", m)
    assert np.array_equal(before, m.cells(sa.REFERENCE))


def test_assign_cells_matches_a_new_mesh(host_ctx):
    src = adapted(host_ctx, 2, 2, 6)
    m = sa.make_mesh(host_ctx, sa.mesh_config(2, 2, 6, 1))
    m.assign_cells(src.cells())
    for mid in (sa.CELLS, sa.CELLS_AND_GHOSTS, sa.PROJ_CELLS, sa.UNION_CELLS, sa.REFERENCE):
        assert np.array_equal(src.cells(mid), m.cells(mid)), mid


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("seed", range(6))
def test_random_cell_arrays_round_trip(tmp_path, dim, seed):
    """Random sparse cell lists on several levels, negative coordinates
    included (ghost-like cells), in for_each_cell order: load(dump(ca)) == ca,
    and the arrays keep the LevelCellArray invariants (offsets ascending from
    0 to the number of x-intervals; x-interval index = cells before the run -
    its start, in for_each_cell order)."""
    rng = np.random.default_rng(seed)
    rows = []
    for level in sorted(rng.choice(np.arange(1, 9), size=3, replace=False)):
        n = 1 << level
        k = int(rng.integers(1, 60))
        pts = np.unique(np.stack([rng.integers(-2, n + 2, k), rng.integers(-2, n + 2, k),
                                  rng.integers(-2, n + 2, k) if dim == 3 else np.zeros(k, dtype=np.int64)], axis=1), axis=0)
        order = np.lexsort((pts[:, 0], pts[:, 1], pts[:, 2]))  # z, then y, then x
        rows += [(level, int(p[0]), int(p[1]), int(p[2])) for p in pts[order]]
    cells = np.array(rows, dtype=np.int32)
    sa.dump_cells(tmp_path / "r", dim, cells)
    d, back = sa.load_cells(tmp_path / "r")
    assert d == dim and np.array_equal(back, cells)
    if not os.path.exists(H5DUMP):
        return
    h5 = str(tmp_path / "r.h5")
    count = 0
    for level in sorted(set(cells[:, 0])):
        xi = read(h5, f"/mesh/level/{level}/dim/0/intervals/data", 4)
        off = read(h5, f"/mesh/level/{level}/dim/1/offsets/data")
        assert off[0] == 0 and off[-1] == len(xi) and all(a <= b for a, b in zip(off, off[1:]))
        for s, e, step, index in xi:
            assert step == 1 and s < e and index == count - s
            count += e - s
    assert count == len(cells)


def test_mesh_swap_exchanges_contents(host_ctx):
    """smr_mesh_swap, the std::swap(mesh, new_mesh) of samurai::load and of
    `mesh = make_mesh(...)`: contents and configurations change places; meshes
    of two contexts are refused."""
    a = sa.make_mesh(host_ctx, sa.mesh_config(2, 2, 4, 1))
    b = adapted(host_ctx, 2, 3, 7)
    ca = {mid: a.cells(mid) for mid in (sa.CELLS, sa.REFERENCE)}
    cb = {mid: b.cells(mid) for mid in (sa.CELLS, sa.REFERENCE)}
    sa.swap_meshes(a, b)
    assert (a.min_level, a.max_level, b.min_level, b.max_level) == (3, 7, 2, 4)
    for mid in (sa.CELLS, sa.REFERENCE):
        assert np.array_equal(a.cells(mid), cb[mid]) and np.array_equal(b.cells(mid), ca[mid])
    other = sa.Context(-1)
    c = sa.make_mesh(other, sa.mesh_config(2, 2, 4, 1))
    with pytest.raises(sa.SamuraiError, match="different contexts"):
        sa.swap_meshes(a, c)
