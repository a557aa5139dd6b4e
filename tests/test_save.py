"""HDF5/XDMF output (io/hdf5.hpp) on the CPU.  The writer is pinned against the
reference's own regression files of demos/FiniteVolume/advection_2d.cpp
(tests/reference/finite_volume/*advection-2d*.h5 and .xdmf, kept as the cells
and values in tests/golden/advection_2d_pred*.npz and the points and
connectivity digests and XDMF text in tests/golden/advection_2d_save.json):
saving the same cells with the same values must give the same bytes.  The
hexahedron numbering and the by_level / by_mesh_id layouts are checked on
cases derived by hand."""
import hashlib
import json
import os
import subprocess
import xml.etree.ElementTree as ET

import numpy as np
import pytest

import samurai_amd as sa

H5DUMP = "/opt/conda/bin/h5dump"
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
STEM = "test_finite_volume_advection_2d_finite-volume-advection-2d-0.01_pred_{s}{suffix}"

pytestmark = [pytest.mark.skipif(not os.path.exists(sa.IO_LIB_PATH), reason="I/O library not built (no HDF5)"),
              pytest.mark.skipif(not os.path.exists(H5DUMP), reason="h5dump not available")]


def read(h5, dset, dtype, tmp):
    out = os.path.join(str(tmp), dset.strip("/").replace("/", "_").replace(" ", "_") + ".bin")
    subprocess.run([H5DUMP, "-d", dset, "-b", "LE", "-o", out, h5], check=True, stdout=subprocess.DEVNULL)
    return np.fromfile(out, dtype=dtype)


def fixture_cells(name):
    fx = np.load(os.path.join(GOLDEN, name))
    cells = np.stack([fx["level"], fx["i"], fx["j"], np.zeros_like(fx["i"])], axis=1).astype(np.int32)
    return cells, fx["u"]


def check_against_reference(tmp_path, stem):
    ref = json.load(open(os.path.join(GOLDEN, "advection_2d_save.json")))[stem]
    h5 = str(tmp_path / (stem + ".h5"))
    pts = read(h5, "/mesh/points", "<f8", tmp_path)
    con = read(h5, "/mesh/connectivity", "<u8", tmp_path)
    assert con.size // 4 == ref["n_cells"] and pts.size // 3 == ref["n_points"]
    assert hashlib.sha256(pts.tobytes()).hexdigest() == ref["points_sha256"]
    assert hashlib.sha256(con.tobytes()).hexdigest() == ref["connectivity_sha256"]
    assert open(tmp_path / (stem + ".xdmf")).read() == ref["xdmf"]
    return read(h5, "/mesh/fields/u", "<f8", tmp_path)


@pytest.mark.parametrize("s", [0, 1])
@pytest.mark.parametrize("when", ["_init", ""])
def test_cell_array_save_reproduces_reference_files(tmp_path, s, when):
    cells, u = fixture_cells(f"advection_2d_pred{s}{'_init' if when else '_t0.01'}.npz")
    stem = STEM.format(s=s, suffix=when)
    sa.save_cell_array(tmp_path, stem, 2, cells, {"u": u})
    assert np.array_equal(check_against_reference(tmp_path, stem).view(np.uint64), u.view(np.uint64))


def test_hexahedra_numbering(tmp_path):
    """Level 1 of the unit cube, 8 cells: the first cell's corners are points
    0..7 in the element order (io/hdf5.hpp:77-82); the second cell (x = 1)
    shares its face x = 1/2 (points 1, 2, 5, 6) and adds 8..11."""
    cells = np.array([(1, x, y, z) for z in range(2) for y in range(2) for x in range(2)], dtype=np.int32)
    sa.save_cell_array(tmp_path, "cube", 3, cells)
    h5 = str(tmp_path / "cube.h5")
    con = read(h5, "/mesh/connectivity", "<u8", tmp_path).reshape(-1, 8)
    pts = read(h5, "/mesh/points", "<f8", tmp_path).reshape(-1, 3)
    assert pts.shape == (27, 3)
    assert list(con[0]) == list(range(8))
    assert list(con[1]) == [1, 8, 9, 2, 5, 10, 11, 6]
    assert np.array_equal(pts[:8], 0.5 * np.array([(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0),
                                                   (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]))
    assert np.array_equal(pts[8:12], 0.5 * np.array([(2, 0, 0), (2, 1, 0), (2, 0, 1), (2, 1, 1)]))
    topo = ET.parse(tmp_path / "cube.xdmf").getroot().find("Domain/Grid/Topology")
    assert topo.get("TopologyType") == "Hexahedron" and topo.get("NumberOfElements") == "8"


def test_by_level_and_by_mesh_id_layouts(tmp_path):
    """Levels 3..4 of a cell array saved by_level: one collection grid per
    level from min_level - 2 = 1, grids with cells only where the level has
    some, and per level the same points and connectivity as a save of that
    level's cells alone.  by_mesh_id nests the one sub-mesh "cell_array"."""
    cells = np.array([(3, 0, 0, 0), (3, 1, 0, 0), (4, 4, 0, 0), (4, 5, 0, 0), (4, 4, 1, 0)], dtype=np.int32)
    u = np.arange(5, dtype=np.float64)
    sa.save_cell_array(tmp_path, "lv", 2, cells, {"u": u}, by_level=True)
    dom = ET.parse(tmp_path / "lv.xdmf").getroot().find("Domain")
    grids = dom.findall("Grid")
    assert [g.get("Name") for g in grids] == ["Level 1", "Level 2", "Level 3", "Level 4"]
    assert all(g.get("GridType") == "Collection" and g.get("CollectionType") == "Spatial" for g in grids)
    assert [len(g.findall("Grid")) for g in grids] == [0, 0, 1, 1]
    for level, sel in ((3, slice(0, 2)), (4, slice(2, 5))):
        sa.save_cell_array(tmp_path, f"only{level}", 2, cells[sel], {"u": u[sel]})
        for dset, dt in (("connectivity", "<u8"), ("points", "<f8")):
            a = read(str(tmp_path / "lv.h5"), f"/level/{level}/mesh/{dset}", dt, tmp_path)
            b = read(str(tmp_path / f"only{level}.h5"), f"/mesh/{dset}", dt, tmp_path)
            assert np.array_equal(a, b), (level, dset)
        assert np.array_equal(read(str(tmp_path / "lv.h5"), f"/level/{level}/mesh/fields/u", "<f8", tmp_path), u[sel])

    sa.save_cell_array(tmp_path, "id", 2, cells, {"u": u}, by_mesh_id=True)
    g = ET.parse(tmp_path / "id.xdmf").getroot().find("Domain/Grid")
    assert g.get("Name") == "cell_array" and g.get("GridType") == "Collection"
    inner = g.find("Grid")
    assert inner.get("Name") == "cell_array"
    assert inner.find("Attribute/DataItem").text == "id.h5:/mesh/cell_array/fields/u"
    sa.save_cell_array(tmp_path, "plain", 2, cells, {"u": u})
    for dset, dt in (("connectivity", "<u8"), ("points", "<f8"), ("fields/u", "<f8")):
        assert np.array_equal(read(str(tmp_path / "id.h5"), f"/mesh/cell_array/{dset}", dt, tmp_path),
                              read(str(tmp_path / "plain.h5"), f"/mesh/{dset}", dt, tmp_path))


def test_mesh_save_without_fields(tmp_path, host_ctx):
    """An MRMesh (host context) saved by mesh id: every id of the mesh has its
    grid, with as many cells as the mesh holds for that id."""
    m = sa.make_mesh(host_ctx, sa.mesh_config(2, 2, 4, 1))
    sa.save(tmp_path / "out", "box", m, by_mesh_id=True)
    dom = ET.parse(tmp_path / "out" / "box.xdmf").getroot().find("Domain")
    names = ["cells", "cells and ghosts", "projection cells", "union cells", "reference"]
    ids = [sa.CELLS, sa.CELLS_AND_GHOSTS, sa.PROJ_CELLS, sa.UNION_CELLS, sa.REFERENCE]
    by_name = {g.get("Name"): g for g in dom.findall("Grid")}
    for name, mid in zip(names, ids):
        n = m.nb_cells(mid)
        if n == 0:
            assert name in by_name and by_name[name].find("Grid") is None
            continue
        topo = by_name[name].find("Grid/Topology")
        assert topo.get("NumberOfElements") == str(n), name


def test_save_errors(tmp_path):
    cells = np.array([(3, 1, 0, 0), (3, 0, 0, 0)], dtype=np.int32)
    with pytest.raises(sa.SamuraiError, match="for_each_cell order"):
        sa.save_cell_array(tmp_path, "bad", 2, cells)
    with pytest.raises(sa.SamuraiError, match="holds a '/'"):
        sa.save_cell_array(tmp_path, "bad", 2, cells[::-1], {"a/b": np.zeros(2)})


def test_io_abi_exports_save_symbols():
    import ctypes
    sa.io_lib()
    L = ctypes.CDLL(sa.IO_LIB_PATH)
    for n in ("smr_save", "smr_save_cell_array"):
        assert hasattr(L, n), n


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("seed", range(4))
def test_random_cell_arrays_points_and_connectivity(tmp_path, dim, seed):
    """Random sparse multi-level cell lists: every cell's corners, read back
    through the connectivity, are its exact corners in the element order; the
    points are exactly the distinct corners, numbered at first sight in
    for_each_cell order."""
    rng = np.random.default_rng(100 + seed)
    rows = []
    for level in sorted(rng.choice(np.arange(1, 7), size=3, replace=False)):
        n = 1 << level
        k = int(rng.integers(1, 40))
        pts = np.unique(np.stack([rng.integers(0, n, k), rng.integers(0, n, k),
                                  rng.integers(0, n, k) if dim == 3 else np.zeros(k, dtype=np.int64)], axis=1), axis=0)
        order = np.lexsort((pts[:, 0], pts[:, 1], pts[:, 2]))
        rows += [(level, int(p[0]), int(p[1]), int(p[2])) for p in pts[order]]
    cells = np.array(rows, dtype=np.int32)
    sa.save_cell_array(tmp_path, "rnd", dim, cells)
    h5 = str(tmp_path / "rnd.h5")
    npc = 1 << dim
    con = read(h5, "/mesh/connectivity", "<u8", tmp_path).reshape(-1, npc).astype(np.int64)
    pts = read(h5, "/mesh/points", "<f8", tmp_path).reshape(-1, 3)
    elem = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)][:npc]
    seen = {}
    for c, row in zip(cells, con):
        h = np.ldexp(1.0, -int(c[0]))
        for e, pid in zip(elem, row):
            corner = tuple((c[1 + d] + e[d]) * h if d < dim else 0.0 for d in range(3))
            assert tuple(pts[pid]) == corner
            if corner not in seen:
                assert pid == len(seen)  # numbered at first sight
                seen[corner] = pid
    assert len(seen) == len(pts)
