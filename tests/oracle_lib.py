"""ctypes binding of the oracle (oracle/liboracle.so) for the tests and the
CPU-baseline leg of bench.py.  The oracle is the checker, never the product."""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
# SMR_ORACLE_LIB selects another build of the oracle (the mutation test's
# libraries with a planted mistake)
LIB_PATH = os.environ.get("SMR_ORACLE_LIB") or os.path.join(ROOT, "oracle", "liboracle.so")

CELLS, CELLS_AND_GHOSTS, PROJ_CELLS, UNION_CELLS, REFERENCE = range(5)

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle not built: {LIB_PATH} (run __graft_entry__.build())")
        L = ctypes.CDLL(LIB_PATH)
        vp, i32, i64, dbl = ctypes.c_void_p, ctypes.c_int, ctypes.c_int64, ctypes.c_double
        P = ctypes.POINTER
        L.orc_last_error.restype = ctypes.c_char_p
        L.orc_create_box.restype = vp
        L.orc_create_box.argtypes = [i32] * 6
        L.orc_create_from_cells.restype = vp
        L.orc_create_from_cells.argtypes = [i32] * 6 + [P(ctypes.c_int32), i64]
        L.orc_free.argtypes = [vp]
        L.orc_set_poison.argtypes = [vp, i32]
        L.orc_nb_cells.restype = i64
        L.orc_nb_cells.argtypes = [vp, i32]
        L.orc_get_cells.argtypes = [vp, i32, P(i64)]
        L.orc_set_cells_values.argtypes = [vp, P(dbl)]
        L.orc_get_cells_values.argtypes = [vp, P(dbl)]
        L.orc_get_storage.argtypes = [vp, P(dbl)]
        L.orc_set_storage.argtypes = [vp, P(dbl)]
        L.orc_set_dirichlet.argtypes = [vp, dbl]
        L.orc_update_ghost_mr.argtypes = [vp]
        L.orc_adapt.argtypes = [vp, dbl, dbl]
        L.orc_harten.argtypes = [vp, i32, dbl, dbl]
        L.orc_tag_storage_size.restype = i64
        L.orc_tag_storage_size.argtypes = [vp]
        L.orc_get_last_tags.argtypes = [vp, P(ctypes.c_uint8), P(dbl)]
        L.orc_upwind_step.argtypes = [vp, P(dbl), dbl]
        L.orc_upwind_only.argtypes = [vp, P(dbl), dbl]
        L.orc_convection_upwind.argtypes = [vp, P(dbl), P(dbl)]
        L.orc_convection_upwind_nonlinear.argtypes = [vp, P(dbl)]
        L.orc_convection_weno5.argtypes = [vp, P(dbl), P(dbl)]
        L.orc_set_relative_detail.argtypes = [vp, i32]
        L.orc_set_refine_boundary.argtypes = [vp, i32]
        L.orc_weno5_flux.restype = dbl
        L.orc_weno5_flux.argtypes = [P(dbl)]
        L.orc_upwind_scalar_burgers_step.argtypes = [vp, P(dbl), dbl]
        L.orc_convection_upwind_heterogeneous.argtypes = [vp, P(P(dbl)), P(dbl)]
        L.orc_upwind_hand_loop.argtypes = [vp, dbl]
        L.orc_apply_level_op.argtypes = [vp, i32, i32]
        L.orc_apply_field_bc_direction.argtypes = [vp, i32, i32, i32]
        L.orc_add_field.argtypes = [vp]
        L.orc_set_extra_values.argtypes = [vp, i32, P(dbl)]
        L.orc_get_extra_values.argtypes = [vp, i32, P(dbl)]
        L.orc_get_extra_storage.argtypes = [vp, i32, P(dbl)]
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _check(rc):
    if rc < 0:
        raise RuntimeError(lib().orc_last_error().decode())
    return rc


class OracleSim:
    """One scalar field on an MR mesh of the unit box (oracle side)."""

    def __init__(self, dim, min_level, max_level, pred_s=1, radius=1, grad_width=1, cells=None):
        """A box mesh at max_level, or the mesh of the given leaves (rows level, x, y, z)."""
        self.dim = dim
        if cells is None:
            self.h = lib().orc_create_box(dim, min_level, max_level, pred_s, radius, grad_width)
        else:
            c = np.ascontiguousarray(np.asarray(cells)[:, :4], dtype=np.int32)
            self.h = lib().orc_create_from_cells(dim, min_level, max_level, pred_s, radius, grad_width,
                                                 c.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), c.shape[0])
        if not self.h:
            raise RuntimeError(lib().orc_last_error().decode())

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_free(self.h)
            self.h = None

    def set_poison(self, on):
        lib().orc_set_poison(self.h, int(on))

    def nb_cells(self, mesh_id=CELLS):
        return lib().orc_nb_cells(self.h, mesh_id)

    def cells(self, mesh_id=CELLS):
        """(n, 5) int64 rows (level, x, y, z, storage index) in traversal order."""
        n = self.nb_cells(mesh_id)
        out = np.zeros((n, 5), dtype=np.int64)
        _check(lib().orc_get_cells(self.h, mesh_id, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))))
        return out

    def set_values(self, v):
        v = np.ascontiguousarray(v, dtype=np.float64)
        assert v.size == self.nb_cells(CELLS)
        _check(lib().orc_set_cells_values(self.h, _dp(v)))

    def values(self):
        out = np.zeros(self.nb_cells(CELLS))
        _check(lib().orc_get_cells_values(self.h, _dp(out)))
        return out

    def storage(self):
        out = np.zeros(self.nb_cells(REFERENCE))
        lib().orc_get_storage(self.h, _dp(out))
        return out

    def set_storage(self, v):
        v = np.ascontiguousarray(v, dtype=np.float64)
        assert v.size == self.nb_cells(REFERENCE)
        lib().orc_set_storage(self.h, _dp(v))

    def add_field(self, values):
        """Another field on this mesh adapted together with the main one
        (make_MRAdapt(u, v...)); values on the leaves.  Returns its number."""
        k = _check(lib().orc_add_field(self.h))
        v = np.ascontiguousarray(values, dtype=np.float64)
        assert v.size == self.nb_cells(CELLS)
        _check(lib().orc_set_extra_values(self.h, k, _dp(v)))
        return k

    def extra_values(self, k):
        out = np.zeros(self.nb_cells(CELLS))
        _check(lib().orc_get_extra_values(self.h, k, _dp(out)))
        return out

    def extra_storage(self, k):
        out = np.zeros(self.nb_cells(REFERENCE))
        _check(lib().orc_get_extra_storage(self.h, k, _dp(out)))
        return out

    def set_dirichlet(self, value=0.0):
        self._bc = float(value)
        lib().orc_set_dirichlet(self.h, float(value))

    def update_ghost_mr(self):
        _check(lib().orc_update_ghost_mr(self.h))

    def adapt(self, eps, reg):
        return _check(lib().orc_adapt(self.h, float(eps), float(reg)))

    def harten(self, ite, eps, reg):
        return _check(lib().orc_harten(self.h, int(ite), float(eps), float(reg)))

    def last_tags(self):
        n = lib().orc_tag_storage_size(self.h)
        t = np.zeros(n, dtype=np.uint8)
        d = np.zeros(n)
        lib().orc_get_last_tags(self.h, t.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)), _dp(d))
        return t, d

    def upwind_step(self, a, dt):
        a = np.ascontiguousarray(a, dtype=np.float64)
        _check(lib().orc_upwind_step(self.h, _dp(a), float(dt)))

    def upwind_only(self, a, dt):
        a = np.ascontiguousarray(a, dtype=np.float64)
        _check(lib().orc_upwind_only(self.h, _dp(a), float(dt)))

    def upwind_scalar_burgers_step(self, k, dt):
        """update_ghost_mr, unp1 = u - dt * upwind_scalar_burgers(k, u), swap (2D)."""
        k = np.ascontiguousarray(k, dtype=np.float64)
        _check(lib().orc_upwind_scalar_burgers_step(self.h, _dp(k), float(dt)))

    def convection_upwind_nonlinear(self):
        """make_convection_upwind<Field>()(u) (scalar Burgers flux u*u) after update_ghost_mr(u): the output storage."""
        out = np.zeros(self.nb_cells(REFERENCE))
        _check(lib().orc_convection_upwind_nonlinear(self.h, _dp(out)))
        return out

    def ghost_updated(self, values, bc_value):
        """The storage of another field on this mesh (values on the cells,
        Dirichlet bc_value) after update_ghost_mr; this sim's own field and BC
        are left as they were."""
        keep, keep_bc = self.storage(), (self._bc if hasattr(self, "_bc") else None)
        st = np.zeros(self.nb_cells(REFERENCE))
        self.set_storage(st)
        self.set_values(np.asarray(values, dtype=np.float64))
        lib().orc_set_dirichlet(self.h, float(bc_value))
        self.update_ghost_mr()
        out = self.storage()
        self.set_storage(keep)
        if keep_bc is not None:
            lib().orc_set_dirichlet(self.h, float(keep_bc))
        return out

    def convection_upwind_heterogeneous(self, velocity_storages):
        """make_convection_upwind<Field>(velocity_field)(u) after update_ghost_mr(u);
        velocity_storages: one ghost-updated storage per component."""
        vs = [np.ascontiguousarray(v, dtype=np.float64) for v in velocity_storages]
        arr = (ctypes.POINTER(ctypes.c_double) * len(vs))(*[_dp(v) for v in vs])
        out = np.zeros(self.nb_cells(REFERENCE))
        _check(lib().orc_convection_upwind_heterogeneous(self.h, arr, _dp(out)))
        return out

    def convection_upwind(self, v):
        """make_convection_upwind(v)(u) after update_ghost_mr(u): the output storage."""
        v = np.ascontiguousarray(v, dtype=np.float64)
        out = np.zeros(self.nb_cells(REFERENCE))
        _check(lib().orc_convection_upwind(self.h, _dp(v), _dp(out)))
        return out

    def set_refine_boundary(self, on=True):
        """The --refine-boundary option for the following adaptations."""
        _check(lib().orc_set_refine_boundary(self.h, int(bool(on))))

    def set_relative_detail(self, on=True):
        """mra_config().relative_detail(on) for the following adaptations."""
        _check(lib().orc_set_relative_detail(self.h, int(bool(on))))

    def convection_weno5(self, v):
        """make_convection_weno5(v)(u) after update_ghost_mr(u) (one-level
        mesh, stencil radius 3): the output storage."""
        v = np.ascontiguousarray(v, dtype=np.float64)
        out = np.zeros(self.nb_cells(REFERENCE))
        _check(lib().orc_convection_weno5(self.h, _dp(v), _dp(out)))
        return out

    def upwind_hand_loop(self, dt):
        """update_ghost_mr, then the README's for_each_interval upwind loop and swap."""
        _check(lib().orc_upwind_hand_loop(self.h, float(dt)))

    def apply_level_op(self, op, level):
        """op 0 projection into level-1, 1 prediction<s,false> into level,
        2 prediction<s,true> from level-1, over cells[level] & reference[level-1]."""
        _check(lib().orc_apply_level_op(self.h, int(op), int(level)))

    def apply_field_bc_direction(self, direction):
        """apply_field_bc(u, direction) at every level, with the sim's Dirichlet value."""
        d = list(direction) + [0] * (3 - len(direction))
        _check(lib().orc_apply_field_bc_direction(self.h, *[int(v) for v in d]))


def weno5_flux(f):
    """The oracle's WENO5 flux of five values."""
    f = np.ascontiguousarray(f, dtype=np.float64)
    return lib().orc_weno5_flux(_dp(f))
