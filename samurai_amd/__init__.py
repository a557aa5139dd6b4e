"""MI355X-native adaptive-MR time step with samurai's API (ctypes binding of
libsamurai_amd.so, whose C ABI is declared in include/samurai_amd.h).

Names follow the reference's C++ API for this path (hpc-maths/samurai v0.33.0):
``mra.make_mesh`` -> :func:`make_mesh`, ``make_scalar_field`` -> :func:`make_scalar_field`,
``make_bc<Dirichlet<1>>`` -> :func:`make_bc_dirichlet`, ``update_ghost_mr`` ->
:func:`update_ghost_mr`, ``unp1 = u - dt * upwind(a, u)`` -> :func:`upwind`,
``make_MRAdapt`` -> :func:`make_MRAdapt`, ``samurai::swap`` -> :func:`swap`.
The library must be built (``python samurai_amd/build.py``); nothing here
falls back to a CPU implementation.
"""
import atexit
import ctypes
import os
import weakref

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsamurai_amd.so")
IO_LIB_PATH = os.path.join(HERE, "libsamurai_amd_io.so")

CELLS, CELLS_AND_GHOSTS, PROJ_CELLS, UNION_CELLS, REFERENCE = range(5)
KEEP, COARSEN, REFINE = 1, 2, 4


class SamuraiError(RuntimeError):
    pass


class MeshConfig(ctypes.Structure):
    _fields_ = [("dim", ctypes.c_int), ("min_level", ctypes.c_int), ("max_level", ctypes.c_int),
                ("prediction_stencil_radius", ctypes.c_int), ("max_stencil_radius", ctypes.c_int),
                ("graduation_width", ctypes.c_int)]


_lib = None


def lib():
    """The loaded C ABI (raises if the shared library has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise SamuraiError(f"{LIB_PATH} is missing: build it with python samurai_amd/build.py")
        # One HIP runtime per process: torch ships its own libamdhip64 and
        # cannot initialize the GPU once another copy of the runtime holds it.
        # Loading torch first makes the library bind to torch's copy (same
        # soname), so torch tensors and this library can share the device.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = ctypes.CDLL(LIB_PATH)
        vp, i32, i64, dbl = ctypes.c_void_p, ctypes.c_int, ctypes.c_int64, ctypes.c_double
        P, PP = ctypes.POINTER, lambda t: ctypes.POINTER(ctypes.POINTER(t))
        _sig(L, "smr_last_error", "restype", ctypes.c_char_p)
        _sig(L, "smr_ctx_create", "argtypes", [i32, ctypes.POINTER(vp)])
        _sig(L, "smr_ctx_destroy", "argtypes", [vp])
        _sig(L, "smr_ctx_synchronize", "argtypes", [vp])
        _sig(L, "smr_mesh_create_box", "argtypes", [vp, P(MeshConfig), ctypes.POINTER(vp)])
        _sig(L, "smr_mesh_create_from_cells", "argtypes", [vp, P(MeshConfig), P(ctypes.c_int32), i64, ctypes.POINTER(vp)])
        _sig(L, "smr_mesh_assign_cells", "argtypes", [vp, i32, i32, P(ctypes.c_int32), i64])
        _sig(L, "smr_mesh_swap", "argtypes", [vp, vp])
        _sig(L, "smr_mesh_destroy", "argtypes", [vp])
        _sig(L, "smr_mesh_nb_cells", "restype", i64)
        _sig(L, "smr_mesh_nb_cells", "argtypes", [vp, i32])
        _sig(L, "smr_mesh_get_cells", "argtypes", [vp, i32, P(i64)])
        _sig(L, "smr_mesh_get_config", "argtypes", [vp, P(MeshConfig)])
        _sig(L, "smr_field_name", "restype", ctypes.c_char_p)
        _sig(L, "smr_field_name", "argtypes", [vp])
        _sig(L, "smr_field_on_mesh", "argtypes", [vp, vp])
        _sig(L, "smr_mesh_rebuild_from_tags", "argtypes", [vp, P(ctypes.c_uint8), i64])
        _sig(L, "smr_mesh_rebuild_from_device_tags", "argtypes", [vp, vp, i64])
        _sig(L, "smr_mesh_validate", "argtypes", [vp])
        _sig(L, "smr_ctx_stream", "restype", vp)
        _sig(L, "smr_ctx_stream", "argtypes", [vp])
        _sig(L, "smr_mesh_upwind_bytes", "restype", i64)
        _sig(L, "smr_mesh_upwind_bytes", "argtypes", [vp])
        _sig(L, "smr_mesh_check_device_lists", "argtypes", [vp])
        _sig(L, "smr_mesh_ghost_update_bytes", "restype", i64)
        _sig(L, "smr_mesh_ghost_update_bytes", "argtypes", [vp])
        _sig(L, "smr_timers_enable", "argtypes", [i32])
        _sig(L, "smr_timers_report", "restype", i64)
        _sig(L, "smr_timers_report", "argtypes", [ctypes.c_char_p, i64])
        _sig(L, "smr_field_create", "argtypes", [vp, ctypes.c_char_p, ctypes.POINTER(vp)])
        _sig(L, "smr_field_destroy", "argtypes", [vp])
        _sig(L, "smr_field_resize", "argtypes", [vp])
        _sig(L, "smr_field_set_cells", "argtypes", [vp, P(dbl)])
        _sig(L, "smr_field_get_cells", "argtypes", [vp, P(dbl)])
        _sig(L, "smr_field_get_storage", "argtypes", [vp, P(dbl)])
        _sig(L, "smr_field_set_storage", "argtypes", [vp, P(dbl)])
        _sig(L, "smr_field_size", "restype", i64)
        _sig(L, "smr_field_size", "argtypes", [vp])
        _sig(L, "smr_field_device_data", "restype", vp)
        _sig(L, "smr_field_device_data", "argtypes", [vp])
        _sig(L, "smr_field_attach_dirichlet", "argtypes", [vp, dbl])
        _sig(L, "smr_field_swap", "argtypes", [vp, vp])
        _sig(L, "smr_update_ghost_mr", "argtypes", [vp])
        _sig(L, "smr_upwind", "argtypes", [P(dbl), dbl, vp, vp])
        _sig(L, "smr_adapt_create", "argtypes", [vp, ctypes.POINTER(vp)])
        _sig(L, "smr_adapt_destroy", "argtypes", [vp])
        _sig(L, "smr_adapt_create_fields", "argtypes", [ctypes.POINTER(vp), i32, ctypes.POINTER(vp)])
        _sig(L, "smr_adapt_nb_fields", "argtypes", [vp])
        _sig(L, "smr_adapt_run", "argtypes", [vp, dbl, dbl])
        _sig(L, "smr_adapt_set_relative_detail", "argtypes", [vp, i32])
        _sig(L, "smr_adapt_set_refine_boundary", "argtypes", [vp, i32])
        _sig(L, "smr_adapt_harten", "argtypes", [vp, i32, dbl, dbl])
        _sig(L, "smr_adapt_tags", "argtypes", [vp, i32, dbl, dbl, vp])
        _sig(L, "smr_adapt_tags_noalloc", "argtypes", [vp, i32, dbl, dbl, vp])
        _sig(L, "smr_adapt_last_size", "restype", i64)
        _sig(L, "smr_adapt_last_size", "argtypes", [vp])
        _sig(L, "smr_adapt_last_tags", "argtypes", [vp, P(ctypes.c_uint8), P(dbl)])
        _sig(L, "smr_batch_create", "argtypes", [vp, i32, P(vp), P(vp), P(dbl), P(dbl), ctypes.POINTER(vp)])
        _sig(L, "smr_batch_destroy", "argtypes", [vp])
        _sig(L, "smr_batch_adapt", "argtypes", [vp, dbl, dbl])
        _sig(L, "smr_batch_advance", "argtypes", [vp])
        _sig(L, "smr_batch_step", "argtypes", [vp, dbl, dbl])
        _sig(L, "smr_batch_nb_cells", "restype", i64)
        _sig(L, "smr_batch_nb_cells", "argtypes", [vp])
        _sig(L, "smr_batch_set_tuning", "argtypes", [vp, i32, i32, i32])
        _sig(L, "smr_batch_get_tuning", "argtypes", [vp, P(i32), P(i32), P(i32)])
        _sig(L, "smr_batch_autotune", "argtypes", [vp, i32, ctypes.c_char_p])
        _sig(L, "smr_batch_load_tuning", "argtypes", [vp, ctypes.c_char_p])
        _sig(L, "smr_convection_upwind", "argtypes", [P(dbl), vp, vp])
        _sig(L, "smr_convection_upwind_nonlinear", "argtypes", [vp, vp])
        _sig(L, "smr_convection_weno5", "argtypes", [P(dbl), vp, vp])
        _sig(L, "smr_upwind_scalar_burgers", "argtypes", [P(dbl), dbl, vp, vp])
        _sig(L, "smr_convection_upwind_field", "argtypes", [ctypes.POINTER(vp), vp, vp])
        _sig(L, "smr_upwind_hand_loop", "argtypes", [dbl, vp, vp])
        _sig(L, "smr_field_sub_scaled", "argtypes", [vp, dbl, vp, vp])
        _sig(L, "smr_mesh_convection_terms", "argtypes", [vp, P(dbl), P(i64), P(i64), P(i64), P(i64), P(i64), P(dbl)])
        _sig(L, "smr_adjoint_create", "argtypes", [vp, i32, P(vp), P(dbl), P(dbl), i32, i32, ctypes.POINTER(vp)])
        _sig(L, "smr_adjoint_destroy", "argtypes", [vp])
        _sig(L, "smr_adjoint_nb_cells", "restype", i64)
        _sig(L, "smr_adjoint_nb_cells", "argtypes", [vp, i32])
        _sig(L, "smr_adjoint_set_velocity", "argtypes", [vp, P(dbl)])
        _sig(L, "smr_adjoint_forward", "argtypes", [vp, vp, vp])
        _sig(L, "smr_adjoint_backward", "argtypes", [vp, vp, vp, P(dbl)])
        _sig(L, "smr_adjoint_set_tuning", "argtypes", [vp, i32, i32, i32])
        _lib = L
    return _lib


def _sig(L, name, kind, value):
    """Declare a C signature; a symbol missing from an older build stays
    undeclared and fails when that entry point is called."""
    if hasattr(L, name):
        setattr(getattr(L, name), kind, value)


def _check(rc):
    if rc < 0:
        raise SamuraiError(f"samurai_amd error {rc}: {lib().smr_last_error().decode()}")
    return rc


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class _Handle:
    """A library object owned by a context.  close() destroys it once; the
    context closes every object still alive before itself, dependents first
    (batched and adjoint problems, adaptations, fields, meshes), so no object
    is destroyed after the stream and device memory it uses."""

    _destroy = None  # name of the smr_*_destroy function

    def _own(self, ctx, h):
        self.h, self._ctx = h, ctx
        ctx._children.add(self)

    def close(self):
        h = getattr(self, "h", None)
        self.h = None
        if h and _lib is not None:
            getattr(_lib, self._destroy)(h)

    def __del__(self):
        self.close()


# the order in which Context.close() destroys the objects still alive
_CLOSE_ORDER = ("Batch", "AdjointProblem", "Adapt", "ScalarField", "MRMesh")
_contexts = weakref.WeakSet()


class Context:
    """A device (``device >= 0``) or host-only (``device < 0``) context."""

    def __init__(self, device=0):
        h = ctypes.c_void_p()
        _check(lib().smr_ctx_create(int(device), ctypes.byref(h)))
        self.h = h
        self.device = device
        self._children = weakref.WeakSet()
        _contexts.add(self)

    def synchronize(self):
        _check(lib().smr_ctx_synchronize(self.h))

    def stream(self):
        """The hipStream_t (as an integer) all work of this context runs on."""
        return lib().smr_ctx_stream(self.h)

    def close(self):
        """Destroys every object of this context still alive, then the context
        (idempotent; also run for every context at interpreter exit)."""
        if getattr(self, "h", None) is None:
            return
        children = list(self._children)
        for kind in _CLOSE_ORDER:
            for obj in children:
                if type(obj).__name__ == kind:
                    obj.close()
        for obj in children:
            obj.close()
        h, self.h = self.h, None
        if _lib is not None:
            _lib.smr_ctx_destroy(h)

    def __del__(self):
        self.close()


@atexit.register
def _close_contexts():
    # before the interpreter and the HIP runtime tear down: no library object
    # is left for a finalizer or a static destructor to release later
    for ctx in list(_contexts):
        ctx.close()


class MRMesh(_Handle):
    """samurai::MRMesh on the unit square/cube (mr/mesh.hpp)."""

    _destroy = "smr_mesh_destroy"

    def __init__(self, ctx, handle, cfg):
        self._own(ctx, handle)
        self.ctx, self.cfg = ctx, cfg
        self.dim = cfg.dim

    @property
    def min_level(self):
        return self.cfg.min_level

    @property
    def max_level(self):
        return self.cfg.max_level

    def nb_cells(self, mesh_id=CELLS):
        return _check(lib().smr_mesh_nb_cells(self.h, mesh_id))

    def cells(self, mesh_id=CELLS):
        """(n, 5) int64 rows (level, x, y, z, storage index) in traversal order."""
        n = self.nb_cells(mesh_id)
        out = np.zeros((n, 5), dtype=np.int64)
        _check(lib().smr_mesh_get_cells(self.h, mesh_id, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))))
        return out

    def upwind_bytes(self):
        """Algorithmic HBM bytes of one upwind launch on this mesh."""
        return _check(lib().smr_mesh_upwind_bytes(self.h))

    def ghost_update_bytes(self):
        """Algorithmic HBM bytes of one update_ghost_mr on this mesh."""
        return _check(lib().smr_mesh_ghost_update_bytes(self.h))

    def validate(self):
        """Host check that every kernel access of this mesh's programs is in bounds."""
        return _check(lib().smr_mesh_validate(self.h))

    def check_device_lists(self):
        """0 when the work lists the device built for this mesh equal the host's, 1 when it built none."""
        return _check(lib().smr_mesh_check_device_lists(self.h))

    def refresh_config(self):
        """Re-read the configuration (the levels change when a restart file is loaded into the mesh)."""
        cfg = MeshConfig()
        _check(lib().smr_mesh_get_config(self.h, ctypes.byref(cfg)))
        self.cfg = cfg

    def assign_cells(self, cells, min_level=None, max_level=None):
        """mesh = MRMesh(ca, cfg) in place, from (n, >=4) rows (level, x, y, z)."""
        c = np.ascontiguousarray(np.asarray(cells).reshape(-1, np.asarray(cells).shape[-1])[:, :4], dtype=np.int32)
        lo = self.min_level if min_level is None else min_level
        hi = self.max_level if max_level is None else max_level
        _check(lib().smr_mesh_assign_cells(self.h, lo, hi, c.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), c.shape[0]))
        self.refresh_config()

    def rebuild_from_tags(self, tags):
        tags = np.ascontiguousarray(tags, dtype=np.uint8)
        return _check(lib().smr_mesh_rebuild_from_tags(
            self.h, tags.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)), tags.size))

    def rebuild_from_device_tags(self, tags_ptr, n):
        """rebuild_from_tags on the device from n tags at device address tags_ptr."""
        return _check(lib().smr_mesh_rebuild_from_device_tags(self.h, ctypes.c_void_p(tags_ptr), int(n)))


def swap_meshes(a, b):
    """std::swap(mesh_a, mesh_b) of two meshes of one context: the contents
    (and configurations) are exchanged, fields stay with their mesh object and
    must be resized before their next use."""
    _check(lib().smr_mesh_swap(a.h, b.h))
    a.refresh_config()
    b.refresh_config()


def mesh_config(dim, min_level, max_level, prediction_stencil_radius=1, max_stencil_radius=1, graduation_width=1):
    return MeshConfig(dim, min_level, max_level, prediction_stencil_radius, max_stencil_radius, graduation_width)


def make_mesh(ctx, cfg):
    """samurai::mra::make_mesh(unit box, cfg): every cell at max_level."""
    h = ctypes.c_void_p()
    _check(lib().smr_mesh_create_box(ctx.h, ctypes.byref(cfg), ctypes.byref(h)))
    return MRMesh(ctx, h, cfg)


def make_mesh_from_cells(ctx, cfg, cells):
    """MRMesh(ca, cfg) from (n, >=4) rows (level, x, y, z)."""
    c = np.ascontiguousarray(np.asarray(cells)[:, :4], dtype=np.int32)
    h = ctypes.c_void_p()
    _check(lib().smr_mesh_create_from_cells(ctx.h, ctypes.byref(cfg), c.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                            c.shape[0], ctypes.byref(h)))
    return MRMesh(ctx, h, cfg)


class ScalarField(_Handle):
    """samurai::ScalarField<mesh, double> stored on the device."""

    _destroy = "smr_field_destroy"

    def __init__(self, mesh, name):
        h = ctypes.c_void_p()
        _check(lib().smr_field_create(mesh.h, name.encode(), ctypes.byref(h)))
        self._own(mesh.ctx, h)
        self.mesh, self.name = mesh, name

    def resize(self):
        _check(lib().smr_field_resize(self.h))

    def size(self):
        return lib().smr_field_size(self.h)

    def device_ptr(self):
        return lib().smr_field_device_data(self.h)

    def set_cells(self, values):
        v = np.ascontiguousarray(values, dtype=np.float64)
        if v.size != self.mesh.nb_cells(CELLS):
            raise ValueError("one value per leaf cell expected")
        _check(lib().smr_field_set_cells(self.h, _dp(v)))

    def cells(self):
        out = np.zeros(self.mesh.nb_cells(CELLS))
        _check(lib().smr_field_get_cells(self.h, _dp(out)))
        return out

    def storage(self):
        out = np.zeros(self.size())
        _check(lib().smr_field_get_storage(self.h, _dp(out)))
        return out

    def set_storage(self, values):
        v = np.ascontiguousarray(values, dtype=np.float64)
        if v.size != self.size():
            raise ValueError("storage size mismatch")
        _check(lib().smr_field_set_storage(self.h, _dp(v)))


def timers_enable(on=True):
    """Phase timers under the reference's timer names (synchronizing: diagnostics only)."""
    lib().smr_timers_enable(int(on))


def timers_reset():
    lib().smr_timers_reset()


def timers_report():
    import json
    n = lib().smr_timers_report(None, 0)
    buf = ctypes.create_string_buffer(int(n) + 1)
    lib().smr_timers_report(buf, n + 1)
    return json.loads(buf.value.decode())


def make_scalar_field(name, mesh):
    return ScalarField(mesh, name)


def make_bc_dirichlet(field, value=0.0):
    """make_bc<Dirichlet<1>>(u, value)."""
    _check(lib().smr_field_attach_dirichlet(field.h, float(value)))


def swap(u, v):
    _check(lib().smr_field_swap(u.h, v.h))


def update_ghost_mr(*fields):
    """update_ghost_mr(u, v, ...) (algorithm/update_ghost_mr.hpp:260-270)."""
    for field in fields:
        _check(lib().smr_update_ghost_mr(field.h))


def upwind(a, dt, u, unp1):
    """unp1 = u - dt * upwind(a, u) on mesh[cells]."""
    av = np.ascontiguousarray(a, dtype=np.float64)
    _check(lib().smr_upwind(_dp(av), float(dt), u.h, unp1.h))


def upwind_scalar_burgers(k, dt, u, unp1):
    """unp1 = u - dt * upwind_scalar_burgers(k, u) (stencil_field.hpp:181-255), 2D."""
    k = np.ascontiguousarray(np.asarray(k, dtype=np.float64).ravel())
    _check(lib().smr_upwind_scalar_burgers(_dp(k), float(dt), u.h, unp1.h))


def sub_scaled(u, a, v, out):
    """out = u - a * v on the cells (unp1 = u - dt * conv(u))."""
    _check(lib().smr_field_sub_scaled(u.h, float(a), v.h, out.h))


def upwind_hand_loop(dt, u, unp1):
    """The README's for_each_interval loop: unp1 = u - dt/dx*(u - u(i-1) + u - u(j-1)) (2D)."""
    _check(lib().smr_upwind_hand_loop(float(dt), u.h, unp1.h))


class ConvectionUpwind:
    """make_convection_upwind<Field>(velocity) (schemes/fv/operators/convection_lin.hpp:15-89),
    with a list of component fields the velocity-field form (:184-264),
    or with velocity None the non-linear make_convection_upwind<Field>()
    (convection_nonlin.hpp:24-79, flux f(u) = u * u); calling it on u returns
    the operator applied to u in a new field (or in `out`)."""

    def __init__(self, velocity=None):
        self.fields = None
        if velocity is not None and len(velocity) > 0 and isinstance(velocity[0], ScalarField):
            # make_convection_upwind<Field>(velocity_field) (convection_lin.hpp:184-264):
            # one scalar field per velocity component
            self.fields, self.velocity = list(velocity), None
        else:
            self.velocity = None if velocity is None else np.ascontiguousarray(np.asarray(velocity, dtype=np.float64).ravel())

    def __call__(self, u, out=None):
        if out is None:
            out = make_scalar_field("convection(" + u.name + ")", u.mesh)
        if self.fields is not None:
            arr = (ctypes.c_void_p * len(self.fields))(*[f.h for f in self.fields])
            _check(lib().smr_convection_upwind_field(arr, u.h, out.h))
        elif self.velocity is None:
            _check(lib().smr_convection_upwind_nonlinear(u.h, out.h))
        else:
            _check(lib().smr_convection_upwind(_dp(self.velocity), u.h, out.h))
        return out


def make_convection_upwind(velocity=None):
    return ConvectionUpwind(velocity)


class ConvectionWeno5:
    """make_convection_weno5<Field>(velocity) (schemes/fv/operators/convection_lin.hpp:95-178)
    on a mesh of one level with ghost layers of width 3 (mesh_config(...,
    max_stencil_radius=3), min_level == max_level); calling it on u returns
    the operator applied to u in a new field (or in `out`)."""

    def __init__(self, velocity):
        self.velocity = np.ascontiguousarray(np.asarray(velocity, dtype=np.float64).ravel())

    def __call__(self, u, out=None):
        if out is None:
            out = make_scalar_field("convection(" + u.name + ")", u.mesh)
        v = np.zeros(3)
        v[: len(self.velocity)] = self.velocity
        _check(lib().smr_convection_weno5(_dp(v), u.h, out.h))
        return out


def make_convection_weno5(velocity):
    return ConvectionWeno5(velocity)


def convection_terms(mesh, velocity=None):
    """The ordered contribution lists behind make_convection_upwind (host); the
    non-linear operator's with velocity None."""
    v = None if velocity is None else np.ascontiguousarray(np.asarray(velocity, dtype=np.float64).ravel())
    i64 = ctypes.c_int64
    nr, nt = i64(0), i64(0)
    null = ctypes.POINTER(i64)()
    vp = _dp(v) if v is not None else ctypes.POINTER(ctypes.c_double)()
    _check(lib().smr_mesh_convection_terms(mesh.h, vp, ctypes.byref(nr), ctypes.byref(nt), null, null, null,
                                           ctypes.POINTER(ctypes.c_double)()))
    row = np.zeros(nr.value, dtype=np.int64)
    ptr = np.zeros(nr.value + 1, dtype=np.int64)
    src = np.zeros(2 * nt.value, dtype=np.int64)
    coef = np.zeros(nt.value)
    P64 = ctypes.POINTER(i64)
    _check(lib().smr_mesh_convection_terms(mesh.h, vp, ctypes.byref(nr), ctypes.byref(nt), row.ctypes.data_as(P64),
                                           ptr.ctypes.data_as(P64), src.ctypes.data_as(P64), _dp(coef)))
    return row, ptr, src, coef


class Adapt(_Handle):
    """samurai::Adapt built by make_MRAdapt(u, v, ...) (mr/adapt.hpp): the
    fields (one mesh) are adapted together."""

    _destroy = "smr_adapt_destroy"

    def __init__(self, *fields):
        if not fields:
            raise ValueError("make_MRAdapt needs at least one field")
        h = ctypes.c_void_p()
        if len(fields) == 1:
            _check(lib().smr_adapt_create(fields[0].h, ctypes.byref(h)))
        else:
            arr = (ctypes.c_void_p * len(fields))(*[f.h for f in fields])
            _check(lib().smr_adapt_create_fields(arr, len(fields), ctypes.byref(h)))
        self._own(fields[0].mesh.ctx, h)
        self.field, self.fields = fields[0], list(fields)

    def __call__(self, epsilon=1e-4, regularity=1.0, relative_detail=False, refine_boundary=False):
        """MRadaptation(mra_config().epsilon(eps).regularity(reg).relative_detail(rel)),
        refine_boundary as the reference's --refine-boundary; returns #harten calls."""
        _check(lib().smr_adapt_set_relative_detail(self.h, int(bool(relative_detail))))
        _check(lib().smr_adapt_set_refine_boundary(self.h, int(bool(refine_boundary))))
        return _check(lib().smr_adapt_run(self.h, float(epsilon), float(regularity)))

    def harten(self, ite, epsilon, regularity):
        return _check(lib().smr_adapt_harten(self.h, int(ite), float(epsilon), float(regularity)))

    def tags_into(self, ite, epsilon, regularity, tags_ptr):
        """The device half of harten(ite): tags into the device array at
        tags_ptr (reference-storage size bytes); the mesh is not changed."""
        _check(lib().smr_adapt_tags(self.h, int(ite), float(epsilon), float(regularity), ctypes.c_void_p(tags_ptr)))

    def tags_into_noalloc(self, ite, epsilon, regularity, tags_ptr):
        """tags_into without allocating: details in the library's static
        per-GPU scratch array (the mesh's programs must exist already)."""
        _check(lib().smr_adapt_tags_noalloc(self.h, int(ite), float(epsilon), float(regularity), ctypes.c_void_p(tags_ptr)))

    def last_tags(self):
        """Tags and details of the last harten call (details: one array per
        field, concatenated; a single field gives one array)."""
        n = lib().smr_adapt_last_size(self.h)
        t = np.zeros(max(n, 0), dtype=np.uint8)
        d = np.zeros(max(n, 0) * len(self.fields))
        _check(lib().smr_adapt_last_tags(self.h, t.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)), _dp(d)))
        return t, d


def make_MRAdapt(*fields):  # noqa: N802 (reference name)
    return Adapt(*fields)


def upwind_variant_count():
    return lib().smr_upwind_variant_count()


def set_upwind_variant(i):
    """Launch shape of the upwind kernel (tuning knob; results are identical)."""
    _check(lib().smr_set_upwind_variant(int(i)))


# launch modes of the batched stage programs
MODE_FUSED, MODE_STAGES, MODE_GRAPH = 0, 1, 2


def default_tuning_file():
    """Where the batch autotuner keeps its choice per GPU architecture."""
    d = os.environ.get("SMR_TUNE_DIR") or os.path.join(os.path.expanduser("~"), ".cache", "samurai_amd")
    return os.path.join(d, "batch_tuning.txt")


class Batch(_Handle):
    """Many independent instances of the advection demo loop advanced together.

    instances: list of (u, unp1, a, dt) with u, unp1 ScalarFields on one mesh
    per instance.  step(eps, reg) is, for every instance, MRadaptation(eps,
    reg); unp1.resize(); update_ghost_mr(u); unp1 = u - dt * upwind(a, u);
    swap(u, unp1) -- bit-identical to those per-instance calls."""

    _destroy = "smr_batch_destroy"

    def __init__(self, ctx, instances):
        n = len(instances)
        vp = ctypes.c_void_p
        us = (vp * max(n, 1))(*[i[0].h for i in instances])
        vs = (vp * max(n, 1))(*[i[1].h for i in instances])
        a = np.zeros((max(n, 1), 3))
        dt = np.zeros(max(n, 1))
        for k, (_, _, ak, dk) in enumerate(instances):
            ak = np.asarray(ak, dtype=np.float64).ravel()
            a[k, :len(ak)] = ak
            dt[k] = dk
        h = vp()
        _check(lib().smr_batch_create(ctx.h, n, us, vs, _dp(a), _dp(dt), ctypes.byref(h)))
        self._own(ctx, h)
        self.ctx, self.instances = ctx, list(instances)

    def adapt(self, epsilon, regularity):
        return _check(lib().smr_batch_adapt(self.h, float(epsilon), float(regularity)))

    def advance(self):
        _check(lib().smr_batch_advance(self.h))

    def step(self, epsilon, regularity):
        return _check(lib().smr_batch_step(self.h, float(epsilon), float(regularity)))

    def nb_cells(self):
        return lib().smr_batch_nb_cells(self.h)

    def set_tuning(self, block, cells_per_thread, mode):
        _check(lib().smr_batch_set_tuning(self.h, int(block), int(cells_per_thread), int(mode)))

    def tuning(self):
        b, c, m = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        _check(lib().smr_batch_get_tuning(self.h, ctypes.byref(b), ctypes.byref(c), ctypes.byref(m)))
        return {"block": b.value, "cells_per_thread": c.value, "mode": m.value}

    def autotune(self, reps=5, cache_file=None):
        if cache_file:
            os.makedirs(os.path.dirname(cache_file) or ".", exist_ok=True)
        _check(lib().smr_batch_autotune(self.h, int(reps), cache_file.encode() if cache_file else None))
        return self.tuning()

    def load_tuning(self, cache_file):
        return bool(_check(lib().smr_batch_load_tuning(self.h, cache_file.encode())))


class AdjointProblem(_Handle):
    """Frozen-mesh upwind steps of many instances and their adjoint.

    meshes: adapted MRMesh objects (not adapted again while this lives);
    dt: per-instance time steps; bc: per-instance Dirichlet values (None: no
    boundary condition).  forward/backward take and return device pointers
    of compact cell vectors (see samurai_amd.torch_adjoint for tensors)."""

    _destroy = "smr_adjoint_destroy"

    def __init__(self, ctx, meshes, dt, bc=None, steps=100, checkpoint_every=10):
        n = len(meshes)
        vp = ctypes.c_void_p
        ms = (vp * max(n, 1))(*[m.h for m in meshes])
        dt = np.ascontiguousarray(np.asarray(dt, dtype=np.float64).reshape(n))
        bcp = None
        if bc is not None:
            bc = np.ascontiguousarray(np.asarray(bc, dtype=np.float64).reshape(n))
            bcp = _dp(bc)
        h = vp()
        _check(lib().smr_adjoint_create(ctx.h, n, ms, _dp(dt), bcp, int(steps), int(checkpoint_every), ctypes.byref(h)))
        self._own(ctx, h)
        self.ctx, self.meshes = ctx, list(meshes)
        self.n, self.steps, self.checkpoint_every = n, int(steps), int(checkpoint_every)

    def nb_cells(self, instance=-1):
        return _check(lib().smr_adjoint_nb_cells(self.h, int(instance)))

    def offsets(self):
        """First compact cell of each instance, plus the total."""
        return np.concatenate([[0], np.cumsum([self.nb_cells(i) for i in range(self.n)])]).astype(np.int64)

    def set_velocity(self, a):
        a3 = np.zeros((max(self.n, 1), 3))
        a = np.asarray(a, dtype=np.float64).reshape(self.n, -1)
        a3[:self.n, :a.shape[1]] = a
        _check(lib().smr_adjoint_set_velocity(self.h, _dp(a3)))

    def set_tuning(self, block, cells_per_thread, mode):
        _check(lib().smr_adjoint_set_tuning(self.h, int(block), int(cells_per_thread), int(mode)))

    def forward_ptr(self, u0_ptr, uT_ptr):
        _check(lib().smr_adjoint_forward(self.h, ctypes.c_void_p(u0_ptr), ctypes.c_void_p(uT_ptr)))

    def backward_ptr(self, gT_ptr, g0_ptr):
        ga = np.zeros((max(self.n, 1), 3))
        _check(lib().smr_adjoint_backward(self.h, ctypes.c_void_p(gT_ptr), ctypes.c_void_p(g0_ptr), _dp(ga)))
        return ga[:self.n]


# ---------------------------------------------------------------------------
# restart files (io/restart.hpp layout), libsamurai_amd_io.so
# ---------------------------------------------------------------------------
_io = None


def io_lib():
    """The restart-file library (HDF5); loaded on first use."""
    global _io
    if _io is None:
        lib()  # the compute library first: the I/O library is linked against it
        if not os.path.exists(IO_LIB_PATH):
            raise SamuraiError(f"{IO_LIB_PATH} is missing: build it with python samurai_amd/build.py (needs HDF5)")
        L = ctypes.CDLL(IO_LIB_PATH)
        vp, i32, i64 = ctypes.c_void_p, ctypes.c_int, ctypes.c_int64
        L.smr_io_last_error.restype = ctypes.c_char_p
        L.smr_restart_dump.argtypes = [ctypes.c_char_p, vp, ctypes.POINTER(vp), i32]
        L.smr_restart_load_mesh.argtypes = [ctypes.c_char_p, vp, ctypes.POINTER(MeshConfig), ctypes.POINTER(vp)]
        L.smr_restart_load_field.argtypes = [ctypes.c_char_p, vp, vp]
        L.smr_restart_load.argtypes = [ctypes.c_char_p, vp, ctypes.POINTER(vp), i32]
        L.smr_restart_dump_cells.argtypes = [ctypes.c_char_p, i32, ctypes.POINTER(ctypes.c_int32), i64]
        L.smr_restart_load_cells.restype = i64
        L.smr_restart_load_cells.argtypes = [ctypes.c_char_p, ctypes.POINTER(i32), ctypes.POINTER(ctypes.c_int32), i64]
        L.smr_save.argtypes = [ctypes.c_char_p, ctypes.c_char_p, vp, ctypes.POINTER(vp), i32, i32]
        L.smr_save_cell_array.argtypes = [ctypes.c_char_p, ctypes.c_char_p, i32, ctypes.POINTER(ctypes.c_int32), i64,
                                          ctypes.POINTER(ctypes.POINTER(ctypes.c_double)), ctypes.POINTER(ctypes.c_char_p), i32, i32]
        _io = L
    return _io


def _io_check(rc):
    if rc < 0:
        raise SamuraiError(f"samurai_amd restart error {rc}: {io_lib().smr_io_last_error().decode()}")
    return rc


def _h5(filename):
    # samurai::dump / load take the name without extension and add ".h5"
    return (os.fspath(filename) + ".h5").encode()


def dump(filename, mesh, *fields):
    """samurai::dump(filename, mesh, fields...) (io/restart.hpp:189-237): the
    leaves of `mesh` and the named scalar fields into filename + ".h5"."""
    arr = (ctypes.c_void_p * max(len(fields), 1))(*[f.h for f in fields])
    _io_check(io_lib().smr_restart_dump(_h5(filename), mesh.h, arr, len(fields)))


def load_mesh(filename, ctx, cfg):
    """The mesh part of samurai::load(filename, mesh) (io/restart.hpp:423-462):
    a new MRMesh from the file's leaves, with cfg's stencil settings and the
    file's min and max levels."""
    h = ctypes.c_void_p()
    _io_check(io_lib().smr_restart_load_mesh(_h5(filename), ctx.h, ctypes.byref(cfg), ctypes.byref(h)))
    out = MeshConfig()
    _check(lib().smr_mesh_get_config(h, ctypes.byref(out)))
    return MRMesh(ctx, h, out)


def load_field(filename, mesh, field):
    """load_field(file, mesh, field) (io/restart.hpp:341-386): the values of
    /fields/{field.name} into `field`, which lives on `mesh`."""
    _io_check(io_lib().smr_restart_load_field(_h5(filename), mesh.h, field.h))


def load(filename, ctx, cfg, *names):
    """samurai::load(filename, mesh, fields...): the mesh, then one scalar field
    per name; returns (mesh, [fields])."""
    mesh = load_mesh(filename, ctx, cfg)
    fields = []
    for name in names:
        f = make_scalar_field(name, mesh)
        load_field(filename, mesh, f)
        fields.append(f)
    return mesh, fields


def load_into(filename, mesh, *fields):
    """samurai::load(filename, mesh, fields...) (io/restart.hpp:423-462): `mesh`
    is replaced in place by the file's leaves and levels, and each field on it
    gets the file's values of its name."""
    arr = (ctypes.c_void_p * max(len(fields), 1))(*[f.h for f in fields])
    _io_check(io_lib().smr_restart_load(_h5(filename), mesh.h, arr, len(fields)))
    mesh.refresh_config()


def dump_cells(filename, dim, cells):
    """samurai::dump(filename, CellArray) (io/restart.hpp:126-150): leaves as
    (n, >=4) rows (level, x, y, z) in for_each_cell order."""
    c = np.ascontiguousarray(np.asarray(cells).reshape(-1, np.asarray(cells).shape[-1])[:, :4], dtype=np.int32)
    _io_check(io_lib().smr_restart_dump_cells(_h5(filename), int(dim), c.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                              c.shape[0]))


def load_cells(filename):
    """samurai::load(filename, CellArray): (dim, (n, 4) int32 rows (level, x, y, z))."""
    dim = ctypes.c_int()
    n = _io_check(io_lib().smr_restart_load_cells(_h5(filename), ctypes.byref(dim), None, 0))
    out = np.zeros((n, 4), dtype=np.int32)
    _io_check(io_lib().smr_restart_load_cells(_h5(filename), ctypes.byref(dim), out.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), n))
    return dim.value, out



SAVE_BY_LEVEL, SAVE_BY_MESH_ID = 1, 2


def _save_options(by_level, by_mesh_id):
    return (SAVE_BY_LEVEL if by_level else 0) | (SAVE_BY_MESH_ID if by_mesh_id else 0)


def save(path, filename, mesh, *fields, by_level=False, by_mesh_id=False):
    """samurai::save(path, filename, Hdf5Options(by_level, by_mesh_id), mesh,
    fields...) (io/hdf5.hpp:1280-1329): path/filename.h5 and .xdmf."""
    arr = (ctypes.c_void_p * max(len(fields), 1))(*[f.h for f in fields])
    _io_check(io_lib().smr_save(os.fspath(path).encode(), filename.encode(), mesh.h, arr, len(fields),
                                _save_options(by_level, by_mesh_id)))


def save_cell_array(path, filename, dim, cells, fields=None, by_level=False, by_mesh_id=False):
    """samurai::save of a CellArray with fields (io/hdf5.hpp:958-1000): cells as
    (n, >=4) rows (level, x, y, z) in for_each_cell order, fields a mapping
    name -> n values."""
    c = np.ascontiguousarray(np.asarray(cells).reshape(-1, np.asarray(cells).shape[-1])[:, :4], dtype=np.int32)
    fields = dict(fields or {})
    vals = [np.ascontiguousarray(v, dtype=np.float64) for v in fields.values()]
    for v in vals:
        if v.shape != (c.shape[0],):
            raise SamuraiError(f"a field has {v.size} values for {c.shape[0]} cells")
    dp = ctypes.POINTER(ctypes.c_double)
    varr = (dp * max(len(vals), 1))(*[v.ctypes.data_as(dp) for v in vals])
    narr = (ctypes.c_char_p * max(len(vals), 1))(*[n.encode() for n in fields])
    _io_check(io_lib().smr_save_cell_array(os.fspath(path).encode(), filename.encode(), int(dim),
                                           c.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), c.shape[0], varr, narr, len(vals),
                                           _save_options(by_level, by_mesh_id)))
