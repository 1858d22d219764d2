"""ctypes binding of libpartia_hip.so (the C ABI declared in include/partia_hip.h).

This is the Python counterpart of the `ccall` layer a Julia backend would use.
The library is loaded from the package directory (partia.jl_amd/libpartia_hip.so);
there is no fallback: if it is missing or cannot be loaded, NativeLibraryError is
raised.
"""
import atexit
import ctypes as C
import os
import weakref

import numpy as np

from .errors import raise_for_status

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.normpath(os.path.join(_HERE, "..", "libpartia_hip.so"))

PH_OK = 0
PH_ERR_ARGUMENT, PH_ERR_DIMENSION, PH_ERR_STRATEGY, PH_ERR_LEAF_ORDER, PH_ERR_CAPACITY = -1, -2, -3, -4, -5
PH_GATHER, PH_SCATTER, PH_SYMMETRIC = 0, 1, 2
PH_PREC_FP64, PH_PREC_FP32, PH_PREC_FP32C = 64, 32, 33
PH_TRAVERSAL_AUTO, PH_TRAVERSAL_LANE, PH_TRAVERSAL_BUCKET, PH_TRAVERSAL_WAVE = 0, 1, 2, 3
GET = dict(order=0, codes=1, left=2, right=3, escape=4, parent=5, leaf_coor=6, leaf_scale=7, node_min=8,
           node_max=9, node_scale=10, columns=11)

# Every symbol the header declares (checked by the CPU test suite).
EXPORTS = (
    "ph_ctx_create", "ph_ctx_destroy", "ph_last_error", "ph_get_stats", "ph_set_count_neighbours",
    "ph_set_traversal",
    "ph_device_alloc", "ph_device_free", "ph_memcpy_htod", "ph_memcpy_dtoh", "ph_synchronize",
    "ph_lbvh_build", "ph_lbvh_build_device", "ph_lbvh_build2d", "ph_lbvh_build2d_device", "ph_lbvh_dims", "ph_query_device", "ph_find_nearest_device",
    "ph_lbvh_destroy", "ph_lbvh_set_columns", "ph_lbvh_info",
    "ph_lbvh_get",
    "ph_point_samples", "ph_point_samples_device", "ph_point_samples_device_noalloc", "ph_structured_grid", "ph_structured_grid_device",
    "ph_line_samples", "ph_line_samples_device", "ph_line_grid_device", "ph_nearest_h", "ph_line_table",
    "ph_batch_create", "ph_batch_run", "ph_batch_times", "ph_batch_destroy",
    "ph_line_samples_adjoint_device", "ph_line_grid_adjoint_device",
    "ph_line_polar_device", "ph_line_polar_adjoint_device",
    "ph_structured_grid_coords", "ph_structured_grid_coords_device",
    "ph_point_samples_adjoint_device", "ph_structured_grid_adjoint_device",
    "ph_set_autotune", "ph_set_launch", "ph_autotune_report", "ph_set_tile_shape",
    "ph_point_single", "ph_point_single_device", "ph_line_samples_ha", "ph_line_samples_ha_device",
    "ph_line_density", "ph_line_density_device",
)

PH_REQ_POINTS, PH_REQ_PLANE, PH_REQ_BOX, PH_REQ_POLAR, PH_REQ_CYLINDER = 0, 1, 2, 3, 4
PH_SINGLE_DENSITY, PH_SINGLE_NUMBER_DENSITY, PH_SINGLE_GRADIENT_DENSITY = 0, 1, 2
PH_COORDS_CARTESIAN, PH_COORDS_CYLINDRICAL, PH_COORDS_SPHERICAL = 0, 1, 2


class NativeLibraryError(RuntimeError):
    """libpartia_hip.so is missing or unusable; there is no CPU fallback."""


class ph_catalog(C.Structure):
    _fields_ = [
        ("n_scalar", C.c_int32), ("scalar_slots", C.c_void_p), ("scalar_shepard", C.c_void_p),
        ("n_grad", C.c_int32), ("grad_slots", C.c_void_p),
        ("n_div", C.c_int32), ("div_slots", C.c_void_p),
        ("n_curl", C.c_int32), ("curl_slots", C.c_void_p),
    ]


class ph_request(C.Structure):
    _fields_ = [("kind", C.c_int32), ("reserved", C.c_int32), ("count", C.c_int64),
                ("px", C.c_void_p), ("py", C.c_void_p), ("pz", C.c_void_p),
                ("position", C.c_double * 3), ("right", C.c_double * 3), ("up", C.c_double * 3),
                ("forward", C.c_double * 3), ("amin", C.c_double * 3), ("amax", C.c_double * 3),
                ("n", C.c_int64 * 3), ("out", C.c_void_p), ("trig", C.c_void_p)]


class ph_tune_entry(C.Structure):
    _fields_ = [("key", C.c_uint64), ("occupancy", C.c_double), ("xcd_map", C.c_int32), ("tile_order", C.c_int32),
                ("best_ms", C.c_double), ("default_ms", C.c_double), ("bucket_leaves", C.c_int32),
                ("block_size", C.c_int32)]


class ph_stats(C.Structure):
    _fields_ = [("build_ms", C.c_double), ("interp_ms", C.c_double), ("kernel_ms", C.c_double),
                ("samples", C.c_int64), ("accepted", C.c_int64), ("rechecks", C.c_int64)]


_lib = None


def lib():
    """Load libpartia_hip.so once and declare every signature."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise NativeLibraryError(
            "libpartia_hip.so not found at %s; build it with `make -C partia.jl_amd` "
            "(or __graft_entry__.build())" % LIB_PATH)
    try:
        L = C.CDLL(LIB_PATH)
    except OSError as e:  # pragma: no cover - depends on the machine
        raise NativeLibraryError("cannot load %s: %s" % (LIB_PATH, e))
    vp, i32, i64, d = C.c_void_p, C.c_int32, C.c_int64, C.c_double
    P = C.POINTER
    sig = {
        "ph_ctx_create": (i32, [C.c_int, P(vp)]),
        "ph_ctx_destroy": (i32, [vp]),
        "ph_last_error": (C.c_char_p, [vp]),
        "ph_get_stats": (i32, [vp, P(ph_stats)]),
        "ph_set_count_neighbours": (i32, [vp, C.c_int]),
        "ph_set_traversal": (i32, [vp, C.c_int]),
        "ph_device_alloc": (i32, [vp, i64, P(vp)]),
        "ph_device_free": (i32, [vp, vp]),
        "ph_memcpy_htod": (i32, [vp, vp, vp, i64]),
        "ph_memcpy_dtoh": (i32, [vp, vp, vp, i64]),
        "ph_synchronize": (i32, [vp]),
        "ph_lbvh_build": (i32, [vp, i64] + [vp] * 6 + [i32, vp, i32, i32, i32, P(vp)]),
        "ph_lbvh_build_device": (i32, [vp, i64] + [vp] * 6 + [i32, vp, i32, i32, i32, P(vp)]),
        "ph_lbvh_build2d": (i32, [vp, i64] + [vp] * 5 + [i32, vp, i32, i32, i32, P(vp)]),
        "ph_lbvh_build2d_device": (i32, [vp, i64] + [vp] * 5 + [i32, vp, i32, i32, i32, P(vp)]),
        "ph_lbvh_dims": (i32, [vp, P(i32)]),
        "ph_query_device": (i32, [vp, vp, i64] + [vp] * 9),
        "ph_find_nearest_device": (i32, [vp, vp, i64] + [vp] * 5),
        "ph_lbvh_destroy": (i32, [vp]),
        "ph_lbvh_set_columns": (i32, [vp, vp, vp, i32, vp]),
        "ph_lbvh_info": (i32, [vp, P(i64), P(i32), P(i32), P(i32), P(i32)]),
        "ph_lbvh_get": (i32, [vp, i32, vp]),
        "ph_point_samples": (i32, [vp, vp, P(ph_catalog), i32, i32, i64, vp, vp, vp, vp]),
        "ph_point_samples_device": (i32, [vp, vp, P(ph_catalog), i32, i32, i64, vp, vp, vp, vp]),
        "ph_point_samples_device_noalloc": (i32, [vp, vp, P(ph_catalog), i32, i32, i64, vp, vp, vp, vp]),
        "ph_point_single": (i32, [vp, vp, i32, i32, i32, i64, vp, vp, vp, vp, vp]),
        "ph_point_single_device": (i32, [vp, vp, i32, i32, i32, i64, vp, vp, vp, vp, vp]),
        "ph_line_samples_ha": (i32, [vp, vp, P(ph_catalog), i32, i32, i64] + [vp] * 8),
        "ph_line_samples_ha_device": (i32, [vp, vp, P(ph_catalog), i32, i32, i64] + [vp] * 8),
        "ph_line_density": (i32, [vp, vp, i32, i32, i64] + [vp] * 8),
        "ph_line_density_device": (i32, [vp, vp, i32, i32, i64] + [vp] * 8),
        "ph_structured_grid": (i32, [vp, vp, P(ph_catalog), i32, i32, vp, i64, vp, i64, vp, i64, vp]),
        "ph_structured_grid_device": (i32, [vp, vp, P(ph_catalog), i32, i32, vp, i64, vp, i64, vp, i64, vp]),
        "ph_line_samples": (i32, [vp, vp, P(ph_catalog), i32, i32, i64] + [vp] * 7),
        "ph_line_samples_device": (i32, [vp, vp, P(ph_catalog), i32, i32, i64] + [vp] * 7),
        "ph_line_grid_device": (i32, [vp, vp, P(ph_catalog), i32, i32, vp, vp, vp, vp, vp, i64, vp, i64, vp]),
        "ph_nearest_h": (i32, [vp, vp, i64, vp, vp, vp, vp]),
        "ph_line_table": (i32, [i32, vp]),
        "ph_batch_create": (i32, [vp, vp, P(ph_catalog), i32, i32, i32, P(ph_request), P(vp)]),
        "ph_batch_run": (i32, [vp]),
        "ph_batch_times": (i32, [vp, vp, vp]),
        "ph_batch_destroy": (i32, [vp]),
        "ph_set_autotune": (i32, [vp, C.c_int]),
        "ph_set_launch": (i32, [vp, d, i32, i32]),
        "ph_set_tile_shape": (i32, [vp, i32, i32]),
        "ph_autotune_report": (i32, [vp, i32, P(ph_tune_entry), P(i32)]),
        "ph_line_samples_adjoint_device": (i32, [vp, vp, P(ph_catalog), i32, i32, i64] + [vp] * 8),
        "ph_line_grid_adjoint_device": (i32, [vp, vp, P(ph_catalog), i32, i32, vp, vp, vp, vp, vp, i64, vp, i64,
                                              vp, vp]),
        "ph_point_samples_adjoint_device": (i32, [vp, vp, P(ph_catalog), i32, i32, i64, vp, vp, vp, vp, vp]),
        "ph_structured_grid_coords": (i32, [vp, vp, P(ph_catalog), i32, i32, i32, vp, i64, vp, i64, vp, i64, vp]),
        "ph_structured_grid_coords_device": (i32, [vp, vp, P(ph_catalog), i32, i32, i32, vp, i64, vp, i64, vp, i64,
                                                   vp]),
        "ph_line_polar_device": (i32, [vp, vp, P(ph_catalog), i32, i32, vp, vp, vp, vp, vp, i64, vp, i64, vp]),
        "ph_line_polar_adjoint_device": (i32, [vp, vp, P(ph_catalog), i32, i32, vp, vp, vp, vp, vp, i64, vp, i64,
                                               vp, vp]),
        "ph_structured_grid_adjoint_device": (i32, [vp, vp, P(ph_catalog), i32, i32, vp, i64, vp, i64, vp, i64,
                                                    vp, vp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def ptr(a):
    """Data pointer of a numpy array (None for None)."""
    if a is None:
        return None
    return C.c_void_p(a.ctypes.data)


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


# Live handles, released before the interpreter exits: batches and trees first,
# then their contexts, while the HIP runtime is still up (objects kept alive until shutdown,
# e.g. by a traceback, would otherwise be freed after the runtime's own teardown).
_trees = weakref.WeakSet()
_contexts = weakref.WeakSet()


def register_tree(obj):
    """Track an object owning a ph_lbvh or ph_batch handle (it must have close())."""
    _trees.add(obj)


@atexit.register
def _release_all():
    for group in (_trees, _contexts):
        for obj in list(group):
            try:
                obj.close()
            except Exception:
                pass


class Context:
    """One GPU + one HIP stream (ph_context)."""

    def __init__(self, device=0):
        L = lib()
        h = C.c_void_p()
        rc = L.ph_ctx_create(int(device), C.byref(h))
        if rc != PH_OK:
            raise_for_status(rc, L.ph_last_error(None).decode())
        self.handle = h
        self.device = device
        _contexts.add(self)

    def check(self, rc):
        if rc != PH_OK:
            raise_for_status(rc, lib().ph_last_error(self.handle).decode())

    def stats(self):
        s = ph_stats()
        self.check(lib().ph_get_stats(self.handle, C.byref(s)))
        return {k: getattr(s, k) for k, _ in ph_stats._fields_}

    def set_count_neighbours(self, enable):
        self.check(lib().ph_set_count_neighbours(self.handle, 1 if enable else 0))

    def set_traversal(self, mode):
        """PH_TRAVERSAL_AUTO, PH_TRAVERSAL_LANE or PH_TRAVERSAL_BUCKET (point and grid calls)."""
        self.check(lib().ph_set_traversal(self.handle, int(mode)))

    def set_autotune(self, enable):
        self.check(lib().ph_set_autotune(self.handle, 1 if enable else 0))

    def set_launch(self, occupancy=1.0, xcd_map=-1, tile_order=-1):
        self.check(lib().ph_set_launch(self.handle, float(occupancy), int(xcd_map), int(tile_order)))

    def set_tile_shape(self, bucket_leaves=0, block_size=0):
        """Grid tile walks: leaves per bucket (4, 8, 16) and threads per block
        (64, 128, 256, 512); 0 restores the default."""
        self.check(lib().ph_set_tile_shape(self.handle, int(bucket_leaves), int(block_size)))

    def autotune_report(self):
        """List of dicts, one per tuned workload key."""
        n = C.c_int32()
        self.check(lib().ph_autotune_report(self.handle, 0, None, C.byref(n)))
        arr = (ph_tune_entry * max(1, n.value))()
        self.check(lib().ph_autotune_report(self.handle, n.value, arr, C.byref(n)))
        return [{k: getattr(arr[i], k) for k, _ in ph_tune_entry._fields_}
                for i in range(n.value)]

    def close(self):
        if getattr(self, "handle", None):
            lib().ph_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class CatalogArgs:
    """Keeps the numpy arrays behind a ph_catalog alive."""

    def __init__(self, sslots=(), sflags=(), gslots=(), dslots=(), cslots=()):
        self.arrays = [np.ascontiguousarray(np.asarray(sslots, dtype=np.int32).ravel()),
                       np.ascontiguousarray(np.asarray(sflags, dtype=np.uint8).ravel()),
                       np.ascontiguousarray(np.asarray(gslots, dtype=np.int32).ravel()),
                       np.ascontiguousarray(np.asarray(dslots, dtype=np.int32).ravel()),
                       np.ascontiguousarray(np.asarray(cslots, dtype=np.int32).ravel())]
        s, f, g, dv, cu = self.arrays
        self.struct = ph_catalog(len(s), s.ctypes.data, f.ctypes.data, len(g), g.ctypes.data,
                                 len(dv) // 3, dv.ctypes.data, len(cu) // 3, cu.ctypes.data)
        self.L = len(s) + 3 * len(g) + len(dv) // 3 + len(cu)

    def ref(self):
        return C.byref(self.struct)
