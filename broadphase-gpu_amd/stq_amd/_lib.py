"""ctypes binding of libstq_amd.so (the C-ABI declared in include/stq.h) and a
minimal HIP runtime binding for device buffers.

The library is loaded from ``broadphase-gpu_amd/lib/libstq_amd.so``; there is
no fallback: if it is missing, loading raises.
"""
from __future__ import annotations

import ctypes
import os

_PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(_PKG, "lib", "libstq_amd.so")

STQ_OK = 0
STQ_ERR_INVALID_ARGUMENT = 1
STQ_ERR_INVALID_BOX = 2
STQ_ERR_CAPACITY = 3
STQ_ERR_OUT_OF_MEMORY = 4
STQ_ERR_HIP = 5
STQ_ERR_TOO_LARGE = 6
STQ_ERR_INTERNAL = 7

PRECISION_F64 = 0
PRECISION_F32_CONFIRM = 1
PRECISION_F64_GROUPED = 2
AXIS_X, AXIS_Y, AXIS_Z, AXIS_VARIANCE, AXIS_CHAIN = 0, 1, 2, 3, 4
SIZING_COUNT_FILL = 0
SIZING_LEARNED = 1

# every entry point of include/stq.h
EXPORTS = [
    "stq_options_default", "stq_create", "stq_destroy", "stq_set_options",
    "stq_find_pairs_device", "stq_find_pairs_device_noalloc", "stq_find_pairs_device_alloc",
    "stq_find_pairs_host", "stq_find_pairs_rows_device", "stq_sort_order_device", "stq_device_memory",
    "stq_radix_sort_device",
    "stq_find_pairs_batched_device", "stq_build_swept_boxes_device", "stq_find_pairs_mesh_device",
    "stq_orient_pairs_device", "stq_ccd_queries_device", "stq_sizeof_ccd_query",
    "stq_mesh_read", "stq_mesh_edges", "stq_recall",
    "stq_autotune_save", "stq_autotune_load",
    "stq_autotune_clear", "stq_autotune_entries", "stq_device_alloc", "stq_copy_to_device",
    "stq_copy_to_host", "stq_free_device", "stq_free_host",
    "stq_sizer_num_params", "stq_sizer_num_features", "stq_sizer_get_params",
    "stq_sizer_set_params", "stq_sizer_reset", "stq_sizer_forward", "stq_sizer_gradient",
    "stq_sizer_train", "stq_sizer_last_items",
    "stq_get_stats", "stq_last_error", "stq_status_string",
    "stq_sizeof_aabb", "stq_sizeof_options", "stq_sizeof_stats",
]


# stq_ccd_query (include/stq.h): the reference's CCDData layout, 256 bytes.
def _ccd_dtype():
    import numpy as np
    return np.dtype([("vs", "<f8", (4, 3)), ("ve", "<f8", (4, 3)), ("ms", "<f8"), ("err", "<f8", (3,)),
                     ("tol", "<f8", (3,)), ("nbr_checks", "<i4"), ("pad", "<i4")])


CCD_QUERY_DTYPE = _ccd_dtype()


class Options(ctypes.Structure):
    _fields_ = [
        ("precision", ctypes.c_int32),
        ("axis_policy", ctypes.c_int32),
        ("sizing", ctypes.c_int32),
        ("autotune", ctypes.c_int32),
        ("chunk", ctypes.c_int32),
        ("window", ctypes.c_int32),
        ("waves_per_cu", ctypes.c_int32),
        ("batch", ctypes.c_int32),
        ("xcd_split", ctypes.c_int32),
        ("train_scenes", ctypes.c_int32),
        ("graph", ctypes.c_int32),
        ("claim", ctypes.c_int32),
        ("tails", ctypes.c_int32),
        ("tail_cost", ctypes.c_int32),
        ("scratch_limit_mb", ctypes.c_int32),
        ("balance", ctypes.c_int32),
    ]


class Stats(ctypes.Structure):
    _fields_ = [
        ("n_boxes", ctypes.c_int64),
        ("n_pairs", ctypes.c_int64),
        ("n_candidates", ctypes.c_int64),
        ("n_lane_steps", ctypes.c_int64),
        ("n_items", ctypes.c_int64),
        ("n_confirm", ctypes.c_int64),
        ("n_tiles", ctypes.c_int64),
        ("n_fallback_tiles", ctypes.c_int64),
        ("n_scenes", ctypes.c_int64),
        ("axis", ctypes.c_int32),
        ("precision", ctypes.c_int32),
        ("sizing", ctypes.c_int32),
        ("graph", ctypes.c_int32),
        ("ms_total", ctypes.c_double),
        ("ms_keys", ctypes.c_double),
        ("ms_sort", ctypes.c_double),
        ("ms_gather", ctypes.c_double),
        ("ms_plan", ctypes.c_double),
        ("ms_sweep_count", ctypes.c_double),
        ("ms_scan", ctypes.c_double),
        ("ms_sweep_fill", ctypes.c_double),
        ("ms_compact", ctypes.c_double),
        ("ms_predict", ctypes.c_double),
        ("sizing_mae", ctypes.c_double),
        ("ms_tune", ctypes.c_double),
        ("reserved", ctypes.c_double * 5),
    ]

    def as_dict(self):
        return {k: (getattr(self, k)) for k, _ in self._fields_ if k not in ("reserved", "pad")}


_LIB = None


def lib():
    """Load libstq_amd.so (raises if it has not been built)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with __graft_entry__.build()")
    L = ctypes.CDLL(LIB_PATH)
    P, I32, I64 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64
    sig = {
        "stq_options_default": (ctypes.c_int, [ctypes.POINTER(Options)]),
        "stq_create": (ctypes.c_int, [ctypes.POINTER(P), ctypes.c_int, ctypes.POINTER(Options)]),
        "stq_destroy": (ctypes.c_int, [P]),
        "stq_set_options": (ctypes.c_int, [P, ctypes.POINTER(Options)]),
        "stq_find_pairs_device": (ctypes.c_int, [P, P, I64, P, I64, ctypes.POINTER(I64)]),
        "stq_find_pairs_device_noalloc": (ctypes.c_int, [P, P, I64, P, I64, ctypes.POINTER(I64)]),
        "stq_find_pairs_device_alloc": (ctypes.c_int, [P, P, I64, ctypes.POINTER(P), ctypes.POINTER(I64)]),
        "stq_find_pairs_host": (ctypes.c_int, [P, P, I64, ctypes.POINTER(P), ctypes.POINTER(I64)]),
        "stq_find_pairs_rows_device": (ctypes.c_int, [P, P, I64, I64, I64, P, I64, ctypes.POINTER(I64)]),
        "stq_sort_order_device": (ctypes.c_int, [P, P, I64, P]),
        "stq_device_memory": (ctypes.c_int, [P, ctypes.POINTER(ctypes.c_size_t),
                                             ctypes.POINTER(ctypes.c_size_t)]),
        "stq_radix_sort_device": (ctypes.c_int, [P, P, P, I64, I32, I32, I32]),
        "stq_find_pairs_batched_device": (ctypes.c_int, [P, P, P, I32, P, I64, P]),
        "stq_build_swept_boxes_device": (ctypes.c_int, [P, P, P, I64, P, I64, P, I64, ctypes.c_double,
                                                         I32, P]),
        "stq_find_pairs_mesh_device": (ctypes.c_int, [P, P, P, I64, P, I64, P, I64, ctypes.c_double, P,
                                                       I64, ctypes.POINTER(I64)]),
        "stq_orient_pairs_device": (ctypes.c_int, [P, P, I64, P, I64]),
        "stq_ccd_queries_device": (ctypes.c_int, [P, P, I64, P, P, I64, P, I64, ctypes.c_double, P]),
        "stq_sizeof_ccd_query": (ctypes.c_int, []),
        "stq_mesh_read": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(P), ctypes.POINTER(I64),
                                         ctypes.POINTER(P), ctypes.POINTER(I64)]),
        "stq_mesh_edges": (ctypes.c_int, [P, I64, ctypes.POINTER(P), ctypes.POINTER(I64)]),
        "stq_recall": (ctypes.c_int, [P, I64, ctypes.c_char_p, I32, ctypes.POINTER(I64), ctypes.POINTER(I64)]),
        "stq_autotune_save": (ctypes.c_int, [P, ctypes.c_char_p]),
        "stq_autotune_load": (ctypes.c_int, [P, ctypes.c_char_p]),
        "stq_autotune_clear": (ctypes.c_int, [P]),
        "stq_autotune_entries": (ctypes.c_int, [P, ctypes.POINTER(I32)]),
        "stq_device_alloc": (ctypes.c_int, [P, ctypes.c_size_t, ctypes.POINTER(P)]),
        "stq_copy_to_device": (ctypes.c_int, [P, P, P, ctypes.c_size_t]),
        "stq_copy_to_host": (ctypes.c_int, [P, P, P, ctypes.c_size_t]),
        "stq_free_device": (ctypes.c_int, [P, P]),
        "stq_free_host": (None, [P]),
        "stq_sizer_num_params": (ctypes.c_int, []),
        "stq_sizer_num_features": (ctypes.c_int, []),
        "stq_sizer_get_params": (ctypes.c_int, [P, P]),
        "stq_sizer_set_params": (ctypes.c_int, [P, P]),
        "stq_sizer_reset": (ctypes.c_int, [P, ctypes.c_uint64]),
        "stq_sizer_forward": (ctypes.c_int, [P, P, I64, P]),
        "stq_sizer_gradient": (ctypes.c_int, [P, P, P, I64, P, P]),
        "stq_sizer_train": (ctypes.c_int, [P, P, P, I64, I32, ctypes.c_float]),
        "stq_sizer_last_items": (ctypes.c_int, [P, P, P, I64, ctypes.POINTER(I64)]),
        "stq_get_stats": (ctypes.c_int, [P, ctypes.POINTER(Stats)]),
        "stq_last_error": (ctypes.c_char_p, [P]),
        "stq_status_string": (ctypes.c_char_p, [ctypes.c_int]),
        "stq_sizeof_aabb": (ctypes.c_int, []),
        "stq_sizeof_options": (ctypes.c_int, []),
        "stq_sizeof_stats": (ctypes.c_int, []),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    assert L.stq_sizeof_aabb() == 72
    assert L.stq_sizeof_options() == ctypes.sizeof(Options)
    assert L.stq_sizeof_stats() == ctypes.sizeof(Stats)
    assert L.stq_sizeof_ccd_query() == CCD_QUERY_DTYPE.itemsize
    _LIB = L
    return L


# ---- HIP runtime (device buffers for callers that hold boxes on the GPU) ----
_HIP = None
HIP_H2D, HIP_D2H, HIP_D2D = 1, 2, 3


def hip():
    global _HIP
    if _HIP is None:
        H = ctypes.CDLL("libamdhip64.so")
        P, S = ctypes.c_void_p, ctypes.c_size_t
        H.hipMalloc.argtypes = [ctypes.POINTER(P), S]
        H.hipMalloc.restype = ctypes.c_int
        H.hipFree.argtypes = [P]
        H.hipFree.restype = ctypes.c_int
        H.hipMemcpy.argtypes = [P, P, S, ctypes.c_int]
        H.hipMemcpy.restype = ctypes.c_int
        H.hipDeviceSynchronize.restype = ctypes.c_int
        H.hipGetDeviceCount.argtypes = [ctypes.POINTER(ctypes.c_int)]
        H.hipGetDeviceCount.restype = ctypes.c_int
        H.hipSetDevice.argtypes = [ctypes.c_int]
        H.hipSetDevice.restype = ctypes.c_int
        _HIP = H
    return _HIP


def _check_hip(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what} failed with hipError {rc}")


def device_count():
    n = ctypes.c_int(0)
    rc = hip().hipGetDeviceCount(ctypes.byref(n))
    return n.value if rc == 0 else 0


class DeviceBuffer:
    """A hipMalloc'd buffer; freed on close()/garbage collection."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        self.ptr = ctypes.c_void_p()
        _check_hip(hip().hipMalloc(ctypes.byref(self.ptr), max(self.nbytes, 1)), "hipMalloc")

    @classmethod
    def from_array(cls, arr):
        import numpy as np
        a = np.ascontiguousarray(arr)
        b = cls(a.nbytes)
        if a.nbytes:
            _check_hip(hip().hipMemcpy(b.ptr, a.ctypes.data, a.nbytes, HIP_H2D), "hipMemcpy H2D")
        return b

    def write(self, arr, offset=0):
        """Copies a host array into the buffer at a byte offset."""
        import numpy as np
        a = np.ascontiguousarray(arr)
        assert offset + a.nbytes <= self.nbytes
        if a.nbytes:
            dst = ctypes.c_void_p(self.ptr.value + offset)
            _check_hip(hip().hipMemcpy(dst, a.ctypes.data, a.nbytes, HIP_H2D), "hipMemcpy H2D")

    def to_array(self, dtype, count):
        import numpy as np
        out = np.empty(count, dtype=dtype)
        if out.nbytes:
            _check_hip(hip().hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, HIP_D2H), "hipMemcpy D2H")
        return out

    @property
    def value(self):
        return self.ptr.value

    def close(self):
        if self.ptr is not None and self.ptr.value:
            hip().hipFree(self.ptr)
        self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
