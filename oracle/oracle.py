"""ctypes front of the CPU oracle (liboracle.so built from stq_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check and by the
CPU-baseline leg of bench.py, never by the product library.  Every function
here restates the reference's CPU broad phase (src/cpu/sweep.cpp:17-158) or
its FP32 rounding rule (src/aabb.cu:56-61); the C source cites each line.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

AABB_DTYPE = np.dtype(
    [("id", "<i4"), ("pad0", "<i4"), ("min", "<f8", (3,)), ("max", "<f8", (3,)),
     ("vids", "<i4", (3,)), ("ref_id", "<i4")])


def build():
    subprocess.run(["make", "-s", "-B", "-C", _HERE], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        P = ctypes.c_void_p
        I64 = ctypes.c_int64
        L.oracle_sweep_rows.restype = I64
        L.oracle_sweep_rows.argtypes = [P, I64, ctypes.c_int, I64, ctypes.c_int,
                                        ctypes.POINTER(ctypes.c_void_p),
                                        ctypes.POINTER(ctypes.c_double)]
        L.oracle_sweep_timed.restype = I64
        L.oracle_sweep_timed.argtypes = [P, I64, ctypes.c_int, ctypes.c_int,
                                         ctypes.POINTER(ctypes.c_void_p),
                                         ctypes.POINTER(ctypes.c_double)]
        L.oracle_bruteforce.restype = I64
        L.oracle_bruteforce.argtypes = [P, I64, ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]
        L.oracle_candidates.restype = I64
        L.oracle_candidates.argtypes = [P, I64, ctypes.c_int, ctypes.c_int,
                                        ctypes.POINTER(I64)]
        L.oracle_sort_u64.restype = None
        L.oracle_sort_u64.argtypes = [P, I64, ctypes.c_int]
        L.oracle_canonicalize.restype = I64
        L.oracle_canonicalize.argtypes = [P, I64]
        L.oracle_round_outward.restype = None
        L.oracle_round_outward.argtypes = [P, I64, P, P]
        L.oracle_boxes_round_outward.restype = None
        L.oracle_boxes_round_outward.argtypes = [P, I64, P]
        L.oracle_pair_predicate.restype = ctypes.c_int
        L.oracle_pair_predicate.argtypes = [P, P]
        L.oracle_free.restype = None
        L.oracle_free.argtypes = [P]
        L.oracle_build_swept_boxes.restype = ctypes.c_int
        L.oracle_build_swept_boxes.argtypes = [P, P, I64, P, I64, P, I64, ctypes.c_double,
                                               ctypes.c_int, P]
        L.oracle_orient_pairs.restype = ctypes.c_int
        L.oracle_orient_pairs.argtypes = [P, I64, P, I64]
        L.oracle_sizeof_aabb.restype = ctypes.c_int
        L.oracle_max_threads.restype = ctypes.c_int
        assert L.oracle_sizeof_aabb() == 72
        _LIB = L
    return _LIB


def _as_boxes(boxes):
    b = np.ascontiguousarray(boxes)
    assert b.dtype.itemsize == 72
    return b


def _take(ptr, n):
    if n == 0:
        lib().oracle_free(ptr)
        return np.zeros((0, 2), dtype=np.int32)
    arr = np.ctypeslib.as_array(ctypes.cast(ptr, ctypes.POINTER(ctypes.c_int32)), shape=(n, 2)).copy()
    lib().oracle_free(ptr)
    return arr


def default_threads():
    return max(1, min(os.cpu_count() or 1, 64))


def sweep(boxes, axis=0, nthreads=None, nrows=None, return_sort_time=False):
    """Reference CPU sweep (sort by axis min, chain scan with break, filters).
    Returns (earlier, later) position pairs in sorted-row order."""
    b = _as_boxes(boxes)
    n = len(b)
    out = ctypes.c_void_p()
    st = ctypes.c_double()
    cnt = lib().oracle_sweep_rows(b.ctypes.data, n, axis, n if nrows is None else nrows,
                                  nthreads or default_threads(), ctypes.byref(out), ctypes.byref(st))
    res = _take(out, cnt)
    return (res, st.value) if return_sort_time else res


def sweep_timed(boxes, axis=0, nthreads=None):
    """(pairs, seconds): the whole-scene sweep and its wall time measured
    inside the C code (sort, sweep, merge; the copy to numpy excluded)."""
    b = _as_boxes(boxes)
    out = ctypes.c_void_p()
    secs = ctypes.c_double()
    cnt = lib().oracle_sweep_timed(b.ctypes.data, len(b), axis, nthreads or default_threads(),
                                   ctypes.byref(out), ctypes.byref(secs))
    return _take(out, cnt), secs.value


def bruteforce(boxes, nthreads=None):
    b = _as_boxes(boxes)
    out = ctypes.c_void_p()
    cnt = lib().oracle_bruteforce(b.ctypes.data, len(b), nthreads or default_threads(), ctypes.byref(out))
    return _take(out, cnt)


def candidates(boxes, axis=0, nthreads=None):
    """(total sweep candidates, longest chain) on an axis."""
    b = _as_boxes(boxes)
    mc = ctypes.c_int64()
    tot = lib().oracle_candidates(b.ctypes.data, len(b), axis, nthreads or default_threads(), ctypes.byref(mc))
    return int(tot), int(mc.value)


def canonical(pairs):
    """(min,max) orientation, lexicographic order, duplicates removed."""
    p = np.ascontiguousarray(np.asarray(pairs, dtype=np.int32).reshape(-1, 2)).copy()
    m = lib().oracle_canonicalize(p.ctypes.data, len(p))
    return p[:m].copy()


def sort_u64(keys, nthreads=None):
    """Ascending copy of uint64 keys (the radix sort under canonical())."""
    k = np.ascontiguousarray(np.asarray(keys, dtype=np.uint64)).copy()
    lib().oracle_sort_u64(k.ctypes.data, len(k), nthreads or default_threads())
    return k


def round_outward(values):
    v = np.ascontiguousarray(np.asarray(values, dtype=np.float64))
    lo = np.empty(v.shape, dtype=np.float32)
    hi = np.empty(v.shape, dtype=np.float32)
    lib().oracle_round_outward(v.ctypes.data, v.size, lo.ctypes.data, hi.ctypes.data)
    return lo, hi


def boxes_round_outward(boxes):
    b = _as_boxes(boxes)
    out = np.empty_like(b)
    lib().oracle_boxes_round_outward(b.ctypes.data, len(b), out.ctypes.data)
    return out


def pair_predicate(boxes, i, j):
    b = _as_boxes(boxes)
    return bool(lib().oracle_pair_predicate(b[i:i + 1].ctypes.data, b[j:j + 1].ctypes.data))


def reference_set(boxes, axis=0, nthreads=None):
    """The canonical FP64 pair set of the reference sweep."""
    return canonical(sweep(boxes, axis=axis, nthreads=nthreads))


def build_swept_boxes(V0, V1, E, F, radius=0.0, float_build=False):
    """Swept boxes of a mesh (vertices, then edges, then faces) as the
    reference's constructBoxes builds them (src/io.cu:15-28,
    src/aabb.cu:56-174); float_build reproduces its float build."""
    from stq_amd.scenes import AABB_DTYPE
    v0 = np.ascontiguousarray(V0, dtype=np.float64).reshape(-1, 3)
    v1 = np.ascontiguousarray(V1, dtype=np.float64).reshape(-1, 3)
    e = np.ascontiguousarray(E, dtype=np.int32).reshape(-1, 2)
    f = np.ascontiguousarray(F, dtype=np.int32).reshape(-1, 3)
    out = np.zeros(len(v0) + len(e) + len(f), dtype=AABB_DTYPE)
    rc = lib().oracle_build_swept_boxes(v0.ctypes.data, v1.ctypes.data, len(v0), e.ctypes.data, len(e),
                                        f.ctypes.data, len(f), float(radius), int(bool(float_build)),
                                        out.ctypes.data)
    if rc != 0:
        raise ValueError("edge or face index outside the vertex range")
    return out


def orient_pairs(boxes, pairs):
    """Pairs oriented as the reference hands them to the narrow phase
    (src/simulation.cu:510-524): (vertex, face); edge-edge (min, max)."""
    b = _as_boxes(boxes)
    p = np.ascontiguousarray(np.asarray(pairs, dtype=np.int32).reshape(-1, 2)).copy()
    if lib().oracle_orient_pairs(b.ctypes.data, len(b), p.ctypes.data, len(p)) != 0:
        raise ValueError("pair of another kind than vertex-face or edge-edge, or id out of range")
    return p
