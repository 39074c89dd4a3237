"""Python front of the gfx950 STQ broad phase (libstq_amd.so).

Mirrors the reference entry ``stq::gpu::run_sweep_sharedqueue``
(src/simulation.cuh:22-27): boxes in the reference FP64 ``Aabb`` layout go in,
the overlapping box-id pairs come out.  All work happens in the HIP library;
this module only moves numpy arrays across the C-ABI.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from . import _lib
from ._lib import (AXIS_CHAIN, AXIS_VARIANCE, AXIS_X, AXIS_Y, AXIS_Z,  # noqa: F401
                   PRECISION_F32_CONFIRM, PRECISION_F64, PRECISION_F64_GROUPED, SIZING_COUNT_FILL,
                   SIZING_LEARNED, DeviceBuffer, Options, Stats)
from .scenes import AABB_DTYPE  # noqa: F401

PAIR_DTYPE = np.dtype([("first", "<i4"), ("second", "<i4")])


class StqError(RuntimeError):
    def __init__(self, status, message):
        super().__init__(f"{message} (status {status})")
        self.status = status


def options(precision=PRECISION_F64, axis_policy=AXIS_X, sizing=SIZING_COUNT_FILL,
            autotune=0, chunk=0, window=0, waves_per_cu=0, batch=0, xcd_split=0,
            train_scenes=0, graph=0, claim=0, tails=0, tail_cost=0, scratch_limit_mb=0,
            balance=0) -> Options:
    o = Options()
    _lib.lib().stq_options_default(ctypes.byref(o))
    o.precision, o.axis_policy, o.sizing = precision, axis_policy, sizing
    o.autotune, o.chunk, o.window, o.waves_per_cu = autotune, chunk, window, waves_per_cu
    o.batch, o.xcd_split, o.train_scenes = batch, xcd_split, train_scenes
    o.graph, o.claim, o.tails, o.tail_cost = graph, claim, tails, tail_cost
    o.scratch_limit_mb, o.balance = scratch_limit_mb, balance
    return o


class Context:
    """One library context: a device, a stream, scratch and tuning state."""

    def __init__(self, device=0, opts: Options | None = None, **kw):
        L = _lib.lib()
        self._L = L
        self.ptr = ctypes.c_void_p()
        o = opts if opts is not None else options(**kw)
        rc = L.stq_create(ctypes.byref(self.ptr), device, ctypes.byref(o))
        if rc != 0:
            raise StqError(rc, "stq_create: " + L.stq_status_string(rc).decode())

    def close(self):
        if self.ptr is not None and self.ptr.value:
            self._L.stq_destroy(self.ptr)
        self.ptr = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc, what):
        if rc != 0:
            msg = self._L.stq_last_error(self.ptr).decode()
            raise StqError(rc, f"{what}: {self._L.stq_status_string(rc).decode()}: {msg}")

    def set_options(self, opts: Options | None = None, **kw):
        o = opts if opts is not None else options(**kw)
        self._check(self._L.stq_set_options(self.ptr, ctypes.byref(o)), "stq_set_options")

    # -- meshes ----------------------------------------------------------------
    @staticmethod
    def _mesh_buffers(V0, V1, E, F):
        v0 = np.ascontiguousarray(V0, dtype=np.float64).reshape(-1, 3)
        v1 = np.ascontiguousarray(V1, dtype=np.float64).reshape(-1, 3)
        e = np.ascontiguousarray(E, dtype=np.int32).reshape(-1, 2)
        f = np.ascontiguousarray(F, dtype=np.int32).reshape(-1, 3)
        assert v0.shape == v1.shape
        bufs = [DeviceBuffer.from_array(a) for a in (v0, v1, e, f)]
        return bufs, (len(v0), len(e), len(f))

    def build_swept_boxes(self, V0, V1, E, F, radius=0.0, float_build=False) -> np.ndarray:
        """Swept boxes of a moving mesh built on the GPU (vertices, edges, faces)."""
        bufs, (nv, ne, nf) = self._mesh_buffers(V0, V1, E, F)
        n = nv + ne + nf
        out = DeviceBuffer(72 * max(n, 1))
        try:
            self._check(self._L.stq_build_swept_boxes_device(
                self.ptr, bufs[0].ptr, bufs[1].ptr, nv, bufs[2].ptr, ne, bufs[3].ptr, nf, float(radius),
                int(bool(float_build)), out.ptr), "stq_build_swept_boxes_device")
            return out.to_array(AABB_DTYPE, n)
        finally:
            out.close()
            for b in bufs:
                b.close()

    # device-resident forms (device pointers as ints); stats() then holds the
    # kernel time on the stream in ms_total
    def build_swept_boxes_device(self, d_v0, d_v1, nv, d_e, ne, d_f, nf, d_boxes, radius=0.0,
                                 float_build=False):
        self._check(self._L.stq_build_swept_boxes_device(
            self.ptr, d_v0, d_v1, nv, d_e, ne, d_f, nf, float(radius), int(bool(float_build)), d_boxes),
            "stq_build_swept_boxes_device")

    def orient_pairs_device(self, d_boxes, n, d_pairs, count):
        self._check(self._L.stq_orient_pairs_device(self.ptr, d_boxes, n, d_pairs, count),
                    "stq_orient_pairs_device")

    def ccd_queries_device(self, d_boxes, n, d_v0, d_v1, nv, d_pairs, count, d_queries,
                           min_separation=0.0):
        self._check(self._L.stq_ccd_queries_device(self.ptr, d_boxes, n, d_v0, d_v1, nv, d_pairs, count,
                                                   float(min_separation), d_queries),
                    "stq_ccd_queries_device")

    def find_pairs_mesh(self, V0, V1, E, F, radius=0.0) -> np.ndarray:
        """Pairs of the mesh's swept boxes (ids = box positions: vertices,
        then edges, then faces)."""
        bufs, (nv, ne, nf) = self._mesh_buffers(V0, V1, E, F)
        cap = 1 << 20
        try:
            while True:
                dout = DeviceBuffer(8 * cap)
                cnt = ctypes.c_int64(0)
                rc = self._L.stq_find_pairs_mesh_device(self.ptr, bufs[0].ptr, bufs[1].ptr, nv, bufs[2].ptr,
                                                        ne, bufs[3].ptr, nf, float(radius), dout.ptr, cap,
                                                        ctypes.byref(cnt))
                if rc == _lib.STQ_ERR_CAPACITY:
                    cap = int(cnt.value)
                    dout.close()
                    continue
                self._check(rc, "stq_find_pairs_mesh_device")
                res = dout.to_array(np.int32, 2 * cnt.value).reshape(-1, 2)
                dout.close()
                return res
        finally:
            for b in bufs:
                b.close()

    def orient_pairs(self, boxes, pairs) -> np.ndarray:
        """Pairs oriented for the narrow phase: (vertex, face), edge-edge (min, max)."""
        b = np.ascontiguousarray(boxes)
        p = np.ascontiguousarray(pairs, dtype=np.int32).reshape(-1, 2)
        db, dp = DeviceBuffer.from_array(b), DeviceBuffer.from_array(p)
        try:
            self._check(self._L.stq_orient_pairs_device(self.ptr, db.ptr, len(b), dp.ptr, len(p)),
                        "stq_orient_pairs_device")
            return dp.to_array(np.int32, 2 * len(p)).reshape(-1, 2)
        finally:
            db.close()
            dp.close()

    def ccd_queries(self, boxes, V0, V1, pairs, min_separation=0.0) -> np.ndarray:
        """Narrow-phase query records (CCD_QUERY_DTYPE) of the pairs."""
        b = np.ascontiguousarray(boxes)
        v0 = np.ascontiguousarray(V0, dtype=np.float64).reshape(-1, 3)
        v1 = np.ascontiguousarray(V1, dtype=np.float64).reshape(-1, 3)
        p = np.ascontiguousarray(pairs, dtype=np.int32).reshape(-1, 2)
        bufs = [DeviceBuffer.from_array(a) for a in (b, v0, v1, p)]
        out = DeviceBuffer(_lib.CCD_QUERY_DTYPE.itemsize * max(len(p), 1))
        try:
            self._check(self._L.stq_ccd_queries_device(self.ptr, bufs[0].ptr, len(b), bufs[1].ptr, bufs[2].ptr,
                                                       len(v0), bufs[3].ptr, len(p), float(min_separation),
                                                       out.ptr), "stq_ccd_queries_device")
            return out.to_array(_lib.CCD_QUERY_DTYPE, len(p))
        finally:
            out.close()
            for x in bufs:
                x.close()

    # -- autotuning table ------------------------------------------------------
    def autotune_save(self, path):
        self._check(self._L.stq_autotune_save(self.ptr, os.fsencode(path)), "stq_autotune_save")

    def autotune_load(self, path):
        self._check(self._L.stq_autotune_load(self.ptr, os.fsencode(path)), "stq_autotune_load")

    def autotune_clear(self):
        self._check(self._L.stq_autotune_clear(self.ptr), "stq_autotune_clear")

    def autotune_entries(self) -> int:
        n = ctypes.c_int32(0)
        self._check(self._L.stq_autotune_entries(self.ptr, ctypes.byref(n)), "stq_autotune_entries")
        return n.value

    def stats(self) -> dict:
        s = Stats()
        self._check(self._L.stq_get_stats(self.ptr, ctypes.byref(s)), "stq_get_stats")
        return s.as_dict()

    # -- device-resident form -------------------------------------------------
    def find_pairs_device(self, d_boxes, n, d_out, capacity):
        """Device boxes in, device pairs out (caller buffer). Returns the count."""
        cnt = ctypes.c_int64(0)
        rc = self._L.stq_find_pairs_device(self.ptr, d_boxes, n, d_out, capacity, ctypes.byref(cnt))
        if rc == _lib.STQ_ERR_CAPACITY:
            raise StqError(rc, f"capacity {capacity} < {cnt.value} pairs")
        self._check(rc, "stq_find_pairs_device")
        return cnt.value

    def find_pairs(self, boxes: np.ndarray) -> np.ndarray:
        """Host numpy boxes in; (P, 2) int32 canonical pairs out (device path,
        library-allocated output, like the reference's d_overlaps)."""
        b = np.ascontiguousarray(boxes)
        assert b.dtype.itemsize == 72
        db = DeviceBuffer.from_array(b)
        out = ctypes.c_void_p()
        cnt = ctypes.c_int64(0)
        try:
            self._check(self._L.stq_find_pairs_device_alloc(self.ptr, db.ptr, len(b), ctypes.byref(out),
                                                            ctypes.byref(cnt)),
                        "stq_find_pairs_device_alloc")
            res = np.empty((cnt.value, 2), dtype=np.int32)
            if cnt.value:
                _lib._check_hip(_lib.hip().hipMemcpy(res.ctypes.data, out, res.nbytes, _lib.HIP_D2H),
                                "hipMemcpy D2H")
        finally:
            if out.value:
                self._L.stq_free_device(self.ptr, out)
            db.close()
        return res

    # -- the learned sizing network ---------------------------------------------
    def sizer_params(self) -> np.ndarray:
        w = np.empty(self._L.stq_sizer_num_params(), dtype=np.float32)
        self._check(self._L.stq_sizer_get_params(self.ptr, w.ctypes.data), "stq_sizer_get_params")
        return w

    def sizer_set_params(self, w):
        w = np.ascontiguousarray(w, dtype=np.float32)
        assert w.size == self._L.stq_sizer_num_params()
        self._check(self._L.stq_sizer_set_params(self.ptr, w.ctypes.data), "stq_sizer_set_params")

    def sizer_reset(self, seed=0x5717):
        self._check(self._L.stq_sizer_reset(self.ptr, seed), "stq_sizer_reset")

    def sizer_forward(self, feat) -> np.ndarray:
        f = np.ascontiguousarray(feat, dtype=np.float32)
        out = np.empty(len(f), dtype=np.float32)
        self._check(self._L.stq_sizer_forward(self.ptr, f.ctypes.data, len(f), out.ctypes.data),
                    "stq_sizer_forward")
        return out

    def sizer_gradient(self, feat, targets):
        f = np.ascontiguousarray(feat, dtype=np.float32)
        t = np.ascontiguousarray(targets, dtype=np.float32)
        g = np.empty(self._L.stq_sizer_num_params(), dtype=np.float32)
        loss = ctypes.c_float()
        self._check(self._L.stq_sizer_gradient(self.ptr, f.ctypes.data, t.ctypes.data, len(f),
                                               g.ctypes.data, ctypes.byref(loss)), "stq_sizer_gradient")
        return float(loss.value), g

    def sizer_train(self, feat, targets, steps, lr=1e-3):
        f = np.ascontiguousarray(feat, dtype=np.float32)
        t = np.ascontiguousarray(targets, dtype=np.float32)
        self._check(self._L.stq_sizer_train(self.ptr, f.ctypes.data, t.ctypes.data, len(f), steps, lr),
                    "stq_sizer_train")

    def sizer_last_items(self):
        """(features (T,32), exact pair counts (T,)) of the last learned-sizing call."""
        n = ctypes.c_int64()
        self._check(self._L.stq_sizer_last_items(self.ptr, None, None, 0, ctypes.byref(n)),
                    "stq_sizer_last_items")
        f = np.empty((n.value, self._L.stq_sizer_num_features()), dtype=np.float32)
        c = np.empty(n.value, dtype=np.uint32)
        self._check(self._L.stq_sizer_last_items(self.ptr, f.ctypes.data, c.ctypes.data, n.value,
                                                 ctypes.byref(n)), "stq_sizer_last_items")
        return f, c

    def find_pairs_host(self, boxes: np.ndarray) -> np.ndarray:
        b = np.ascontiguousarray(boxes)
        out = ctypes.c_void_p()
        cnt = ctypes.c_int64(0)
        self._check(self._L.stq_find_pairs_host(self.ptr, b.ctypes.data, len(b), ctypes.byref(out),
                                                ctypes.byref(cnt)), "stq_find_pairs_host")
        try:
            if cnt.value:
                res = np.ctypeslib.as_array(ctypes.cast(out, ctypes.POINTER(ctypes.c_int32)),
                                            shape=(cnt.value, 2)).copy()
            else:
                res = np.zeros((0, 2), dtype=np.int32)
        finally:
            self._L.stq_free_host(out)
        return res

    def find_pairs_batched_device(self, d_boxes, scene_offsets, d_out, capacity):
        """Device-resident batched call.  Returns (status, pair_offsets); the
        status is STQ_OK or STQ_ERR_CAPACITY (pair_offsets[-1] then holds the
        number needed), anything else raises."""
        offs = np.ascontiguousarray(scene_offsets, dtype=np.int64)
        ns = len(offs) - 1
        poffs = np.zeros(ns + 1, dtype=np.int64)
        rc = self._L.stq_find_pairs_batched_device(self.ptr, d_boxes, offs.ctypes.data, ns, d_out,
                                                   capacity, poffs.ctypes.data)
        if rc != _lib.STQ_ERR_CAPACITY:
            self._check(rc, "stq_find_pairs_batched_device")
        return rc, poffs

    def find_pairs_batched(self, boxes: np.ndarray, scene_offsets: np.ndarray, capacity=None):
        """Scenes back to back; returns (pairs (P,2), pair_offsets (S+1))."""
        b = np.ascontiguousarray(boxes)
        offs = np.ascontiguousarray(scene_offsets, dtype=np.int64)
        ns = len(offs) - 1
        poffs = np.zeros(ns + 1, dtype=np.int64)
        db = DeviceBuffer.from_array(b)
        cap = int(capacity) if capacity is not None else 1 << 20
        try:
            while True:
                dout = DeviceBuffer(8 * max(cap, 1))
                rc = self._L.stq_find_pairs_batched_device(self.ptr, db.ptr, offs.ctypes.data, ns, dout.ptr,
                                                           cap, poffs.ctypes.data)
                if rc == _lib.STQ_ERR_CAPACITY and capacity is None:
                    cap = int(poffs[-1])
                    dout.close()
                    continue
                self._check(rc, "stq_find_pairs_batched_device")
                res = dout.to_array(np.int32, 2 * int(poffs[-1])).reshape(-1, 2)
                dout.close()
                return res, poffs
        finally:
            db.close()


def _take_host(ptr, dtype, count):
    """Copies `count` items from a library-malloc'd block and frees it."""
    L = _lib.lib()
    try:
        if count and ptr.value:
            return np.ctypeslib.as_array(ctypes.cast(ptr, ctypes.POINTER(np.ctypeslib.as_ctypes_type(dtype))),
                                         shape=(count,)).copy()
        return np.zeros(0, dtype=dtype)
    finally:
        L.stq_free_host(ptr)


def read_mesh(path):
    """(V, F) of an OBJ or PLY triangle mesh (the reference's parseMesh input)."""
    L = _lib.lib()
    v, f = ctypes.c_void_p(), ctypes.c_void_p()
    nv, nf = ctypes.c_int64(0), ctypes.c_int64(0)
    rc = L.stq_mesh_read(os.fsencode(path), ctypes.byref(v), ctypes.byref(nv), ctypes.byref(f), ctypes.byref(nf))
    if rc != 0:
        raise StqError(rc, f"stq_mesh_read: {L.stq_status_string(rc).decode()} ({path})")
    V = _take_host(v, np.float64, 3 * nv.value).reshape(-1, 3)
    F = _take_host(f, np.int32, 3 * nf.value).reshape(-1, 3)
    return V, F


def mesh_edges(F):
    """Unique undirected edges of a triangle list, in igl::edges order."""
    L = _lib.lib()
    f = np.ascontiguousarray(F, dtype=np.int32).reshape(-1, 3)
    e = ctypes.c_void_p()
    ne = ctypes.c_int64(0)
    rc = L.stq_mesh_edges(f.ctypes.data, len(f), ctypes.byref(e), ctypes.byref(ne))
    if rc != 0:
        raise StqError(rc, "stq_mesh_edges: " + L.stq_status_string(rc).decode())
    return _take_host(e, np.int32, 2 * ne.value).reshape(-1, 2)


def recall(pairs, json_path, canonical=False):
    """(true positives, total) of a JSON ground-truth pair list."""
    L = _lib.lib()
    p = np.ascontiguousarray(pairs, dtype=np.int32).reshape(-1, 2)
    tp, tot = ctypes.c_int64(0), ctypes.c_int64(0)
    rc = L.stq_recall(p.ctypes.data, len(p), os.fsencode(json_path), int(bool(canonical)), ctypes.byref(tp),
                      ctypes.byref(tot))
    if rc != 0:
        raise StqError(rc, "stq_recall: " + L.stq_status_string(rc).decode())
    return tp.value, tot.value


def canonical_sort(pairs: np.ndarray) -> np.ndarray:
    """Lexicographic order of (first, second) pairs (numpy, for comparisons)."""
    p = np.asarray(pairs, dtype=np.int32).reshape(-1, 2)
    if len(p) == 0:
        return p.copy()
    key = (p[:, 0].astype(np.int64) << 32) | p[:, 1].astype(np.int64)
    return p[np.argsort(key, kind="stable")]
