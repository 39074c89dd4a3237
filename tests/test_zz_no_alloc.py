"""The allocation-free entry point (stq_find_pairs_device_noalloc) through the
C-ABI: on two calls in a row, the second with a smaller input, its pairs
equal those of stq_find_pairs_device on the same device boxes (canonical
sets compared in full, integer ids: bit-exact) and the oracle's set; a scene
above STQ_NOALLOC_MAX_BOXES is rejected before anything runs, and a short
output buffer gives STQ_ERR_CAPACITY with the count."""
import ctypes

import numpy as np
import pytest

import stq_amd
from oracle import oracle
from stq_amd import _lib, scenes

pytestmark = pytest.mark.gpu

MAX_BOXES = 262144  # STQ_NOALLOC_MAX_BOXES


def call(ctx, fn, db, n, dout, cap):
    cnt = ctypes.c_int64(0)
    rc = fn(ctx.ptr, db.ptr, n, dout.ptr if dout is not None else None, cap, ctypes.byref(cnt))
    return rc, cnt.value


def pairs_of(dout, cnt):
    return dout.to_array(np.int32, 2 * cnt).reshape(-1, 2)


@pytest.mark.parametrize("precision", [stq_amd.PRECISION_F64, stq_amd.PRECISION_F32_CONFIRM])
def test_noalloc_matches_device_entry(precision):
    L = _lib.lib()
    big = scenes.density_scene(1, 250_000, 2.0, seed=901)       # cloth-like, near the maximum
    small = scenes.uniform_scene(60_000, seed=902, extent_scale=3.0)
    assert len(big) <= MAX_BOXES
    with stq_amd.Context(0, precision=precision, axis_policy=stq_amd.AXIS_CHAIN) as ctx:
        for b in (big, small):  # the second call is smaller: the array keeps the last contents
            n = len(b)
            db = stq_amd.DeviceBuffer.from_array(b)
            rc, need = call(ctx, L.stq_find_pairs_device, db, n, None, 0)
            assert rc in (_lib.STQ_OK, _lib.STQ_ERR_CAPACITY)
            cap = max(need, 1)
            d1, d2 = stq_amd.DeviceBuffer(8 * cap), stq_amd.DeviceBuffer(8 * cap)
            rc1, c1 = call(ctx, L.stq_find_pairs_device, db, n, d1, cap)
            rc2, c2 = call(ctx, L.stq_find_pairs_device_noalloc, db, n, d2, cap)
            print(f"n={n} precision={precision}: device {rc1}/{c1}, noalloc {rc2}/{c2}", flush=True)
            assert rc1 == _lib.STQ_OK and rc2 == _lib.STQ_OK, ctx._L.stq_last_error(ctx.ptr)
            assert c1 == c2 == need
            p1, p2 = pairs_of(d1, c1), pairs_of(d2, c2)
            assert np.all(p2[:, 0] < p2[:, 1])
            g = oracle.canonical(p2)
            assert len(g) == len(p2), "duplicate pairs"
            assert np.array_equal(g, oracle.canonical(p1))
            assert np.array_equal(g, oracle.reference_set(b))
            # a buffer one pair short: the count comes back, nothing else is promised
            if need > 0:
                rc3, c3 = call(ctx, L.stq_find_pairs_device_noalloc, db, n, d2, need - 1)
                assert rc3 == _lib.STQ_ERR_CAPACITY and c3 == need
            for x in (db, d1, d2):
                x.close()


def test_noalloc_rejects_more_than_the_maximum():
    L = _lib.lib()
    b = scenes.uniform_scene(MAX_BOXES + 1, seed=903)
    with stq_amd.Context(0) as ctx:
        db = stq_amd.DeviceBuffer.from_array(b)
        dout = stq_amd.DeviceBuffer(8)
        rc, _ = call(ctx, L.stq_find_pairs_device_noalloc, db, len(b), dout, 1)
        assert rc == _lib.STQ_ERR_TOO_LARGE
        assert b"STQ_NOALLOC_MAX_BOXES" in L.stq_last_error(ctx.ptr)
        db.close()
        dout.close()
