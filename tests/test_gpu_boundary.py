"""The boundary's building blocks on the GPU, through the C-ABI:
  * stq_find_pairs_rows_device: ranges of sorted rows partition the pair set
    (each pair in the range of its sorted-earlier box), as the reference's
    tidstart ranges do (src/sweep.cu:234-240);
  * stq_sort_order_device: the order the reference sweeps (FP64 min on x,
    ties by input position), against numpy's stable sort, with -0.0/+0.0;
  * stq_radix_sort_device: the key sort itself against numpy's stable
    argsort, for 4- and 8-byte keys, the three forms, digits equal for every
    key, segmented keys (scene bits above a 32-bit key, as the batched form
    sorts them), ragged tails and tiny sizes;
  * the learned single pass through stq_find_pairs_device_alloc and
    stq_find_pairs_host, against the oracle.
Bar: bit-exact (integer ids, permutations)."""
import ctypes

import numpy as np
import pytest

import stq_amd
from oracle import oracle
from stq_amd import _lib, scenes

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = stq_amd.Context(0)
    yield c
    c.close()


def sort_order(ctx, b):
    db = stq_amd.DeviceBuffer.from_array(b)
    d = stq_amd.DeviceBuffer(4 * max(len(b), 1))
    rc = ctx._L.stq_sort_order_device(ctx.ptr, db.ptr, len(b), d.ptr)
    out = d.to_array(np.int32, len(b)) if rc == 0 else None
    db.close()
    d.close()
    return rc, out


def rows_pairs(ctx, db, n, r0, r1):
    cnt = ctypes.c_int64(0)
    rc = ctx._L.stq_find_pairs_rows_device(ctx.ptr, db.ptr, n, r0, r1, None, 0, ctypes.byref(cnt))
    assert rc in (0, _lib.STQ_ERR_CAPACITY)
    cap = max(cnt.value, 1)
    d = stq_amd.DeviceBuffer(8 * cap)
    rc = ctx._L.stq_find_pairs_rows_device(ctx.ptr, db.ptr, n, r0, r1, d.ptr, cap, ctypes.byref(cnt))
    assert rc == 0
    p = d.to_array(np.int32, 2 * cnt.value).reshape(-1, 2)
    d.close()
    return p


def test_sort_order_is_the_reference_order(ctx):
    b = scenes.uniform_scene(70_001, seed=21)
    b["min"][::7, 0] = b["min"][3, 0]          # ties: broken by input position
    b["min"][5::11, 0] = -0.0                  # -0.0 and +0.0 are equal keys
    b["min"][6::11, 0] = 0.0
    b["max"][:, 0] = np.maximum(b["max"][:, 0], b["min"][:, 0])
    rc, order = sort_order(ctx, b)
    assert rc == 0
    key = b["min"][:, 0] + 0.0                 # -0.0 -> +0.0
    ref = np.lexsort((np.arange(len(b)), key))
    assert np.array_equal(order, ref)
    bad = b.copy()
    bad["max"][9, 0] = bad["min"][9, 0] - 1.0
    assert sort_order(ctx, bad)[0] == _lib.STQ_ERR_INVALID_BOX


@pytest.mark.parametrize("cuts", [(0.0, 0.31, 0.77), (0.5,), (0.999,)])
def test_row_ranges_partition_the_pair_set(ctx, cuts):
    b = scenes.uniform_scene(60_000, seed=22, extent_scale=4.0)
    n = len(b)
    ref = oracle.reference_set(b)
    _, order = sort_order(ctx, b)
    rank = np.empty(n, np.int64)
    rank[order] = np.arange(n)
    db = stq_amd.DeviceBuffer.from_array(b)
    edges = [0] + [int(c * n) for c in cuts] + [n + 17]  # the last range passes the end
    parts = []
    for r0, r1 in zip(edges[:-1], edges[1:]):
        p = rows_pairs(ctx, db, n, r0, r1)
        earlier = np.minimum(rank[p[:, 0]], rank[p[:, 1]]) if len(p) else np.zeros(0)
        assert np.all((earlier >= r0) & (earlier < r1)), (r0, r1)
        parts.append(p)
    db.close()
    allp = np.concatenate(parts)
    assert len(allp) == len(ref)
    assert np.array_equal(oracle.canonical(allp), ref)


def radix(ctx, keys, bits, form):
    kb = keys.dtype.itemsize
    # values distinct from the positions, so a sort that drops them shows
    vals = ((np.arange(len(keys), dtype=np.uint64) * np.uint64(2654435761)) % np.uint64(1 << 32)).astype(np.uint32)
    dk, dv = stq_amd.DeviceBuffer.from_array(keys), stq_amd.DeviceBuffer.from_array(vals)
    rc = ctx._L.stq_radix_sort_device(ctx.ptr, dk.ptr, dv.ptr, len(keys), kb, bits, form)
    assert rc == 0, ctx._L.stq_last_error(ctx.ptr)
    k, v = dk.to_array(keys.dtype, len(keys)), dv.to_array(np.uint32, len(keys))
    dk.close()
    dv.close()
    return k, v


def key_cases(rng):
    n_big = 4096 * 37 + 1234                   # ragged last tile
    yield "u32 random", rng.integers(0, 2 ** 32, n_big, dtype=np.uint64).astype(np.uint32), 32
    yield "u32 few values", rng.integers(0, 5, n_big).astype(np.uint32), 32
    yield "u32 constant high digits", (0xAB000000 | rng.integers(0, 4096, n_big)).astype(np.uint32), 32
    yield "u32 24 bits", rng.integers(0, 2 ** 24, n_big).astype(np.uint32), 24
    yield "u32 constant", np.full(n_big, 77, np.uint32), 32
    yield "u64 random", rng.integers(0, 2 ** 63, n_big, dtype=np.uint64) * np.uint64(2), 64
    scene = rng.integers(0, 512, n_big).astype(np.uint64)
    yield "u64 scene|key32", (scene << np.uint64(32)) | rng.integers(0, 2 ** 32, n_big, dtype=np.uint64), 41
    yield "u64 float keys", rng.random(n_big).view(np.uint64) ^ np.uint64(1 << 63), 64
    for m in (0, 1, 2, 3, 4095, 4097):
        yield f"u32 n={m}", rng.integers(0, 2 ** 32, m, dtype=np.uint64).astype(np.uint32), 32


@pytest.mark.parametrize("form", [0, 1, 2])
def test_radix_sort_matches_numpy_stable_argsort(ctx, form):
    rng = np.random.default_rng(23)
    for name, keys, bits in key_cases(rng):
        if form == 2 and keys.dtype != np.uint32:
            continue
        k, v = radix(ctx, keys, bits, form)
        ref = np.argsort(keys, kind="stable")
        vals = ((np.arange(len(keys), dtype=np.uint64) * np.uint64(2654435761)) % np.uint64(1 << 32)).astype(np.uint32)
        assert np.array_equal(v, vals[ref]), (name, form)
        assert np.array_equal(k, keys[ref]), (name, form)


@pytest.mark.parametrize("precision", [stq_amd.PRECISION_F64, stq_amd.PRECISION_F32_CONFIRM])
def test_learned_pass_through_alloc_and_host_entries(precision):
    train = [scenes.density_scene(i % 2, 40_000, 2.0, seed=300 + i) for i in range(3)]
    b = scenes.cloth_scene(g=150, seed=24, sigma=6e-3)
    ref = oracle.reference_set(b)
    with stq_amd.Context(0, precision=precision, sizing=stq_amd.SIZING_LEARNED, train_scenes=3) as c:
        for t in train:
            c.find_pairs(t)  # count then fill while the network trains
        got = c.find_pairs(b)  # stq_find_pairs_device_alloc, learned single pass
        st = c.stats()
        assert st["sizing"] == 1 and st["n_tiles"] > 0 and st["ms_sweep_count"] == 0.0
        assert np.array_equal(oracle.canonical(got), ref) and len(got) == len(ref)
        got_h = c.find_pairs_host(b)  # host in, host out, learned single pass
        assert c.stats()["ms_sweep_count"] == 0.0
        assert np.array_equal(oracle.canonical(got_h), ref) and len(got_h) == len(ref)
        print(f"learned alloc/host: {len(ref)} pairs, fallback items {st['n_fallback_tiles']}", flush=True)
