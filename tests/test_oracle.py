"""Pins the CPU oracle (oracle/stq_oracle.c) before anything is checked against it.

* known-answer fixtures derived by hand (tests/golden/known_answers.json),
* O(N^2) brute force on random scenes of every kind,
* invariance of the pair set under the sweep axis,
* the FP32 rounding rule against hand-derived bit patterns and libm,
* the FP32 superset property of the reference's default (float) build.
A planted mistake in the oracle (dropped axis, flipped inequality, wrong
type rule, wrong padding) fails at least one of these.
"""
import math
import struct

import numpy as np
import pytest

from oracle import oracle
from stq_amd import scenes


def test_known_answers(known_answers):
    for name, (boxes, expected) in known_answers.items():
        got = oracle.canonical(oracle.sweep(boxes)) if len(boxes) else np.zeros((0, 2), np.int32)
        assert np.array_equal(got, oracle.canonical(expected)), name
        if len(boxes) > 1:
            bf = oracle.canonical(oracle.bruteforce(boxes))
            assert np.array_equal(bf, got), name


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_known_answers_every_axis(known_answers, axis):
    for name, (boxes, expected) in known_answers.items():
        if len(boxes) < 2:
            continue
        got = oracle.canonical(oracle.sweep(boxes, axis=axis))
        assert np.array_equal(got, oracle.canonical(expected)), (name, axis)


def _scenes_small():
    yield "uniform", scenes.uniform_scene(4000, seed=11, extent_scale=6.0)
    yield "uniform_small_pool", scenes.uniform_scene(3000, seed=12, extent_scale=8.0, vid_pool=200)
    yield "cloth", scenes.cloth_scene(g=30, seed=13, sigma=2e-2)
    yield "clustered", scenes.clustered_scene(n=4000, seed=14, clusters=8, sigma=0.05, hmin=2e-3)
    yield "dense_cloth", scenes.cloth_scene(g=24, seed=15, sigma=5e-2)


@pytest.mark.parametrize("name,boxes", list(_scenes_small()))
def test_sweep_equals_bruteforce(name, boxes):
    bf = oracle.canonical(oracle.bruteforce(boxes))
    assert len(bf) > 20, f"{name}: scene too sparse to pin anything"
    for axis in (0, 1, 2):
        sw = oracle.canonical(oracle.sweep(boxes, axis=axis))
        assert np.array_equal(sw, bf), (name, axis)


def test_sweep_orientation_and_uniqueness():
    b = scenes.uniform_scene(3000, seed=21, extent_scale=8.0)
    raw = oracle.sweep(b)
    # every raw pair is (sorted-earlier, sorted-later): its min is <= the other's
    assert np.all(b["min"][raw[:, 0], 0] <= b["min"][raw[:, 1], 0])
    can = oracle.canonical(raw)
    assert len(can) == len(raw)  # no duplicates emitted


def test_planted_mistakes_are_caught():
    """Each plausible restatement error changes the set on a random scene."""
    b = scenes.uniform_scene(3000, seed=31, extent_scale=8.0, vid_pool=300)
    truth = oracle.canonical(oracle.bruteforce(b))
    n = len(b)
    mn, mx, v = b["min"], b["max"], b["vids"].astype(np.int64)

    def numpy_set(drop_axis=None, open_intervals=False, ignore_type=False, ignore_covertex=False):
        i, j = np.triu_indices(n, 1)
        ok = np.ones(len(i), bool)
        for k in range(3):
            if k == drop_axis:
                continue
            if open_intervals:
                ok &= (mn[i, k] < mx[j, k]) & (mn[j, k] < mx[i, k])
            else:
                ok &= (mn[i, k] <= mx[j, k]) & (mn[j, k] <= mx[i, k])
        face = lambda x: x[:, 2] >= 0
        edge = lambda x: (x[:, 2] < 0) & (x[:, 1] >= 0)
        vert = lambda x: (x[:, 2] < 0) & (x[:, 1] < 0)
        if not ignore_type:
            a, c = v[i], v[j]
            ok &= (vert(a) & face(c)) | (face(a) & vert(c)) | (edge(a) & edge(c))
        if not ignore_covertex:
            a, c = v[i], v[j]
            share = np.zeros(len(i), bool)
            for s in range(3):
                for t in range(3):
                    share |= a[:, s] == c[:, t]
            ok &= ~share
        return np.stack([i[ok], j[ok]], 1).astype(np.int32)

    assert np.array_equal(numpy_set(), truth)
    for kw in ({"drop_axis": 0}, {"drop_axis": 2}, {"ignore_type": True}, {"ignore_covertex": True}):
        assert not np.array_equal(numpy_set(**kw), truth), kw
    # open intervals differ once boxes touch exactly
    t = b.copy()
    t["min"][1] = t["max"][0]
    t["max"][1] = t["max"][0] + 0.01
    t["vids"][0] = [0, -1, -1]
    t["vids"][1] = [1, 2, 3]
    assert [0, 1] in oracle.canonical(oracle.sweep(t)).tolist()


def _bits(x):
    return struct.unpack("<I", struct.pack("<f", float(x)))[0]


def test_rounding_golden(rounding_vectors):
    vals = np.array([r["value"] for r in rounding_vectors])
    lo, hi = oracle.round_outward(vals)
    for r, a, b in zip(rounding_vectors, lo, hi):
        assert _bits(a) == r["lo_bits"], r
        assert _bits(b) == r["hi_bits"], r


def test_rounding_contains_and_is_tight():
    rng = np.random.default_rng(7)
    vals = np.concatenate([
        rng.uniform(-1, 1, 20000),
        rng.uniform(0, 1, 20000) * 10.0 ** rng.integers(-40, 38, 20000),
        -rng.uniform(0, 1, 20000) * 10.0 ** rng.integers(-40, 38, 20000),
        np.array([0.0, -0.0, 1.0, -1.0, 2.0 ** -149, -(2.0 ** -149), 2.0 ** -126]),
    ])
    lo, hi = oracle.round_outward(vals)
    assert np.all(lo.astype(np.float64) < vals)
    assert np.all(hi.astype(np.float64) > vals)
    # at most two float ulps apart around the nearest float
    f = vals.astype(np.float32)
    up = np.nextafter(f, np.float32(np.inf))
    dn = np.nextafter(f, np.float32(-np.inf))
    assert np.array_equal(lo, dn) and np.array_equal(hi, up)


def test_fp32_build_is_superset():
    """The reference default (float, outward) set contains the FP64 set."""
    for seed in (41, 42):
        b = scenes.uniform_scene(4000, seed=seed, extent_scale=6.0)
        f64 = oracle.reference_set(b)
        f32 = oracle.reference_set(oracle.boxes_round_outward(b))
        s64 = set(map(tuple, f64.tolist()))
        s32 = set(map(tuple, f32.tolist()))
        assert s64 <= s32


def test_candidates_count_matches_definition():
    b = scenes.uniform_scene(2000, seed=51, extent_scale=5.0)
    tot, mc = oracle.candidates(b, axis=0)
    order = np.lexsort((np.arange(len(b)), b["min"][:, 0]))
    smin = b["min"][order, 0]
    smax = b["max"][order, 0]
    hi = np.searchsorted(smin, smax, side="right") - 1
    lens = hi - np.arange(len(b))
    assert tot == int(lens.sum()) and mc == int(lens.max())


# ---- swept-box construction (src/io.cu:15-28, src/aabb.cu:56-174) -----------------
SPECIAL = np.array([0.0, -0.0, 1e-45, -1e-45, 1e-40, 0.1, -0.1, 1.0, 3.0e38, -3.0e38, 3.5e38, -3.5e38,
                    1e300, -1e300, 16777217.0, 0.333333333333], dtype=np.float64)


def _mesh(seed, nv=300):
    rng = np.random.default_rng(seed)
    V0 = rng.uniform(-2, 2, (nv, 3))
    V0.ravel()[:len(SPECIAL)] = SPECIAL
    V1 = V0 + rng.normal(0, 1e-2, V0.shape)
    V1.ravel()[len(SPECIAL):2 * len(SPECIAL)] = SPECIAL[::-1]
    E = rng.integers(0, nv, (500, 2))
    F = rng.integers(0, nv, (400, 3))
    return V0, V1, E, F


def test_build_fp64_matches_hull():
    """FP64 build: exactly min/max over the simplex's t0/t1 points -/+ r, the
    id/vertex-id/ref_id conventions, vertices then edges then faces."""
    V0, V1, E, F = _mesh(1)
    for r in (0.0, 1e-3):
        b = oracle.build_swept_boxes(V0, V1, E, F, radius=r)
        nv, ne = len(V0), len(E)
        assert len(b) == nv + ne + len(F)
        assert np.array_equal(b["id"], np.arange(len(b)))
        lo = np.minimum(V0, V1)
        hi = np.maximum(V0, V1)
        for s0, idx in ((0, np.arange(nv)[:, None]), (nv, E), (nv + ne, F)):
            m = len(idx)
            assert np.array_equal(b["min"][s0:s0 + m], lo[idx].min(axis=1) - r)
            assert np.array_equal(b["max"][s0:s0 + m], hi[idx].max(axis=1) + r)
            assert np.array_equal(b["ref_id"][s0:s0 + m], np.arange(m))
        i = np.arange(nv)
        assert np.array_equal(b["vids"][:nv], np.stack([i, -i - 1, -i - 1], 1))
        assert np.array_equal(b["vids"][nv:nv + ne], np.stack([E[:, 0], E[:, 1], -E[:, 0] - 1], 1))
        assert np.array_equal(b["vids"][nv + ne:], F)
    # r = 0 equals the scene generator's own construction
    Vc, Ec, Fc = scenes.cloth_mesh(30)
    Vc1 = Vc + np.random.default_rng(3).normal(0, 2e-3, Vc.shape)
    a = oracle.build_swept_boxes(Vc, Vc1, Ec, Fc)
    g = scenes.swept_boxes(Vc, Vc1, Ec, Fc)
    for k in ("id", "min", "max", "vids"):
        assert np.array_equal(a[k], g[k])


def _np_next(x, up):
    big = np.float32(3.40282347e38)
    with np.errstate(over="ignore", invalid="ignore"):
        return np.nextafter(x, (x + big) if up else (x - big)).astype(np.float32)


def test_build_float_matches_numpy_nextafter():
    """Float build against numpy's float32 nextafter and float32 arithmetic
    (src/aabb.cu:56-61,86-91,120-125,162-167), special values included."""
    V0, V1, E, F = _mesh(2)
    nv, ne = len(V0), len(E)
    lo = np.minimum(V0, V1)
    hi = np.maximum(V0, V1)
    for r in (0.0, 1e-3, 1e-45):
        b = oracle.build_swept_boxes(V0, V1, E, F, radius=r, float_build=True)
        with np.errstate(over="ignore", invalid="ignore"):
            rf = _np_next(np.float32(r), True)
            for s0, idx in ((0, np.arange(nv)[:, None]), (nv, E), (nv + ne, F)):
                m = len(idx)
                wl = _np_next(lo[idx].min(axis=1).astype(np.float32), False) - rf
                wh = _np_next(hi[idx].max(axis=1).astype(np.float32), True) + rf
                got_l = b["min"][s0:s0 + m]
                got_h = b["max"][s0:s0 + m]
                assert np.array_equal(got_l, wl.astype(np.float64)), r
                assert np.array_equal(got_h, wh.astype(np.float64)), r
    # outward (for finite float images and r = 0): the float box contains the FP64 box
    b64 = oracle.build_swept_boxes(V0, V1, E, F)
    b32 = oracle.build_swept_boxes(V0, V1, E, F, float_build=True)
    fin = np.isfinite(b32["min"]) & np.isfinite(b32["max"]) & (np.abs(b64["min"]) < 3.4e38) & \
        (np.abs(b64["max"]) < 3.4e38)
    assert np.all(b32["min"][fin] <= b64["min"][fin]) and np.all(b32["max"][fin] >= b64["max"][fin])


def test_build_rejects_bad_indices():
    V0, V1, E, F = _mesh(3, nv=20)
    E = E % 20
    F = F % 20
    E[3, 1] = 20
    with pytest.raises(ValueError):
        oracle.build_swept_boxes(V0, V1, E, F)
    E[3, 1] = 0
    F[2, 0] = -1
    with pytest.raises(ValueError):
        oracle.build_swept_boxes(V0, V1, E, F)


def test_orient_pairs_known_answers():
    """(vertex, face) whatever the input order; edge-edge (min, max); other kinds rejected."""
    b = scenes.empty_boxes(6)
    b["vids"][0] = (0, -1, -1)   # vertex
    b["vids"][1] = (1, 2, 3)     # face
    b["vids"][2] = (4, 5, -5)    # edge
    b["vids"][3] = (6, 7, -7)    # edge
    b["vids"][4] = (8, -9, -9)   # vertex
    b["vids"][5] = (9, 10, 11)   # face
    got = oracle.orient_pairs(b, [[0, 1], [1, 0], [3, 2], [2, 3], [5, 4], [4, 5]])
    assert got.tolist() == [[0, 1], [0, 1], [2, 3], [2, 3], [4, 5], [4, 5]]
    for bad in ([[0, 4]], [[1, 5]], [[0, 2]], [[2, 5]], [[0, 6]], [[-1, 0]]):
        with pytest.raises(ValueError):
            oracle.orient_pairs(b, bad)


def test_radix_canonical_form_matches_numpy():
    """oracle.canonical (the radix sort under every set comparison) against
    numpy's sort and unique: negative ids, reversed pairs, repeats, sizes on
    both sides of the threaded path."""
    rng = np.random.default_rng(31)
    for n in (0, 1, 7, 65535, 65536, 300_001):
        k = rng.integers(0, 2 ** 64 - 1, n, dtype=np.uint64)
        assert np.array_equal(oracle.sort_u64(k), np.sort(k)), n
        p = rng.integers(-2 ** 31, 2 ** 31 - 1, (n, 2)).astype(np.int32)
        if n > 10:
            p[: n // 3] = p[n // 3: 2 * (n // 3)][:, ::-1]  # the same pairs reversed
        a = np.minimum(p[:, 0], p[:, 1]).astype(np.int64)
        b = np.maximum(p[:, 0], p[:, 1]).astype(np.int64)
        ref = np.unique(np.stack([a, b], 1), axis=0).astype(np.int32).reshape(-1, 2)
        assert np.array_equal(oracle.canonical(p), ref), n
