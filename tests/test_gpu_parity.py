"""GPU parity: the HIP broad phase (through the C-ABI) against the CPU oracle.

Bar: the canonical pair set is bit-identical (integer ids; the FP64 path does
only comparisons).  Sizes where the oracle finishes in seconds are compared
pair for pair; the 3M-box cloth is compared on a spatial crop (the pairs of a
crop are exactly the full run's pairs with both boxes inside it) plus
size-independent properties over the whole output.
"""
import numpy as np
import pytest

import stq_amd
from oracle import oracle
from stq_amd import scenes

pytestmark = pytest.mark.gpu

FORMS = [
    ("f64", dict(precision=stq_amd.PRECISION_F64)),
    ("f32c", dict(precision=stq_amd.PRECISION_F32_CONFIRM)),
    ("f64g", dict(precision=stq_amd.PRECISION_F64_GROUPED)),
]


@pytest.fixture(scope="module")
def ctx():
    c = stq_amd.Context(0)
    yield c
    c.close()


def run(ctx, boxes, **kw):
    ctx.set_options(**kw)
    return ctx.find_pairs(boxes)


def same_set(got, ref):
    g = oracle.canonical(got)
    assert len(g) == len(got), "duplicate pairs in the output"
    assert np.all(got[:, 0] < got[:, 1]) if len(got) else True
    return np.array_equal(g, oracle.canonical(ref))


@pytest.mark.parametrize("form,kw", FORMS)
@pytest.mark.parametrize("axis", [0, 1, 2, 3, 4])
def test_known_answers(ctx, known_answers, form, kw, axis):
    for name, (boxes, expected) in known_answers.items():
        got = run(ctx, boxes, axis_policy=axis, **kw)
        assert same_set(got, expected), (name, form, axis)


@pytest.mark.parametrize("form,kw", FORMS)
def test_uniform_100k(ctx, form, kw):
    b = scenes.uniform_scene(seed=0)
    ref = oracle.reference_set(b)
    got = run(ctx, b, **kw)
    print(f"uniform 100k {form}: {len(got)} pairs, stats {ctx.stats()}", flush=True)
    assert same_set(got, ref)


@pytest.mark.parametrize("form,kw", FORMS)
@pytest.mark.parametrize("name,maker", [
    ("uniform_dense", lambda: scenes.uniform_scene(60000, seed=5, extent_scale=4.0, vid_pool=5000)),
    ("cloth_g120", lambda: scenes.cloth_scene(g=120, seed=6, sigma=1e-2)),
    ("clustered_300k", lambda: scenes.clustered_scene(n=300_000, seed=7, clusters=16)),
])
def test_scenes_vs_oracle(ctx, form, kw, name, maker):
    b = maker()
    ref = oracle.reference_set(b)
    assert len(ref) > 100
    for axis in (0, 4):
        got = run(ctx, b, axis_policy=axis, **kw)
        assert same_set(got, ref), (name, form, axis)


# Windows 64..256 sort each group's rows in LDS; 512 and up use the
# register-blocked window sort (1, 2, 4 and 8 keys per thread at 512, 1024,
# 2048 and 4096).
@pytest.mark.parametrize("window,chunk,batch,xcd", [(64, 1, 2, 1), (64, 7, 4, 2), (256, 64, 8, 1),
                                                   (512, 16, 2, 2), (1024, 300, 4, 1),
                                                   (2048, 100000, 0, 2), (128, 33, 8, 2)])
def test_planning_parameters_do_not_change_the_set(ctx, window, chunk, batch, xcd):
    b = scenes.uniform_scene(20000, seed=8, extent_scale=4.0)
    ref = oracle.reference_set(b)
    for kw in (dict(), dict(precision=stq_amd.PRECISION_F32_CONFIRM),
               dict(precision=stq_amd.PRECISION_F64_GROUPED)):
        got = run(ctx, b, window=window, chunk=chunk, batch=batch, xcd_split=xcd, **kw)
        assert same_set(got, ref), (window, chunk, batch, xcd, kw)


@pytest.mark.parametrize("window", [512, 1024])
def test_planning_windows_of_the_batched_planner(window):
    """The batched planner's register-blocked window sort at 512 and 1024."""
    sc = [scenes.uniform_scene(n, seed=80 + i, extent_scale=4.0) for i, n in enumerate((3000, 17000, 40))]
    sc.append(scenes.cloth_scene(g=60, seed=84, sigma=1e-2))
    allb, offs = scenes.concat_scenes(sc)
    c = stq_amd.Context(0)
    try:
        for kw in (dict(), dict(precision=stq_amd.PRECISION_F32_CONFIRM),
                   dict(precision=stq_amd.PRECISION_F64_GROUPED)):
            c.set_options(window=window, chunk=64, axis_policy=4, **kw)
            pairs, poffs = c.find_pairs_batched(allb, offs)
            for s, b in enumerate(sc):
                assert same_set(pairs[poffs[s]:poffs[s + 1]], oracle.reference_set(b)), (window, kw, s)
    finally:
        c.close()


def test_deterministic_output_order(ctx):
    b = scenes.uniform_scene(50000, seed=9, extent_scale=3.0)
    ctx.set_options()
    a1 = ctx.find_pairs(b)
    a2 = ctx.find_pairs(b)
    assert np.array_equal(a1, a2)


def test_rounding_edges_fp32_confirm_is_exact(ctx):
    """Coordinates on and next to float boundaries, tiny and huge values: the
    FP32 outward sweep plus FP64 confirmation equals the FP64 set."""
    rng = np.random.default_rng(10)
    n = 20000
    grid = np.float32(1.0) + np.arange(-50, 50, dtype=np.float32) * np.finfo(np.float32).eps
    c = rng.choice(grid.astype(np.float64), size=(n, 3))
    c += rng.integers(-2, 3, size=(n, 3)) * 2.0 ** -60      # sub-float-ulp offsets
    h = rng.choice([0.0, 2.0 ** -30, 2.0 ** -24, 1e-7], size=(n, 3))
    b = scenes.empty_boxes(n)
    b["min"] = c - h
    b["max"] = c + h
    b["vids"] = scenes._vids_random(rng, n, 1000)
    # a few huge and tiny boxes
    b["min"][:50] = -1e300
    b["max"][:50] = -1e299
    b["min"][50:100] = 1e-300
    b["max"][50:100] = 2e-300
    ref = oracle.reference_set(b)
    f64 = run(ctx, b, precision=stq_amd.PRECISION_F64)
    f32 = run(ctx, b, precision=stq_amd.PRECISION_F32_CONFIRM)
    assert same_set(f64, ref)
    assert same_set(f32, ref)
    assert ctx.stats()["n_confirm"] >= len(ref)


def test_sweep_keys_outliers_infinite_bounds_and_key_widths(ctx):
    """The reduced-precision form sorts quantised sweep keys (16 bits up to
    2^22 boxes per scene, 24 above).  Far outliers, infinite bounds, a scene
    whose sweep coordinates are all equal and scenes on both sides of 2^20
    boxes, single and batched, give the oracle's set exactly; the outliers do
    not coarsen the other boxes' keys (the candidate count barely moves)."""
    f32 = dict(precision=stq_amd.PRECISION_F32_CONFIRM, axis_policy=stq_amd.AXIS_X)
    rng = np.random.default_rng(21)
    b = scenes.uniform_scene(200000, seed=21)
    run(ctx, b, **f32)
    base = ctx.stats()["n_candidates"]
    assert base >= oracle.candidates(b, axis=0)[0]
    o = b.copy()
    k = rng.choice(len(o), size=40, replace=False)
    far = np.where(rng.random((40, 1)) < 0.5, -1e6, 1e6) * (1.0 + rng.random((40, 3)))
    o["min"][k] = far
    o["max"][k] = far + 1.0
    o["min"][k[:5], 0] = -np.inf          # infinite bounds on the sweep axis
    o["max"][k[5:10], 0] = np.inf
    o["min"][k[10:12]] = -np.inf
    o["max"][k[10:12]] = np.inf           # boxes that meet everything
    got = run(ctx, o, **f32)
    assert same_set(got, oracle.reference_set(o))
    cand = ctx.stats()["n_candidates"]
    print(f"candidates: {base} without outliers, {cand} with them", flush=True)
    assert cand <= 1.02 * base + 2 * 12 * len(o)
    # every sweep coordinate equal: one key for all
    e = scenes.uniform_scene(3000, seed=22)
    e["min"][:, 0] = 0.25
    e["max"][:, 0] = 0.25
    assert same_set(run(ctx, e, **f32), oracle.reference_set(e))
    # past 2^22 boxes: 24-bit keys, single and as the largest scene of a batch
    big = scenes.uniform_scene((1 << 22) + 4096, seed=23, extent_scale=0.05)
    ref_big = oracle.reference_set(big)
    assert same_set(run(ctx, big, **f32), ref_big)
    small = scenes.uniform_scene(30000, seed=24)
    allb, offs = scenes.concat_scenes([small, big])
    ctx.set_options(**f32)
    pairs, poffs = ctx.find_pairs_batched(allb, offs)
    assert same_set(pairs[poffs[0]:poffs[1]], oracle.reference_set(small))
    assert same_set(pairs[poffs[1]:poffs[2]], ref_big)


def test_invalid_boxes_rejected(ctx):
    b = scenes.uniform_scene(1000, seed=11)
    ctx.set_options()
    bad = b.copy()
    bad["min"][5, 1] = np.nan
    with pytest.raises(stq_amd.StqError) as e:
        ctx.find_pairs(bad)
    assert e.value.status == 2
    inv = b.copy()
    inv["min"][7, 2] = inv["max"][7, 2] + 1.0
    with pytest.raises(stq_amd.StqError) as e:
        ctx.find_pairs(inv)
    assert e.value.status == 2
    # the context stays usable
    assert same_set(ctx.find_pairs(b), oracle.reference_set(b))


def test_capacity_and_device_entry(ctx):
    b = scenes.uniform_scene(30000, seed=12, extent_scale=3.0)
    ref = oracle.reference_set(b)
    ctx.set_options()
    db = stq_amd.DeviceBuffer.from_array(b)
    small = stq_amd.DeviceBuffer(8 * 4)
    with pytest.raises(stq_amd.StqError) as e:
        ctx.find_pairs_device(db.ptr, len(b), small.ptr, 4)
    assert e.value.status == 3
    out = stq_amd.DeviceBuffer(8 * len(ref))
    cnt = ctx.find_pairs_device(db.ptr, len(b), out.ptr, len(ref))
    assert cnt == len(ref)
    got = out.to_array(np.int32, 2 * cnt).reshape(-1, 2)
    assert same_set(got, ref)


def test_host_entry_matches(ctx):
    b = scenes.cloth_scene(g=60, seed=13, sigma=2e-2)
    ref = oracle.reference_set(b)
    for form, kw in FORMS:
        ctx.set_options(**kw)
        assert same_set(ctx.find_pairs_host(b), ref), form


def test_batched_equals_independent(ctx):
    rng = np.random.default_rng(14)
    sc = []
    for i in range(12):
        n = int(rng.integers(0, 5000)) if i % 5 else int(rng.integers(0, 3))
        sc.append(scenes.uniform_scene(n, seed=100 + i, extent_scale=5.0) if i % 2 else
                  scenes.cloth_scene(g=int(rng.integers(2, 25)), seed=200 + i, sigma=2e-2))
    allb, offs = scenes.concat_scenes(sc)
    for form, kw in FORMS:
        ctx.set_options(**kw)
        pairs, poffs = ctx.find_pairs_batched(allb, offs)
        for s, b in enumerate(sc):
            got = pairs[poffs[s]:poffs[s + 1]]
            ref = oracle.reference_set(b) if len(b) > 1 else np.zeros((0, 2), np.int32)
            assert same_set(got, ref), (form, s)


def test_cloth_full_size_properties(ctx):
    """The 708 x 708 cloth (3,001,923 boxes)."""
    b = scenes.cloth_scene()
    assert len(b) == 3_001_923
    ctx.set_options(precision=stq_amd.PRECISION_F64)
    f64 = ctx.find_pairs(b)
    st = ctx.stats()
    print(f"cloth f64: {len(f64)} pairs, stats {st}", flush=True)
    ctx.set_options(precision=stq_amd.PRECISION_F32_CONFIRM)
    f32 = ctx.find_pairs(b)
    print(f"cloth f32c: {len(f32)} pairs, stats {ctx.stats()}", flush=True)
    c64 = oracle.canonical(f64)
    assert len(c64) == len(f64)
    assert np.array_equal(c64, oracle.canonical(f32))
    # every pair satisfies the rule (vectorised check over all pairs)
    i, j = c64[:, 0], c64[:, 1]
    ok = np.ones(len(i), bool)
    for k in range(3):
        ok &= (b["min"][i, k] <= b["max"][j, k]) & (b["min"][j, k] <= b["max"][i, k])
    vi, vj = b["vids"][i].astype(np.int64), b["vids"][j].astype(np.int64)
    for s in range(3):
        for t in range(3):
            ok &= vi[:, s] != vj[:, t]
    assert ok.all()
    # crop: boxes with min.x < 0.05 form a sub-scene the oracle sweeps in seconds
    crop = np.nonzero(b["min"][:, 0] < 0.05)[0]
    sub = b[crop].copy()
    ref = oracle.reference_set(sub)
    inside = np.zeros(len(b), bool)
    inside[crop] = True
    pos = np.full(len(b), -1, np.int64)
    pos[crop] = np.arange(len(crop))
    keep = inside[i] & inside[j]
    mine = np.stack([pos[i[keep]], pos[j[keep]]], 1)
    assert len(ref) > 100000
    assert np.array_equal(oracle.canonical(mine), ref)
    # and the whole set against the oracle's full sweep
    assert np.array_equal(c64, oracle.reference_set(b))


@pytest.mark.parametrize("form,kw", FORMS)
def test_clustered_8m_full_size(ctx, form, kw):
    """The 8M-box clustered scene with Pareto extents (64 clusters, sigma
    0.02): the whole pair set against the oracle, with tail items on (the
    default) and with the learned single pass."""
    b = scenes.clustered_scene()
    assert len(b) == 8_000_000
    ref = oracle.reference_set(b)
    got = run(ctx, b, axis_policy=0, **kw)
    st = ctx.stats()
    print(f"clustered 8M {form}: {len(got)} pairs, {st['ms_total']:.1f} ms, items {st['n_items']}, "
          f"lane steps {st['n_lane_steps']:.3e}, candidates {st['n_candidates']:.3e}", flush=True)
    assert same_set(got, ref)
    ctx.sizer_reset()
    for rep in range(2):
        got = run(ctx, b, axis_policy=0, sizing=stq_amd.SIZING_LEARNED, train_scenes=1, **kw)
        assert same_set(got, ref), rep


def _long_row_scene(n=40_000, seed=21, frac=0.004, scale=300.0, extent=2.0):
    """Small uniform boxes plus a few whose sweep-axis extent is `scale` times
    longer: their rows have chains far longer than their group's other rows."""
    b = scenes.uniform_scene(n, seed=seed, extent_scale=extent)
    rng = np.random.default_rng(seed)
    k = rng.choice(n, size=max(1, int(frac * n)), replace=False)
    c = 0.5 * (b["min"][k] + b["max"][k])
    h = 0.5 * (b["max"][k] - b["min"][k])
    h[:, 0] *= scale
    b["min"][k] = c - h
    b["max"][k] = c + h
    return b


@pytest.mark.parametrize("form,kw", FORMS)
@pytest.mark.parametrize("sizing", [stq_amd.SIZING_COUNT_FILL, stq_amd.SIZING_LEARNED])
def test_long_rows_tail_items(ctx, form, kw, sizing):
    b = _long_row_scene()
    ref = oracle.reference_set(b)
    steps = {}
    outs = {}
    for tails in (2, 1):
        ctx.sizer_reset()
        for rep in range(3):  # learned: one training call, then the single pass
            got = run(ctx, b, axis_policy=0, tails=tails, sizing=sizing, train_scenes=1, **kw)
            assert same_set(got, ref), (form, sizing, tails, rep)
        steps[tails] = ctx.stats()["n_lane_steps"]
        outs[tails] = got
        assert np.array_equal(run(ctx, b, axis_policy=0, tails=tails, sizing=sizing, train_scenes=1,
                                  **kw), got)
    print(f"{form} sizing={sizing}: lane steps without tails {steps[2]}, with {steps[1]}", flush=True)
    assert steps[1] < steps[2]


@pytest.mark.parametrize("form,kw", FORMS)
def test_long_rows_tail_items_forced_fallback(ctx, form, kw):
    """Every item (tail items too) overflows its predicted slots and is redone."""
    b = _long_row_scene(seed=22, extent=12.0, scale=40.0)
    ref = oracle.reference_set(b)
    w = np.zeros(6337, np.float32)
    w[-1] = -5.0
    ctx.sizer_set_params(w)
    got = run(ctx, b, axis_policy=0, sizing=stq_amd.SIZING_LEARNED, train_scenes=-1, **kw)
    assert ctx.stats()["n_fallback_tiles"] > 0
    assert same_set(got, ref)
    assert np.array_equal(got, run(ctx, b, axis_policy=0, **kw))
