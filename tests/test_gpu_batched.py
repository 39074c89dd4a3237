"""Batched scenes on the GPU: one sync-free launch sequence over many scenes
(and its hipGraph replays) against the CPU oracle, scene by scene.

Bar: the canonical pair set of every scene is bit-identical to the oracle's
(integer ids local to the scene); replays of a captured graph give the same
pairs in the same order as the eager run of the same batch.
"""
import numpy as np
import pytest

import stq_amd
from stq_amd import _lib, scenes
from oracle import mlp, oracle

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


def ref_set(b):
    return oracle.reference_set(b) if len(b) > 1 else np.zeros((0, 2), np.int32)


def check_scenes(pairs, poffs, sc, tag):
    assert poffs[0] == 0 and np.all(np.diff(poffs) >= 0)
    assert len(pairs) == poffs[-1]
    for s, b in enumerate(sc):
        got = pairs[poffs[s]:poffs[s + 1]]
        c = oracle.canonical(got)
        assert len(c) == len(got), (tag, s, "duplicates")
        if len(got):
            assert np.all(got[:, 0] < got[:, 1]), (tag, s)
        assert np.array_equal(c, ref_set(b)), (tag, s)


def mixed_scenes(seed, count, nmax=6000):
    rng = np.random.default_rng(seed)
    sc = []
    for i in range(count):
        r = i % 7
        if r == 3:
            n = int(rng.integers(0, 3))  # empty and one-box scenes
        else:
            n = int(rng.integers(2, nmax))
        if i % 2:
            sc.append(scenes.uniform_scene(n, seed=seed * 100 + i, extent_scale=float(rng.uniform(2, 8))))
        else:
            g = max(2, int(np.sqrt(max(n, 12) / 6)))
            sc.append(scenes.cloth_scene(g=g, seed=seed * 100 + i, sigma=2e-2))
    return sc


@pytest.mark.parametrize("form,kw", FORMS)
@pytest.mark.parametrize("axis", [0, 2, 3, 4])
@pytest.mark.parametrize("graph", [0, 1, 2])
def test_batched_matches_oracle(ctx, form, kw, axis, graph):
    sc = mixed_scenes(5 + axis, 23)
    allb, offs = scenes.concat_scenes(sc)
    ctx.set_options(axis_policy=axis, graph=graph, **kw)
    for rep in range(3):  # eager, then (graph modes) capture and replay
        pairs, poffs = ctx.find_pairs_batched(allb, offs)
        check_scenes(pairs, poffs, sc, (form, axis, graph, rep))


@pytest.mark.parametrize("form,kw", FORMS)
def test_batched_equals_single_scene_calls(ctx, form, kw):
    sc = mixed_scenes(41, 9, nmax=20000)
    allb, offs = scenes.concat_scenes(sc)
    ctx.set_options(**kw)
    pairs, poffs = ctx.find_pairs_batched(allb, offs)
    for s, b in enumerate(sc):
        single = ctx.find_pairs(b) if len(b) > 1 else np.zeros((0, 2), np.int32)
        assert np.array_equal(oracle.canonical(pairs[poffs[s]:poffs[s + 1]]), oracle.canonical(single))


@pytest.mark.parametrize("form,kw", FORMS)
def test_graph_replays_follow_new_box_values(ctx, form, kw):
    """Same shape and buffers on every call: the first call runs eagerly, the
    second captures, the rest replay.  Between calls the box values change and
    unrelated eager work runs on the same context; every replay is checked
    against the oracle and against an eager run of the same input."""
    ctx.set_options(graph=0, **kw)
    sizes = [3000, 0, 5000, 1, 4000, 2500]
    variants = []
    for v in range(5):
        sc = []
        for i, n in enumerate(sizes):
            if i % 2:
                sc.append(scenes.uniform_scene(n, seed=1000 * v + i, extent_scale=4.0))
            else:
                sc.append(scenes.density_scene(1, n, 1.0 + v, seed=1000 * v + i))
        # keep the shape: density_scene may round the cloth size
        sc = [s[:n] if len(s) >= n else np.concatenate([s, scenes.uniform_scene(n - len(s), seed=7)])
              for s, n in zip(sc, sizes)]
        variants.append(sc)
    allb0, offs = scenes.concat_scenes(variants[0])
    cap = 1 << 20
    db = _lib.DeviceBuffer(allb0.nbytes)
    dout = _lib.DeviceBuffer(8 * cap)
    eager = stq_amd.Context(0, **kw, graph=2)
    modes = []
    try:
        for v, sc in enumerate(variants):
            allb, o2 = scenes.concat_scenes(sc)
            assert np.array_equal(o2, offs)
            db.write(allb)
            rc, poffs = ctx.find_pairs_batched_device(db.ptr, offs, dout.ptr, cap)
            assert rc == _lib.STQ_OK
            modes.append(ctx.stats()["graph"])
            pairs = dout.to_array(np.int32, 2 * int(poffs[-1])).reshape(-1, 2)
            check_scenes(pairs, poffs, sc, (form, v))
            # eager work in between: a single-scene call on the same context
            assert len(ctx.find_pairs(sc[0])) == poffs[1] - poffs[0]
            e_pairs, e_poffs = eager.find_pairs_batched(allb, offs)
            assert np.array_equal(e_poffs, poffs)
            assert np.array_equal(e_pairs, pairs), (form, v)
        print(f"graph states {modes}", flush=True)
        # eager until the buffers stop changing, then one capture, then replays
        assert modes[0] == 0 and modes.count(1) == 1 and modes[-2:] == [2, 2], modes
    finally:
        eager.close()
        db.close()
        dout.close()


@pytest.mark.parametrize("form,kw", FORMS)
def test_batched_capacity(ctx, form, kw):
    sc = mixed_scenes(77, 8)
    allb, offs = scenes.concat_scenes(sc)
    for graph in (2, 1):
        ctx.set_options(graph=graph, **kw)
        db = _lib.DeviceBuffer.from_array(allb)
        small = _lib.DeviceBuffer(8 * 16)
        try:
            rc, poffs = ctx.find_pairs_batched_device(db.ptr, offs, small.ptr, 16)
            need = int(poffs[-1])
            assert need > 16 and rc == _lib.STQ_ERR_CAPACITY
            rc0, p0 = ctx.find_pairs_batched_device(db.ptr, offs, None, 0)
            assert rc0 == _lib.STQ_ERR_CAPACITY and np.array_equal(p0, poffs)
            exact = _lib.DeviceBuffer(8 * need)
            rc, poffs2 = ctx.find_pairs_batched_device(db.ptr, offs, exact.ptr, need)
            assert rc == _lib.STQ_OK and np.array_equal(poffs2, poffs)
            check_scenes(exact.to_array(np.int32, 2 * need).reshape(-1, 2), poffs2, sc, (form, graph))
            exact.close()
        finally:
            db.close()
            small.close()


def test_batched_offsets_not_starting_at_zero(ctx):
    sc = mixed_scenes(90, 5)
    allb, offs = scenes.concat_scenes(sc)
    pad = scenes.uniform_scene(17, seed=3)
    both = np.concatenate([pad, allb])
    ctx.set_options(precision=stq_amd.PRECISION_F32_CONFIRM, graph=2)
    pairs, poffs = ctx.find_pairs_batched(both, offs + 17)
    check_scenes(pairs, poffs, sc, "offset")


@pytest.mark.parametrize("form,kw", FORMS)
def test_batched_invalid_box(ctx, form, kw):
    sc = mixed_scenes(91, 4)
    allb, offs = scenes.concat_scenes(sc)
    bad = allb.copy()
    k = int(offs[2]) + 3
    bad["min"][k, 1] = np.nan
    ctx.set_options(graph=2, **kw)
    with pytest.raises(stq_amd.StqError) as e:
        ctx.find_pairs_batched(bad, offs)
    assert e.value.status == _lib.STQ_ERR_INVALID_BOX
    bad = allb.copy()
    bad["min"][k, 0] = bad["max"][k, 0] + 1.0
    with pytest.raises(stq_amd.StqError):
        ctx.find_pairs_batched(bad, offs)
    # the context stays usable
    pairs, poffs = ctx.find_pairs_batched(allb, offs)
    check_scenes(pairs, poffs, sc, "after error")


def test_batched_stream_512_scenes(ctx):
    """512 scenes (1e3 to 5e4 boxes, both kinds, densities 0.5 to 4), graph
    replay; every scene against the oracle."""
    items = scenes.batched_scenes(count=512, seed=3, nmin=1e3, nmax=5e4)
    sc = [b for (_, _, b) in items]
    allb, offs = scenes.concat_scenes(sc)
    ctx.set_options(precision=stq_amd.PRECISION_F32_CONFIRM, axis_policy=stq_amd.AXIS_CHAIN, graph=1)
    db = _lib.DeviceBuffer.from_array(allb)
    rc, poffs = ctx.find_pairs_batched_device(db.ptr, offs, None, 0)
    need = int(poffs[-1])
    dout = _lib.DeviceBuffer(8 * need)
    try:
        for rep in range(2):
            rc, p2 = ctx.find_pairs_batched_device(db.ptr, offs, dout.ptr, need)
            assert rc == _lib.STQ_OK and np.array_equal(p2, poffs)
        st = ctx.stats()
        print(f"512 scenes, {len(allb)} boxes, {need} pairs: {st['ms_total']:.2f} ms", flush=True)
        check_scenes(dout.to_array(np.int32, 2 * need).reshape(-1, 2), poffs, sc, "stream")
    finally:
        db.close()
        dout.close()


# ---- learned sizing inside the batched pipeline ------------------------------------
def _countfill(kw, allb, offs):
    c = stq_amd.Context(0, graph=2, **kw)
    try:
        return c.find_pairs_batched(allb, offs)
    finally:
        c.close()


@pytest.mark.parametrize("form,kw", FORMS)
@pytest.mark.parametrize("graph", [0, 2])
def test_batched_learned_exact_and_same_order(form, kw, graph):
    """The first call trains on exact counts (count-then-fill); later calls use
    the single learned pass.  Pairs and their order equal the count-then-fill
    pipeline's on every call."""
    sc = mixed_scenes(120, 23)
    allb, offs = scenes.concat_scenes(sc)
    ref_pairs, ref_offs = _countfill(kw, allb, offs)
    check_scenes(ref_pairs, ref_offs, sc, form)
    c = stq_amd.Context(0, sizing=stq_amd.SIZING_LEARNED, train_scenes=10, graph=graph, **kw)
    try:
        for rep in range(4):
            pairs, poffs = c.find_pairs_batched(allb, offs)
            st = c.stats()
            print(f"{form} graph={graph} call {rep}: items {st['n_items']} fallback {st['n_fallback_tiles']} "
                  f"mae {st['sizing_mae']:.2f} state {st['graph']}", flush=True)
            assert np.array_equal(poffs, ref_offs), rep
            assert np.array_equal(pairs, ref_pairs), rep
    finally:
        c.close()


@pytest.mark.parametrize("form,kw", FORMS)
def test_batched_learned_forced_fallback(form, kw):
    """A network that predicts no pairs: every item with more than the slack
    overflows and is redone exactly."""
    sc = mixed_scenes(121, 12)
    allb, offs = scenes.concat_scenes(sc)
    ref_pairs, ref_offs = _countfill(kw, allb, offs)
    c = stq_amd.Context(0, sizing=stq_amd.SIZING_LEARNED, train_scenes=-1, graph=0, **kw)
    try:
        w = np.zeros(mlp.NPARAMS, np.float32)
        w[-1] = -5.0
        c.sizer_set_params(w)
        for rep in range(3):
            pairs, poffs = c.find_pairs_batched(allb, offs)
            st = c.stats()
            assert st["n_fallback_tiles"] > 0
            assert np.array_equal(pairs, ref_pairs) and np.array_equal(poffs, ref_offs), rep
    finally:
        c.close()


def test_batched_learned_budget_growth():
    """A network that predicts huge counts asks for more slots than the first
    scratch budget holds: the items past the budget are redone exactly, and
    the next call has room for all of them."""
    kw = dict(precision=stq_amd.PRECISION_F32_CONFIRM)
    items = scenes.batched_scenes(count=96, seed=8, nmin=1e3, nmax=4e4)
    sc = [b for (_, _, b) in items]
    allb, offs = scenes.concat_scenes(sc)
    ref_pairs, ref_offs = _countfill(kw, allb, offs)
    c = stq_amd.Context(0, sizing=stq_amd.SIZING_LEARNED, train_scenes=-1, graph=0, **kw)
    try:
        w = np.zeros(mlp.NPARAMS, np.float32)
        w[-1] = 12.0
        c.sizer_set_params(w)
        need = int(ref_offs[-1])
        db = _lib.DeviceBuffer.from_array(allb)
        dout = _lib.DeviceBuffer(8 * need)
        fb = []
        for rep in range(3):
            rc, poffs = c.find_pairs_batched_device(db.ptr, offs, dout.ptr, need)
            assert rc == _lib.STQ_OK
            fb.append(c.stats()["n_fallback_tiles"])
            pairs = dout.to_array(np.int32, 2 * need).reshape(-1, 2)
            assert np.array_equal(pairs, ref_pairs) and np.array_equal(poffs, ref_offs), rep
        db.close()
        dout.close()
        print(f"{need} pairs; fallback items per call {fb}", flush=True)
        assert fb[0] > 0 and fb[1] == 0 and fb[2] == 0
    finally:
        c.close()
    check_scenes(ref_pairs, ref_offs, sc, "budget")


@pytest.mark.parametrize("change", [dict(tails=1, tail_cost=400), dict(claim=2), dict(balance=2),
                                    dict(sizing=stq_amd.SIZING_COUNT_FILL)])
def test_graph_follows_option_changes(change):
    """A captured graph is replayed only for the options it was captured with:
    after a change of tail items, claiming, work split or sizing, the calls
    give what an eager context with the new options gives (pairs, order and
    item count), and the old graph is never replayed."""
    sc = mixed_scenes(77, 11, nmax=30000)
    allb, offs = scenes.concat_scenes(sc)
    base = dict(precision=stq_amd.PRECISION_F32_CONFIRM, axis_policy=4,
                sizing=stq_amd.SIZING_LEARNED, train_scenes=1)
    c = stq_amd.Context(0, graph=0, **base)
    db = _lib.DeviceBuffer.from_array(allb)
    cap = 1 << 22
    dout = _lib.DeviceBuffer(8 * cap)
    try:
        states = []
        for _ in range(4):  # train/eager, eager, capture, replay
            rc, poffs = c.find_pairs_batched_device(db.ptr, offs, dout.ptr, cap)
            assert rc == _lib.STQ_OK
            states.append(c.stats()["graph"])
        assert states[-1] == 2, states
        new = {**base, **change}
        c.set_options(graph=0, **new)
        eager = stq_amd.Context(0, graph=2, **new)
        try:
            eager.find_pairs_batched(allb, offs)  # trains when learned
            ref_pairs, ref_offs = eager.find_pairs_batched(allb, offs)
            ref_items = eager.stats()["n_items"]
        finally:
            eager.close()
        after = []
        for _ in range(4):
            rc, poffs = c.find_pairs_batched_device(db.ptr, offs, dout.ptr, cap)
            assert rc == _lib.STQ_OK
            st = c.stats()
            after.append(st["graph"])
            pairs = dout.to_array(np.int32, 2 * int(poffs[-1])).reshape(-1, 2)
            assert np.array_equal(poffs, ref_offs), change
            assert np.array_equal(pairs, ref_pairs), change
            assert st["n_items"] == ref_items, change
        assert after[0] != 2, (change, after)
        print(f"graph states {states} -> {after}", flush=True)
    finally:
        dout.close()
        db.close()
        c.close()
