"""Learned sizing on the GPU: the HIP network against the numpy oracle, and the
single learned sweep pass against the CPU pair-set oracle.

Tolerances: forward and gradient in FP32 against the FP64 oracle,
|gpu - cpu| <= 1e-4 (relative to the largest magnitude of the vector) plus
1e-6 absolute; the pair sets are compared exactly.
"""
import numpy as np
import pytest

import stq_amd
from oracle import mlp, oracle
from stq_amd import scenes

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = stq_amd.Context(0)
    yield c
    c.close()


def _rand(seed, n):
    rng = np.random.default_rng(seed)
    w = rng.normal(0, 0.2, mlp.NPARAMS).astype(np.float32)
    x = rng.normal(0, 1, (n, mlp.NF)).astype(np.float32)
    t = rng.normal(2, 1, n).astype(np.float32)
    return w, x, t


def _close(a, b, rel=1e-4, atol=1e-6):
    scale = max(np.abs(b).max(), 1e-30)
    assert np.abs(np.asarray(a, np.float64) - b).max() <= rel * scale + atol, np.abs(a - b).max() / scale


def test_forward_matches_oracle(ctx):
    w, x, _ = _rand(0, 1000)
    ctx.sizer_set_params(w)
    _close(ctx.sizer_forward(x), mlp.forward(w, x))


@pytest.mark.parametrize("n", [1, 31, 32, 33, 500, 5000])
def test_gradient_matches_oracle(ctx, n):
    w, x, t = _rand(1 + n, n)
    ctx.sizer_set_params(w)
    loss, g = ctx.sizer_gradient(x, t)
    L, G = mlp.gradient(w, x, t)
    assert abs(loss - L) <= 1e-4 * abs(L) + 1e-6
    _close(g, G)


def test_adam_step_matches_oracle(ctx):
    w, x, t = _rand(7, 3000)
    ctx.sizer_reset()
    ctx.sizer_set_params(w)
    ctx.sizer_train(x, t, steps=1, lr=1e-3)
    _, G = mlp.gradient(w, x, t)
    w1, _, _ = mlp.adam_step(w.astype(np.float64), G, np.zeros_like(G), np.zeros_like(G), 1e-3, 1)
    _close(ctx.sizer_params(), w1, rel=1e-5)


def _stream(kind_seed, count, nmin=20_000, nmax=120_000):
    rng = np.random.default_rng(kind_seed)
    out = []
    for i in range(count):
        n = int(rng.integers(nmin, nmax))
        rho = float(10 ** rng.uniform(np.log10(0.5), np.log10(4)))
        out.append(scenes.density_scene(i % 2, n, rho, seed=kind_seed * 1000 + i))
    return out


@pytest.mark.parametrize("prec", [stq_amd.PRECISION_F64, stq_amd.PRECISION_F32_CONFIRM])
def test_learned_pipeline_is_exact(ctx, prec):
    ctx.sizer_reset()
    ctx.set_options(precision=prec, sizing=stq_amd.SIZING_LEARNED, train_scenes=4)
    for i, b in enumerate(_stream(21, 8)):
        got = ctx.find_pairs(b)
        st = ctx.stats()
        ref = oracle.reference_set(b)
        print(f"scene {i} n={len(b)} pairs={len(ref)} items={st['n_tiles']} fallback={st['n_fallback_tiles']} "
              f"mae={st['sizing_mae']:.2f}", flush=True)
        assert np.array_equal(oracle.canonical(got), ref), i
        assert len(oracle.canonical(got)) == len(got)


def test_forced_fallback_is_exact(ctx):
    """Predictions of zero pairs give every item the slack (256 slots): items
    with more pairs overflow and are redone exactly."""
    ctx.sizer_reset()
    w = np.zeros(mlp.NPARAMS, np.float32)
    w[-1] = -5.0
    ctx.sizer_set_params(w)
    ctx.set_options(precision=stq_amd.PRECISION_F32_CONFIRM, sizing=stq_amd.SIZING_LEARNED, train_scenes=-1)
    b = scenes.cloth_scene(g=150, seed=5, sigma=1e-2)
    got = ctx.find_pairs(b)
    st = ctx.stats()
    assert st["n_fallback_tiles"] > 0
    assert np.array_equal(oracle.canonical(got), oracle.reference_set(b))
    # same order as the count-then-fill path
    ctx.set_options(precision=stq_amd.PRECISION_F32_CONFIRM)
    assert np.array_equal(ctx.find_pairs(b), got)


def test_scratch_limit_overflow_is_exact():
    """One scene's learned pass holds exactly its predicted slots; with
    options.scratch_limit_mb below that, the items whose slots end beyond the
    limit get none and are redone exactly.  Outputs with and without the
    limit equal the count-then-fill path's, order included."""
    c = stq_amd.Context(0)
    try:
        w = np.zeros(mlp.NPARAMS, np.float32)
        w[-1] = 6.0  # log1p of the prediction: about 400 pairs per item
        c.sizer_set_params(w)
        b = scenes.cloth_scene(g=300, seed=8, sigma=2e-3)
        kw = dict(precision=stq_amd.PRECISION_F32_CONFIRM, sizing=stq_amd.SIZING_LEARNED, train_scenes=-1)
        c.set_options(**kw)
        free = c.find_pairs(b)
        f_free = c.stats()["n_fallback_tiles"]
        c.set_options(**kw, scratch_limit_mb=8)  # one million pair slots
        limited = c.find_pairs(b)
        f_lim = c.stats()["n_fallback_tiles"]
        c.set_options(precision=stq_amd.PRECISION_F32_CONFIRM)
        ref = c.find_pairs(b)
        print(f"fallback items: {f_free} without a limit, {f_lim} with 8 MiB", flush=True)
        assert np.array_equal(free, ref)
        assert np.array_equal(limited, ref)
        assert f_lim > f_free
    finally:
        c.close()


def test_learned_capacity_reported(ctx):
    """The learned single pass into a caller's device buffer that is too small:
    the call reports the count needed and writes nothing (compaction and the
    exact refill of overflowed items check the total on the device); with that
    capacity the set is exact and in the count-then-fill order."""
    ctx.sizer_reset()
    w = np.zeros(mlp.NPARAMS, np.float32)
    w[-1] = -5.0  # predictions of about zero pairs: many items overflow too
    ctx.sizer_set_params(w)
    ctx.set_options(precision=stq_amd.PRECISION_F32_CONFIRM, sizing=stq_amd.SIZING_LEARNED, train_scenes=-1)
    b = scenes.cloth_scene(g=150, seed=6, sigma=1e-2)
    ref = oracle.reference_set(b)
    db = stq_amd.DeviceBuffer.from_array(b)
    sentinel = np.full(2 * 1000, -7, np.int32)
    small = stq_amd.DeviceBuffer.from_array(sentinel)
    with pytest.raises(stq_amd.StqError) as e:
        ctx.find_pairs_device(db.ptr, len(b), small.ptr, 1000)
    assert e.value.status == 3
    st = ctx.stats()
    assert st["n_pairs"] == len(ref) > 1000  # nothing fits, so no item is compacted or refilled
    assert np.array_equal(small.to_array(np.int32, 2 * 1000), sentinel)
    out = stq_amd.DeviceBuffer(8 * len(ref))
    cnt = ctx.find_pairs_device(db.ptr, len(b), out.ptr, len(ref))
    got = out.to_array(np.int32, 2 * cnt).reshape(-1, 2)
    assert cnt == len(ref) and ctx.stats()["n_fallback_tiles"] > 0
    assert np.array_equal(oracle.canonical(got), ref)
    ctx.set_options(precision=stq_amd.PRECISION_F32_CONFIRM)
    assert np.array_equal(ctx.find_pairs(b), got)
    for x in (db, small, out):
        x.close()


def test_trained_network_beats_untrained(ctx):
    """Train on a stream, then size a held-out scene: fewer fallback items and a
    smaller count error than the untrained network.  Each network is measured
    on the second of two calls, once the scratch budget (which follows the
    slots the predictions ask for) has adapted, so only prediction misses
    count as fallbacks."""
    held = scenes.cloth_scene(g=220, seed=77, sigma=1.5e-2)
    ref = oracle.reference_set(held)

    def measure():
        ctx.set_options(precision=stq_amd.PRECISION_F32_CONFIRM, sizing=stq_amd.SIZING_LEARNED, train_scenes=-1)
        for _ in range(2):
            got = ctx.find_pairs(held)
            assert np.array_equal(oracle.canonical(got), ref)
        st = ctx.stats()
        return st["n_fallback_tiles"] / max(st["n_tiles"], 1), st["sizing_mae"]

    ctx.sizer_reset()
    f0, m0 = measure()
    ctx.set_options(precision=stq_amd.PRECISION_F32_CONFIRM, sizing=stq_amd.SIZING_LEARNED, train_scenes=1 << 30)
    for b in _stream(33, 24, 30_000, 250_000):
        ctx.find_pairs(b)
    f1, m1 = measure()
    print(f"untrained: fallback rate {f0:.3f}, MAE {m0:.1f}; trained: fallback rate {f1:.3f}, MAE {m1:.1f}", flush=True)
    assert m1 < m0
    assert f1 <= f0
