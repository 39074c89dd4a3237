"""The work split by estimated cost (options.balance) on the GPU.

Moving items between the XCD slices changes only which wave sweeps an item:
every pair, and the output order (fixed by the items' offsets), must be the
same as with equal numbers of items per slice, in every form and sizing mode,
single-scene and batched (eager, capture and replay).  Bar: bit-identical
output arrays, and the set equal to the oracle's.
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
SIZING = [("count", dict()), ("learned", dict(sizing=stq_amd.SIZING_LEARNED, train_scenes=1))]


def same_set(got, ref):
    g = oracle.canonical(got)
    return len(g) == len(got) and np.array_equal(g, oracle.canonical(ref))


@pytest.mark.parametrize("form,kw", FORMS)
@pytest.mark.parametrize("sizing,skw", SIZING)
def test_balance_keeps_pairs_and_order(form, kw, sizing, skw):
    # clusters of very different density: slices of equal item counts carry
    # very different work
    b = scenes.clustered_scene(n=150_000, seed=21, clusters=6)
    ref = oracle.reference_set(b)
    outs = {}
    for bal in (2, 1, 0):
        c = stq_amd.Context(0, balance=bal, **kw, **skw)
        try:
            first = c.find_pairs(b)       # count-then-fill (trains the network when learned)
            second = c.find_pairs(b)      # the learned single pass when learned
            assert np.array_equal(first, second), (form, sizing, bal)
            outs[bal] = second
        finally:
            c.close()
    assert same_set(outs[2], ref), (form, sizing)
    for bal in (1, 0):
        assert np.array_equal(outs[bal], outs[2]), (form, sizing, bal)


@pytest.mark.parametrize("form,kw", FORMS)
@pytest.mark.parametrize("sizing,skw", SIZING)
def test_balance_batched_keeps_pairs_and_order(form, kw, sizing, skw):
    rng = np.random.default_rng(22)
    sc = []
    for i in range(17):
        n = int(rng.integers(500, 30000))
        if i % 3 == 0:
            sc.append(scenes.cloth_scene(g=max(2, int(np.sqrt(n / 6))), seed=220 + i, sigma=2e-2))
        else:
            sc.append(scenes.uniform_scene(n, seed=220 + i, extent_scale=float(rng.uniform(1, 8))))
    allb, offs = scenes.concat_scenes(sc)
    outs = {}
    for bal in (2, 0):
        c = stq_amd.Context(0, balance=bal, axis_policy=4, **kw, **skw)
        try:
            runs = [c.find_pairs_batched(allb, offs) for _ in range(4)]  # train/eager, capture, replays
        finally:
            c.close()
        for pairs, poffs in runs[1:]:
            assert np.array_equal(pairs, runs[0][0]) and np.array_equal(poffs, runs[0][1]), (form, bal)
        outs[bal] = runs[-1]
    pairs, poffs = outs[2]
    for s, b in enumerate(sc):
        assert same_set(pairs[poffs[s]:poffs[s + 1]], oracle.reference_set(b)), (form, sizing, s)
    assert np.array_equal(outs[0][0], pairs) and np.array_equal(outs[0][1], poffs), (form, sizing)


def test_balance_option_is_validated():
    c = stq_amd.Context(0)
    try:
        with pytest.raises(stq_amd.StqError):
            c.set_options(balance=3)
    finally:
        c.close()
