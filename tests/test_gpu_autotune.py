"""The autotuner on the GPU: tuned calls stay exact, the table is reused,
saved and loaded, and trial runs leave the learned-sizing state alone."""
import os

import numpy as np
import pytest

import stq_amd
from stq_amd import _lib, scenes
from oracle import oracle

pytestmark = pytest.mark.gpu


def same_set(got, ref):
    c = oracle.canonical(got)
    return len(c) == len(got) and np.array_equal(c, oracle.canonical(ref))


@pytest.mark.parametrize("prec", [stq_amd.PRECISION_F64, stq_amd.PRECISION_F32_CONFIRM])
def test_tune_reuse_save_load(tmp_path, prec):
    b = scenes.cloth_scene(g=120, seed=9, sigma=1.5e-2)
    ref = oracle.reference_set(b)
    c = stq_amd.Context(0, precision=prec, autotune=1)
    try:
        got = c.find_pairs(b)
        st = c.stats()
        print(f"tuned in {st['ms_tune']:.1f} ms", flush=True)
        assert st["ms_tune"] > 0 and c.autotune_entries() == 1
        assert same_set(got, ref)
        got2 = c.find_pairs(b)
        assert c.stats()["ms_tune"] == 0
        assert np.array_equal(got2, got)
        path = tmp_path / "tune.txt"
        c.autotune_save(path)
        text = path.read_text()
        assert text.startswith("stq-autotune 2\n") and len(text.splitlines()) == 2
        c.autotune_clear()
        assert c.autotune_entries() == 0
        c.autotune_load(path)
        assert c.autotune_entries() == 1
        got3 = c.find_pairs(b)
        assert c.stats()["ms_tune"] == 0
        assert np.array_equal(got3, got)
    finally:
        c.close()


def test_load_rejects_malformed(tmp_path):
    c = stq_amd.Context(0, autotune=1)
    try:
        bad = tmp_path / "bad.txt"
        for text in ["nonsense\n", "stq-autotune 2\n1 9 0 2048 8 32 200 1.0\n",
                     "stq-autotune 2\n1 4 300 2048 8 32 200 1.0\n", "stq-autotune 2\n1 4 0 2048 8\n",
                     "stq-autotune 2\n1 4 0 2048 8 32 -5 1.0\n", "stq-autotune 1\n"]:
            bad.write_text(text)
            with pytest.raises(stq_amd.StqError):
                c.autotune_load(bad)
            assert c.autotune_entries() == 0
        with pytest.raises(stq_amd.StqError):
            c.autotune_load(tmp_path / "missing.txt")
    finally:
        c.close()


def test_batched_tuning_is_exact():
    items = scenes.batched_scenes(count=40, seed=12, nmin=1e3, nmax=3e4)
    sc = [x for (_, _, x) in items]
    allb, offs = scenes.concat_scenes(sc)
    c = stq_amd.Context(0, precision=stq_amd.PRECISION_F32_CONFIRM, autotune=1)
    try:
        for rep in range(2):
            pairs, poffs = c.find_pairs_batched(allb, offs)
            for s, b in enumerate(sc):
                assert same_set(pairs[poffs[s]:poffs[s + 1]], oracle.reference_set(b)), (rep, s)
        assert c.autotune_entries() == 1
    finally:
        c.close()


def test_trials_do_not_count_as_training_scenes():
    """train_scenes = 2: calls 1 and 2 run count-then-fill (a count sweep is
    timed), call 3 the single learned pass (no count sweep), although the
    first call also ran a dozen trial passes."""
    b = scenes.uniform_scene(60_000, seed=31, extent_scale=4.0)
    ref = oracle.reference_set(b)
    c = stq_amd.Context(0, precision=stq_amd.PRECISION_F32_CONFIRM, sizing=stq_amd.SIZING_LEARNED,
                        train_scenes=2, autotune=1)
    try:
        count_ms = []
        for rep in range(3):
            got = c.find_pairs(b)
            assert same_set(got, ref), rep
            count_ms.append(c.stats()["ms_sweep_count"])
        assert count_ms[0] > 0 and count_ms[1] > 0 and count_ms[2] == 0, count_ms
    finally:
        c.close()
