"""The two stream-shaped workloads at their full parameters, on the GPU,
against the oracle (full canonical pair sets, integer ids: bit-exact).

* The batched workload: 512 scenes of 1e3..5e5 boxes (log-uniform),
  uniform-like or cloth-like, density factor U[0.5, 4] -- 37.4M boxes, the
  same batch the benchmark times -- in the chosen form: reduced precision
  with FP64 confirmation, learned single pass (the network trained on the
  first call's exact counts), autotuned launch parameters, hipGraph capture
  and replay.  Every scene of the replay is compared with the oracle, and the
  replay with the eager run of the same batch (same pairs, same order).
* The learned-sizing stream: scenes.learned_stream_params (5e4..4e6 boxes,
  density factor log-uniform in [0.25, 8]); the first 200 scenes train the
  network online (count then fill), the next 10 use the single learned pass
  and are compared in full with the oracle; the fallback rate and the count
  error of those 10 are printed.
"""
import numpy as np
import pytest

import stq_amd
from oracle import oracle
from stq_amd import _lib, scenes

pytestmark = pytest.mark.gpu


def same_set(got, ref):
    c = oracle.canonical(got)
    return len(c) == len(got) and np.array_equal(c, ref)


@pytest.mark.timeout(900)
def test_batched_512_scenes_chosen_form_full_size():
    items = scenes.batched_scenes(count=512, seed=3, nmin=1e3, nmax=5e5)
    sc = [b for (_, _, b) in items]
    del items
    allb, offs = scenes.concat_scenes(sc)
    print(f"batch: {len(sc)} scenes, {len(allb)} boxes", flush=True)
    db = stq_amd.DeviceBuffer.from_array(allb)
    del allb
    ctx = stq_amd.Context(0, precision=stq_amd.PRECISION_F32_CONFIRM, axis_policy=stq_amd.AXIS_CHAIN,
                          sizing=stq_amd.SIZING_LEARNED, train_scenes=200, autotune=1)
    try:
        rc, poffs = ctx.find_pairs_batched_device(db.ptr, offs, None, 0)  # tunes, trains, sizes
        need = int(poffs[-1])
        dout = stq_amd.DeviceBuffer(8 * max(need, 1))
        outs, states = [], []
        for _ in range(3):  # eager, capture + launch, replay
            rc, p2 = ctx.find_pairs_batched_device(db.ptr, offs, dout.ptr, need)
            assert rc == _lib.STQ_OK and np.array_equal(p2, poffs)
            st = ctx.stats()
            states.append(st["graph"])
            outs.append(dout.to_array(np.int32, 2 * need).reshape(-1, 2))
        print(f"pairs {need}, graph states {states}, items {st['n_items']}, "
              f"fallback items {st['n_fallback_tiles']}, mae {st['sizing_mae']:.1f}", flush=True)
        assert states == [0, 1, 2]
        assert st["sizing"] == 1 and st["ms_sweep_count"] == 0.0  # the learned single pass
        assert np.array_equal(outs[0], outs[2]) and np.array_equal(outs[1], outs[2])
        pairs = outs[2]
        del outs
        bad = []
        for s, b in enumerate(sc):
            ref = oracle.reference_set(b) if len(b) > 1 else np.zeros((0, 2), np.int32)
            if not same_set(pairs[poffs[s]:poffs[s + 1]], ref):
                bad.append(s)
            if s % 128 == 127:
                print(f"compared {s + 1} scenes", flush=True)
        assert not bad, f"scenes differing from the oracle: {bad[:10]}"
        dout.close()
    finally:
        ctx.close()
        db.close()


@pytest.mark.timeout(900)
def test_learned_stream_at_its_parameters():
    params = scenes.learned_stream_params(count=2000)
    ntrain, nheld = 200, 10
    ctx = stq_amd.Context(0, precision=stq_amd.PRECISION_F32_CONFIRM, axis_policy=stq_amd.AXIS_CHAIN,
                          sizing=stq_amd.SIZING_LEARNED, train_scenes=ntrain)
    cap = 1 << 24
    dout = stq_amd.DeviceBuffer(8 * cap)
    items = fallback = 0
    abs_err = 0.0
    try:
        for i, (kind, n, rho, seed) in enumerate(params[:ntrain + nheld]):
            b = scenes.density_scene(kind, n, rho, seed)
            db = stq_amd.DeviceBuffer.from_array(b)
            try:
                cnt = ctx.find_pairs_device(db.ptr, len(b), dout.ptr, cap)
            except stq_amd.StqError as e:
                assert e.status == _lib.STQ_ERR_CAPACITY
                cnt = ctx.stats()["n_pairs"]
                dout.close()
                cap = int(cnt * 1.25) + 1
                dout = stq_amd.DeviceBuffer(8 * cap)
                cnt = ctx.find_pairs_device(db.ptr, len(b), dout.ptr, cap)
            st = ctx.stats()
            if i >= ntrain:
                assert st["sizing"] == 1 and st["ms_sweep_count"] == 0.0 and st["n_tiles"] > 0
                got = dout.to_array(np.int32, 2 * cnt).reshape(-1, 2)
                ref = oracle.reference_set(b)
                ok = same_set(got, ref)
                items += st["n_tiles"]
                fallback += st["n_fallback_tiles"]
                abs_err += st["sizing_mae"] * st["n_tiles"]
                print(f"held-out scene {i}: kind {kind} n={n} rho={rho:.2f} pairs={cnt} "
                      f"items={st['n_tiles']} fallback={st['n_fallback_tiles']} "
                      f"mae={st['sizing_mae']:.1f} exact={ok}", flush=True)
                assert ok, i
            elif i % 50 == 49:
                print(f"trained on {i + 1} scenes (mae of the last {st['sizing_mae']:.1f})", flush=True)
            db.close()
        print(f"held-out: {nheld} scenes, {items} items, fallback rate {fallback / max(items, 1):.5f}, "
              f"count MAE {abs_err / max(items, 1):.1f} pairs per item", flush=True)
    finally:
        dout.close()
        ctx.close()
