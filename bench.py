"""Benchmark of the gfx950 STQ broad phase on one MI355X.

Headline (the JSON line's `value`): the batched workload, 512 independent
scenes back to back (sizes log-uniform in [1e3, 5e5] boxes, uniform-like or
cloth-like, density factor U[0.5, 4]; 37,422,181 FP64 boxes, about 4.6e8
pairs), boxes resident in HBM, pairs written to a device buffer.  One step is
one batched call over all 512 scenes.  Forms on the same scenes:

  chosen      : sort keys and chain ends from the FP32 outward-rounded
                sweep-axis bounds quantised outward per scene (16-bit keys
                here), 14-bit outward levels with kind codes in the hot
                loop, FP64 confirmation of
                every candidate, learned sizing (one emitting pass into slots
                predicted by the network, trained online on the first call's
                exact counts), autotuned launch parameters, hipGraph replay
  plain       : the reference method on this library's kernels: FP64 keys
                and comparisons, axis x, count pass + fill pass, no tail
                items, default launch parameters (src/simulation.cu:555-782)
  plain_tuned : FP64 as plain, with tail items and autotuned launch
                parameters (the best this library does without reducing
                precision)

Timed region: K calls bracketed by device synchronisations on both sides;
ms_per_step is the host clock over them.  The roofline is that of the
dominant kernel, the emitting sweep of the chosen form (k_qsweep, learned
pass), timed by HIP events on the library's stream around that kernel inside
every timed call (event nodes of the replayed graph).  The CPU baseline is the
oracle's restatement of the reference CPU sweep (src/cpu/sweep.cpp) over all
512 scenes on this host's cores.

Further legs in the same line: the 708 x 708 swept cloth (3,001,923 boxes,
single-scene call), the 8M-box clustered heavy-tailed scene (`skewed`), a
bounded sample of the learned-sizing stream (`learned_stream`), the
PCIe-inclusive host-to-host call (`end_to_end_host`), the kernels either side
of the broad phase on the cloth mesh (`handoff`) and the 100,000-box uniform
scene with its whole CPU sweep (`small`).
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "broadphase-gpu_amd"))
sys.path.insert(0, ROOT)

METRIC = "boxes/s & pairs/s at exact pair set: chosen form vs plain FP64 STQ, 1 MI355X"
HBM_PEAK = 8.0e12
# one wave64 VALU instruction per cycle per CU (four 16-lane SIMDs, four cycles
# each), 256 CUs at 2.4 GHz: the issue ceiling of the comparison-bound sweep
VALU_ISSUE_PEAK = 256 * 2.4e9


def progress(msg):
    """One line on stderr when a leg ends (the JSON line alone goes to stdout)."""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--legs", default="batched,cloth,skewed,stream,handoff,small,host",
                    help="comma list of legs; the batched leg is the headline")
    ap.add_argument("--batched-forms", default="chosen,plain,plain_tuned,fp64_grouped,chosen_item_split",
                    help="forms of the batched leg (chosen is the headline)")
    ap.add_argument("--grid", type=int, default=708, help="cloth grid size g (g x g vertices)")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--forms", default="chosen,plain,plain_tuned,fp64_grouped", help="forms of the cloth leg")
    ap.add_argument("--cpu-sample", type=float, default=0.25,
                    help="cloth: fraction of sorted rows the CPU baseline sweeps (0 = skip)")
    ap.add_argument("--cpu-batched", type=int, default=1,
                    help="batched: run the CPU baseline over all scenes (0 = skip)")
    ap.add_argument("--no-check", action="store_true", help="skip the pair-set comparisons")
    ap.add_argument("--train-scenes", type=int, default=80,
                    help="cloth: scenes of the online training stream run before timing the chosen form")
    ap.add_argument("--batched-scenes", type=int, default=512,
                    help="scenes of the batched leg (sizes log-uniform in [1e3, 5e5])")
    ap.add_argument("--stream-scenes", type=int, default=240,
                    help="scenes of the learned-sizing stream leg: the first 200 train, the rest are "
                         "held out (sizes log-uniform in [5e4, 4e6])")
    ap.add_argument("--cpu-sample-1t", type=float, default=0.02,
                    help="cloth: fraction of sorted rows the 1-thread CPU timing sweeps (0 = skip)")
    ap.add_argument("--tune-table", default="",
                    help="autotune table: loaded when the file exists (no tuning), else written after "
                         "tuning; counter runs of one workload use it to share launch parameters")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="only run a device-to-device copy of this size (traffic calibration)")
    return ap.parse_args()


def cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def pair_digest(pairs):
    """Order-independent digest of a pair list: (count, sum of a 64-bit mix of
    each canonical (min, max) key, mod 2^64)."""
    import numpy as np
    p = np.asarray(pairs, dtype=np.int32).reshape(-1, 2)
    a = np.minimum(p[:, 0], p[:, 1]).astype(np.uint64)
    b = np.maximum(p[:, 0], p[:, 1]).astype(np.uint64)
    z = (a << np.uint64(32)) | b
    with np.errstate(over="ignore"):
        z = z + np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
        return len(p), int(z.sum(dtype=np.uint64))


def tuned_table(ctx):
    """The context's autotuning table as a list of dicts (read back through
    the library's own save format)."""
    import tempfile
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "tune.txt")
        ctx.autotune_save(path)
        with open(path) as fh:
            lines = fh.read().split("\n")[1:]
    keys = ("key", "axis_policy", "window", "chunk", "batch", "waves_per_cu", "tail_cost", "trial_ms")
    out = []
    for ln in lines:
        if ln.strip():
            v = ln.split()
            out.append(dict(zip(keys, [v[0]] + [int(x) for x in v[1:7]] + [float(v[7])])))
    return out


def load_traffic(name, form):
    """HBM bytes and VALU wave instructions per launch of a workload's dominant
    kernel, from the committed rocprof counter runs (profiles/<name>.json)."""
    path = os.path.join(ROOT, "profiles", name + ".json")
    try:
        with open(path) as fh:
            entry = json.load(fh).get(form, {})
        return entry.get("hbm_bytes_per_launch"), entry.get("valu_wave_instructions_per_launch")
    except (OSError, ValueError):
        return None, None


def roofline(alg_bytes, kernel_ms, kernel, traffic=None, valu=None):
    achieved = alg_bytes / (kernel_ms * 1e-3)
    issue = None
    if valu:
        rate = valu / (kernel_ms * 1e-3)
        issue = {"bound": "valu issue", "achieved": rate, "peak": VALU_ISSUE_PEAK,
                 "unit": "wave instructions/s", "frac": rate / VALU_ISSUE_PEAK, "valu_per_launch": valu}
    return {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
            "frac": achieved / HBM_PEAK, "traffic": traffic, "kernel": kernel,
            "algorithmic_bytes_per_launch": alg_bytes, "kernel_ms": kernel_ms, "issue": issue}


BATCHED_FORMS = {
    "chosen": dict(precision=1, axis_policy=4, sizing=1, train_scenes=200, autotune=1),
    "plain": dict(precision=0, axis_policy=0, tails=2),
    "plain_tuned": dict(precision=0, axis_policy=0, tails=1, autotune=1),
    # the chosen form's organisation and sizing with every test in FP64: what
    # the reduced precision alone gains
    "fp64_grouped": dict(precision=2, axis_policy=4, sizing=1, train_scenes=200, autotune=1),
    # the chosen form with the XCD slices cut at equal numbers of items instead
    # of equal predicted cost: what the prediction's work split gains
    "chosen_item_split": dict(precision=1, axis_policy=4, sizing=1, train_scenes=200, autotune=1,
                              balance=2),
}


def bench_batched(args, H):
    """The headline: 512 scenes back to back, device-resident, one batched
    call per step.  Each form: one sizing call (tunes, and for the chosen
    form trains the network on the batch's exact counts), warm-up calls (an
    eager run, then the capture), then `steps` timed calls (graph replays)
    between device synchronisations.  Per-scene pair sets are compared by
    digest across forms and with the oracle (the CPU baseline's own run)."""
    import numpy as np
    import stq_amd
    from stq_amd import _lib, scenes
    t0 = time.perf_counter()
    items = scenes.batched_scenes(count=args.batched_scenes, seed=3, nmin=1e3, nmax=5e5)
    sc = [b for (_, _, b) in items]
    allb, offs = scenes.concat_scenes(sc)
    del items
    gen_s = time.perf_counter() - t0
    db = _lib.DeviceBuffer.from_array(allb)
    nbox = len(allb)
    want = [f for f in args.batched_forms.split(",") if f]
    out, digests = {}, {}
    for name in want:
        kw = BATCHED_FORMS[name]
        ctx = stq_amd.Context(0, **kw)
        if kw.get("autotune") and args.tune_table and os.path.exists(args.tune_table):
            ctx.autotune_load(args.tune_table)
        rc, poffs = ctx.find_pairs_batched_device(db.ptr, offs, None, 0)  # sizes, tunes, trains
        train = ctx.stats()
        if kw.get("autotune") and args.tune_table and not os.path.exists(args.tune_table):
            ctx.autotune_save(args.tune_table)
        tuned = tuned_table(ctx) if kw.get("autotune") else None
        need = int(poffs[-1])
        dout = _lib.DeviceBuffer(8 * max(need, 1))
        for _ in range(max(args.warmup, 2)):  # eager, then capture
            ctx.find_pairs_batched_device(db.ptr, offs, dout.ptr, need)
        stats = []
        H.hipDeviceSynchronize()
        t1 = time.perf_counter()
        for _ in range(args.steps):
            rc, p2 = ctx.find_pairs_batched_device(db.ptr, offs, dout.ptr, need)
            stats.append(ctx.stats())
        H.hipDeviceSynchronize()
        t2 = time.perf_counter()
        assert rc == _lib.STQ_OK and np.array_equal(p2, poffs)
        ms = (t2 - t1) * 1e3 / max(args.steps, 1)
        st = stats[-1]
        sweep_ms = [x["ms_sweep_fill"] for x in stats]
        kernel_timing = "HIP events around the kernel inside each timed graph replay"
        if min(sweep_ms) <= 0.0:
            # the runtime recorded no events inside the graph: time the same
            # kernel in eager calls of the same batch
            ctx.set_options(**{**kw, "graph": 2})
            sweep_ms = []
            for _ in range(max(args.steps // 2, 3)):
                ctx.find_pairs_batched_device(db.ptr, offs, dout.ptr, need)
                sweep_ms.append(ctx.stats()["ms_sweep_fill"])
            kernel_timing = "HIP events around the kernel in eager calls of the same batch"
        phases = {k: float(np.mean([x[k] for x in stats])) for k in
                  ("ms_total", "ms_keys", "ms_sort", "ms_gather", "ms_plan", "ms_predict",
                   "ms_sweep_count", "ms_scan", "ms_sweep_fill", "ms_compact")}
        out[name] = dict(ms_per_step=ms, boxes_per_s=nbox / (ms * 1e-3), pairs_per_s=need / (ms * 1e-3),
                         stream_ms=float(np.median([x["ms_total"] for x in stats])),
                         sweep_kernel_ms=float(np.mean(sweep_ms)), sweep_kernel_timing=kernel_timing,
                         phases_ms=phases, pairs=need, items=st["n_items"],
                         candidates=st["n_candidates"], lane_steps=st["n_lane_steps"],
                         confirms=st["n_confirm"], graph_state=st["graph"],
                         fallback_items=st["n_fallback_tiles"],
                         mae_pairs_per_item=st["sizing_mae"] if kw.get("sizing") else None,
                         training_call_mae=train["sizing_mae"] if kw.get("sizing") else None,
                         first_call_ms=train["ms_total"], tune_ms=train["ms_tune"], autotuned=tuned,
                         form=kw)
        if not args.no_check:
            pairs = dout.to_array(np.int32, 2 * need).reshape(-1, 2)
            digests[name] = [pair_digest(pairs[poffs[s]:poffs[s + 1]]) for s in range(len(poffs) - 1)]
            del pairs
        dout.close()
        ctx.close()
        progress(f"batched form {name}: {ms:.2f} ms per batch, {need} pairs")
    db.close()
    res = {"workload": f"{args.batched_scenes} scenes, sizes log-uniform in [1e3, 5e5], uniform-like or "
                       f"cloth-like, density factor U[0.5, 4] (seed 3): {nbox} FP64 boxes, device-resident",
           "boxes": nbox, "scenes": len(offs) - 1, "generate_s": gen_s, "forms": out}
    if "chosen" in out:
        for base in ("plain", "plain_tuned", "fp64_grouped", "chosen_item_split"):
            if base in out:
                res[f"speedup_vs_{base}"] = out["chosen"]["boxes_per_s"] / out[base]["boxes_per_s"]
        ref = next(iter(digests.values()), None)
        res["per_scene_digests_equal_across_forms"] = (all(d == ref for d in digests.values())
                                                       if len(digests) > 1 else None)
    cpu = None
    if args.cpu_batched:
        from oracle import oracle
        threads = max(1, min(int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)), 64))
        secs, npairs, mism = 0.0, 0, 0
        for i, b in enumerate(sc):
            pairs, t = oracle.sweep_timed(b, axis=0, nthreads=threads)
            secs += t
            npairs += len(pairs)
            if digests and not args.no_check:
                mism += int(pair_digest(pairs) != next(iter(digests.values()))[i])
        one = None
        sub = sc[::32]
        if sub:
            t1s = sum(oracle.sweep_timed(b, axis=0, nthreads=1)[1] for b in sub)
            one = {"value": sum(len(b) for b in sub) / t1s, "unit": "boxes/s",
                   "sample": f"every 32nd scene ({len(sub)} scenes, {sum(len(b) for b in sub)} boxes), "
                             f"{t1s:.2f} s on 1 thread"}
        progress(f"batched cpu baseline: {nbox / secs:.3g} boxes/s on {threads} threads")
        cpu = {"value": nbox / secs, "unit": "boxes/s", "cores": threads, "kind": "port",
               "one_thread": one, "seconds": secs, "pairs": npairs,
               "sample": (f"the whole workload: all {len(sc)} scenes ({nbox} boxes) through the oracle's "
                          f"restatement of src/cpu/sweep.cpp (FP64, axis x; sort, chain sweep, merge per "
                          f"scene, {secs:.2f} s); CPU {cpu_model()}, {threads} threads"),
               "per_scene_digest_mismatches_vs_gpu": mism if digests and not args.no_check else None}
    res["cpu_baseline"] = cpu
    return res


def bench_skewed(args, H):
    """8M boxes around 64 Gaussian clusters (sigma 0.02) with Pareto(1.5)
    half-extents (min 5e-5, cap 0.05): a few very long chains.  The chosen
    form (reduced precision, autotuned, tail items) against plain FP64 with
    default parameters, both count-then-fill; the pair sets are compared."""
    import numpy as np
    import stq_amd
    from stq_amd import _lib, scenes
    b = scenes.clustered_scene()
    n = len(b)
    db = _lib.DeviceBuffer.from_array(b)
    del b
    forms = {"chosen": dict(precision=stq_amd.PRECISION_F32_CONFIRM, autotune=1),
             "plain": dict(precision=stq_amd.PRECISION_F64, axis_policy=stq_amd.AXIS_X, tails=2),
             "plain_tuned": dict(precision=stq_amd.PRECISION_F64, axis_policy=stq_amd.AXIS_X, tails=1,
                                 autotune=1),
             "fp64_grouped": dict(precision=stq_amd.PRECISION_F64_GROUPED, autotune=1)}
    out, dig = {}, {}
    for name, kw in forms.items():
        ctx = stq_amd.Context(0, **kw)
        try:
            ctx.find_pairs_device(db.ptr, n, None, 0)
        except stq_amd.StqError as e:
            if e.status != _lib.STQ_ERR_CAPACITY:
                raise
        tune = ctx.stats()["ms_tune"]
        cap = ctx.stats()["n_pairs"]
        dout = _lib.DeviceBuffer(8 * max(cap, 1))
        for _ in range(max(args.warmup, 1)):
            ctx.find_pairs_device(db.ptr, n, dout.ptr, cap)
        ms = []
        for _ in range(args.steps):
            ctx.find_pairs_device(db.ptr, n, dout.ptr, cap)
            ms.append(ctx.stats()["ms_total"])
        st = ctx.stats()
        t = float(np.median(ms))
        out[name] = dict(ms=t, boxes_per_s=n / (t * 1e-3), pairs=cap, candidates=st["n_candidates"],
                         lane_steps=st["n_lane_steps"], items=st["n_items"], tune_ms=tune,
                         autotuned=tuned_table(ctx) if kw.get("autotune") else None)
        if not args.no_check:
            dig[name] = pair_digest(dout.to_array(np.int32, 2 * cap).reshape(-1, 2))
        dout.close()
        ctx.close()
    db.close()
    return {"workload": f"{n} boxes, 64 Gaussian clusters (sigma 0.02), Pareto(1.5) half-extents in "
                        f"[5e-5, 0.05] (seed 2), device-resident",
            "forms": out,
            "speedup_vs": {k: out["chosen"]["boxes_per_s"] / v["boxes_per_s"] for k, v in out.items()
                           if k != "chosen"},
            "pair_sets_equal": (all(d == dig["plain"] for d in dig.values()) if len(dig) > 1 else None)}


def bench_stream(args, H):
    """The learned-sizing stream: scenes of 5e4 to 4e6 boxes, density factor
    log-uniform in [0.25, 8].  The first 200 run count-then-fill and train the
    network on their exact per-item counts; the rest use the single learned
    pass.  Reports the count error and the fallback rate on the held-out
    scenes, and checks every 8th of them against plain FP64 by digest; on
    those it also times the same precision with count-then-fill sizing, the
    form the learned single pass replaces."""
    import numpy as np
    import stq_amd
    from stq_amd import _lib, scenes
    params = scenes.learned_stream_params(count=2000)[:args.stream_scenes]
    ntrain = min(200, args.stream_scenes)
    ctx = stq_amd.Context(0, precision=stq_amd.PRECISION_F32_CONFIRM, axis_policy=stq_amd.AXIS_CHAIN,
                          sizing=stq_amd.SIZING_LEARNED, train_scenes=ntrain)
    plain = stq_amd.Context(0, precision=stq_amd.PRECISION_F64, tails=2)
    ctf = stq_amd.Context(0, precision=stq_amd.PRECISION_F32_CONFIRM, axis_policy=stq_amd.AXIS_CHAIN,
                          sizing=stq_amd.SIZING_COUNT_FILL)
    ctf_ms, learned_ms = 0.0, 0.0
    cap = 1 << 26
    dout = _lib.DeviceBuffer(8 * cap)
    rows = []
    t0 = time.perf_counter()
    checked, mismatches = 0, 0
    for i, (kind, nb, rho, seed) in enumerate(params):
        b = scenes.density_scene(kind, nb, rho, seed=seed)
        db = _lib.DeviceBuffer.from_array(b)
        try:
            cnt = ctx.find_pairs_device(db.ptr, len(b), dout.ptr, cap)
        except stq_amd.StqError as e:
            if e.status != _lib.STQ_ERR_CAPACITY:
                raise
            cnt = ctx.stats()["n_pairs"]
            dout.close()
            cap = int(cnt * 1.25)
            dout = _lib.DeviceBuffer(8 * cap)
            cnt = ctx.find_pairs_device(db.ptr, len(b), dout.ptr, cap)
        st = ctx.stats()
        rows.append((i, len(b), cnt, st["n_tiles"], st["n_fallback_tiles"], st["sizing_mae"], st["ms_total"]))
        if i >= ntrain and (i - ntrain) % 8 == 0 and not args.no_check:
            got = pair_digest(dout.to_array(np.int32, 2 * cnt).reshape(-1, 2))
            c2 = plain.find_pairs_device(db.ptr, len(b), dout.ptr, cap)
            ref = pair_digest(dout.to_array(np.int32, 2 * c2).reshape(-1, 2))
            c3 = ctf.find_pairs_device(db.ptr, len(b), dout.ptr, cap)
            ctf_ms += ctf.stats()["ms_total"]
            learned_ms += st["ms_total"]
            checked += 1
            mismatches += int(got != ref) + int(c3 != cnt)
        db.close()
        if (i + 1) % 100 == 0:
            progress(f"stream: {i + 1} of {args.stream_scenes} scenes")
    dout.close()
    ctx.close()
    plain.close()
    ctf.close()
    held = [r for r in rows if r[0] >= ntrain]
    tr = [r for r in rows if r[0] < ntrain]

    def agg(rs):
        if not rs:
            return None
        items = sum(r[3] for r in rs)
        return {"scenes": len(rs), "boxes": int(sum(r[1] for r in rs)), "pairs": int(sum(r[2] for r in rs)),
                "items": int(items), "fallback_items": int(sum(r[4] for r in rs)),
                "fallback_rate": sum(r[4] for r in rs) / max(items, 1),
                "mae_pairs_per_item": sum(r[5] * r[3] for r in rs) / max(items, 1),
                "ms_total": sum(r[6] for r in rs)}
    return {"workload": f"the first {args.stream_scenes} of the 2000 scenes of scenes.learned_stream_params "
                        f"(seed 4), 5e4..4e6 boxes log-uniform, density factor "
                        f"log-uniform in [0.25, 8]; first {ntrain} train the network online",
            "seconds": time.perf_counter() - t0, "training": agg(tr), "held_out": agg(held),
            "held_out_checked_vs_plain_fp64": checked, "mismatches": mismatches,
            "checked_ms": {"learned_single_pass": learned_ms, "count_then_fill": ctf_ms,
                           "speedup": ctf_ms / learned_ms if learned_ms > 0 else None}}


def bench_small(args, H):
    """The 100,000-box uniform scene (centres in the unit cube, half-extents
    10^U(-4,-2), 1-3 vertex ids from 60,000): the reduced-precision form
    (autotuned, count then fill; learned sizing has nothing to size at about
    1,000 pairs) and plain FP64 on the GPU, median of `steps` calls, and the
    oracle's CPU sweep of the whole scene on the host threads."""
    import numpy as np
    import stq_amd
    from stq_amd import _lib, scenes
    b = scenes.uniform_scene(100_000, seed=0)
    n = len(b)
    db = stq_amd.DeviceBuffer.from_array(b)
    forms = {"chosen": dict(precision=stq_amd.PRECISION_F32_CONFIRM, axis_policy=stq_amd.AXIS_CHAIN, autotune=1),
             "plain": dict(precision=stq_amd.PRECISION_F64, axis_policy=stq_amd.AXIS_X, tails=2)}
    out, sets = {"workload": f"uniform {n} boxes (seed 0)"}, {}
    for name, kw in forms.items():
        ctx = stq_amd.Context(0, **kw)
        try:
            ctx.find_pairs_device(db.ptr, n, None, 0)
        except stq_amd.StqError as e:
            if e.status != _lib.STQ_ERR_CAPACITY:
                raise
        cap = max(ctx.stats()["n_pairs"], 1)
        d_out = stq_amd.DeviceBuffer(8 * cap)
        for _ in range(max(args.warmup, 2)):
            ctx.find_pairs_device(db.ptr, n, d_out.ptr, cap)
        ms = []
        for _ in range(max(args.steps, 1)):
            cnt = ctx.find_pairs_device(db.ptr, n, d_out.ptr, cap)
            ms.append(ctx.stats()["ms_total"])
        t = float(np.median(ms))
        out[name] = {"ms": t, "boxes_per_s": n / (t * 1e-3), "pairs": cnt}
        sets[name] = pair_digest(d_out.to_array(np.int32, 2 * cnt))
        d_out.close()
        ctx.close()
    out["speedup_vs_plain_fp64"] = out["plain"]["ms"] / out["chosen"]["ms"]
    out["pair_sets_equal"] = sets["chosen"] == sets["plain"]
    if args.cpu_sample > 0:
        from oracle import oracle
        threads = max(1, min(int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)), 64))
        oracle.sweep(b, axis=0, nthreads=threads)  # warm
        t0 = time.perf_counter()
        pairs = oracle.sweep(b, axis=0, nthreads=threads)
        cpu_s = time.perf_counter() - t0
        out["cpu_baseline"] = {"ms": cpu_s * 1e3, "boxes_per_s": n / cpu_s, "cores": threads, "kind": "port",
                               "sample": "oracle restatement of src/cpu/sweep.cpp over the whole scene",
                               "pairs_equal_gpu": pair_digest(pairs) == sets["plain"]}
    db.close()
    return out


def bench_handoff(args, H):
    """The kernels either side of the broad phase on the cloth mesh: swept-box
    build (FP64) from the vertices at t0/t1 and the edge and face lists, then,
    on the broad phase's pairs, their orientation for the narrow phase and the
    256-byte query records.  Each kernel's time comes from HIP events on the
    library's stream (stats ms_total), median of `steps` calls; roofline
    against HBM with these algorithmic bytes per launch:
      build:   48 nv (V0, V1) + 8 ne + 12 nf (indices) + 72 n (boxes written)
      orient:  16 P (pairs read and written) + 12 n (vertex ids of the boxes)
      queries: 8 P (pairs) + 256 P (records written) + 12 n + 48 nv
    """
    import numpy as np
    import stq_amd
    from stq_amd import _lib, scenes

    rng = np.random.default_rng(args.seed)
    V0, E, F = scenes.cloth_mesh(args.grid)
    V1 = V0 + rng.normal(0.0, 2e-3, size=V0.shape)
    nv, ne, nf = len(V0), len(E), len(F)
    n = nv + ne + nf
    bufs = [stq_amd.DeviceBuffer.from_array(np.ascontiguousarray(a)) for a in
            (V0, V1, E.astype(np.int32), F.astype(np.int32))]
    d_boxes = stq_amd.DeviceBuffer(72 * n)
    ctx = stq_amd.Context(0, precision=stq_amd.PRECISION_F32_CONFIRM, axis_policy=stq_amd.AXIS_CHAIN)

    def timed(fn):
        for _ in range(max(args.warmup, 1)):
            fn()
        ms = []
        for _ in range(max(args.steps, 1)):
            fn()
            ms.append(ctx.stats()["ms_total"])
        return float(np.median(ms))

    def roof(alg, ms):
        gbs = alg / (ms * 1e-3) / 1e9
        return {"ms": ms, "algorithmic_bytes": alg, "achieved_GBs": gbs, "frac_hbm": gbs * 1e9 / HBM_PEAK}

    out = {"workload": f"cloth {args.grid}x{args.grid}: nv={nv}, ne={ne}, nf={nf}"}
    ms = timed(lambda: ctx.build_swept_boxes_device(bufs[0].ptr, bufs[1].ptr, nv, bufs[2].ptr, ne,
                                                    bufs[3].ptr, nf, d_boxes.ptr))
    out["build"] = roof(48.0 * nv + 8.0 * ne + 12.0 * nf + 72.0 * n, ms)
    if not args.no_check:
        # ids exactly, bounds by value: numpy's maximum/minimum may give the
        # other zero where -0.0 meets +0.0, which no overlap test can see
        ref = scenes.swept_boxes(V0, V1, E, F)
        got = d_boxes.to_array(scenes.AABB_DTYPE, n)
        out["build"]["equal_to_host_build"] = bool(all(np.array_equal(got[k], ref[k]) for k in
                                                       ("id", "min", "max", "vids", "ref_id")))
        del ref, got
    try:
        ctx.find_pairs_device(d_boxes.ptr, n, None, 0)
    except stq_amd.StqError as e:
        if e.status != _lib.STQ_ERR_CAPACITY:
            raise
    P = ctx.stats()["n_pairs"]
    d_pairs = stq_amd.DeviceBuffer(8 * max(P, 1))
    ctx.find_pairs_device(d_boxes.ptr, n, d_pairs.ptr, P)
    ms = timed(lambda: ctx.orient_pairs_device(d_boxes.ptr, n, d_pairs.ptr, P))
    out["orient"] = roof(16.0 * P + 12.0 * n, ms) | {"pairs": P}
    d_q = stq_amd.DeviceBuffer(_lib.CCD_QUERY_DTYPE.itemsize * max(P, 1))
    ms = timed(lambda: ctx.ccd_queries_device(d_boxes.ptr, n, bufs[0].ptr, bufs[1].ptr, nv, d_pairs.ptr, P,
                                              d_q.ptr))
    out["ccd_queries"] = roof(264.0 * P + 12.0 * n + 48.0 * nv, ms) | {"pairs": P}
    for b in bufs + [d_boxes, d_pairs, d_q]:
        b.close()
    ctx.close()
    H.hipDeviceSynchronize()
    return out


def bench_cloth(args, H):
    """The 708 x 708 swept cloth as one single-scene call: the chosen form
    (reduced precision, learned single pass trained on a stream of other
    scenes, autotuned) and plain FP64 count-then-fill; pair sets compared in
    full; roofline of the emitting sweep; CPU baseline on a sample of rows."""
    import numpy as np
    import stq_amd
    from stq_amd import _lib, scenes

    boxes = scenes.cloth_scene(g=args.grid, seed=args.seed)
    n = len(boxes)
    d_boxes = stq_amd.DeviceBuffer.from_array(boxes)
    forms = {
        "chosen": dict(precision=stq_amd.PRECISION_F32_CONFIRM, axis_policy=stq_amd.AXIS_CHAIN,
                       sizing=stq_amd.SIZING_LEARNED, train_scenes=max(args.train_scenes, 1),
                       autotune=1),
        "plain": dict(precision=stq_amd.PRECISION_F64, axis_policy=stq_amd.AXIS_X,
                      sizing=stq_amd.SIZING_COUNT_FILL, tails=2),
        "plain_tuned": dict(precision=stq_amd.PRECISION_F64, axis_policy=stq_amd.AXIS_X,
                            sizing=stq_amd.SIZING_COUNT_FILL, tails=1, autotune=1),
        "fp64_grouped": dict(precision=stq_amd.PRECISION_F64_GROUPED, axis_policy=stq_amd.AXIS_CHAIN,
                             sizing=stq_amd.SIZING_LEARNED, train_scenes=max(args.train_scenes, 1),
                             autotune=1),
    }
    want = [f for f in args.forms.split(",") if f]
    res, outputs, training = {}, {}, None
    for name in want:
        ctx = stq_amd.Context(0, **forms[name])
        table = args.tune_table if forms[name].get("autotune") else None
        loaded = bool(table) and os.path.exists(table)
        if loaded:  # same launch parameters as the run that wrote the table
            ctx.autotune_load(table)
        if forms[name]["sizing"] == stq_amd.SIZING_LEARNED:
            # the online training phase: exact counts of a stream of other
            # scenes (cloth-like and uniform-like, mixed sizes and densities)
            rng = np.random.default_rng(1234)
            maes, t0 = [], time.perf_counter()
            for i in range(max(args.train_scenes, 1)):
                nb = int(10 ** rng.uniform(np.log10(5e4), np.log10(6e5)))
                rho = float(10 ** rng.uniform(np.log10(0.25), np.log10(8.0)))
                sc = scenes.density_scene(i % 2, nb, rho, seed=50_000 + i)
                dsc = stq_amd.DeviceBuffer.from_array(sc)
                try:
                    ctx.find_pairs_device(dsc.ptr, len(sc), None, 0)
                except stq_amd.StqError as e:
                    if e.status != _lib.STQ_ERR_CAPACITY:
                        raise
                maes.append(ctx.stats()["sizing_mae"])
                dsc.close()
            if name == "chosen":
                training = {"scenes": len(maes), "seconds": time.perf_counter() - t0,
                            "mae_first": maes[0], "mae_last": maes[-1]}
        try:  # size the output once (the count is part of every step anyway)
            ctx.find_pairs_device(d_boxes.ptr, n, None, 0)
        except stq_amd.StqError as e:
            if e.status != _lib.STQ_ERR_CAPACITY:
                raise
        cap = ctx.stats()["n_pairs"]
        if table and not loaded:
            ctx.autotune_save(table)
        tuned = tuned_table(ctx) if forms[name].get("autotune") else None
        d_out = stq_amd.DeviceBuffer(8 * max(cap, 1))
        for _ in range(args.warmup):
            ctx.find_pairs_device(d_boxes.ptr, n, d_out.ptr, cap)
        H.hipDeviceSynchronize()
        sweeps, totals, fallback = [], [], []
        t0 = time.perf_counter()
        for _ in range(args.steps):
            cnt = ctx.find_pairs_device(d_boxes.ptr, n, d_out.ptr, cap)
            st = ctx.stats()
            sweeps.append(st["ms_sweep_fill"])
            totals.append(st["ms_total"])
            fallback.append(st["n_fallback_tiles"])
        H.hipDeviceSynchronize()
        t1 = time.perf_counter()
        ms = (t1 - t0) * 1e3 / max(args.steps, 1)
        res[name] = dict(ms_per_step=ms, boxes_per_s=n / (ms * 1e-3), pairs_per_s=cnt / (ms * 1e-3),
                         pairs=cnt, stream_ms=float(np.median(totals)),
                         fill_sweep_ms=float(np.mean(sweeps)), stats=st, form=forms[name],
                         fallback_items=int(np.max(fallback)) if fallback else 0, autotuned=tuned)
        if not args.no_check:
            outputs[name] = d_out.to_array(np.int32, 2 * cnt).reshape(-1, 2)
        d_out.close()
        ctx.close()
        progress(f"cloth form {name}: {ms:.2f} ms per call, {cnt} pairs")
    d_boxes.close()

    exact = None
    if not args.no_check and "plain" in outputs and len(outputs) > 1:
        def canon(p):
            key = (p[:, 0].astype(np.int64) << 32) | p[:, 1].astype(np.int64)
            key.sort()
            return key
        ref = canon(outputs["plain"])
        exact = all(bool(len(v) == len(ref) and np.array_equal(canon(v), ref))
                    for k, v in outputs.items() if k != "plain")
    del outputs

    main_form = "chosen" if "chosen" in res else want[0]
    r = res[main_form]
    P = r["pairs"]
    traffic, valu = load_traffic("sweep_traffic_cloth708", main_form)
    cpu = None
    if args.cpu_sample > 0:
        from oracle import oracle
        threads = max(1, min(int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)), 64))
        nrows = max(1, int(n * args.cpu_sample))
        t0 = time.perf_counter()
        pairs, sort_s = oracle.sweep(boxes, axis=0, nthreads=threads, nrows=nrows, return_sort_time=True)
        t1 = time.perf_counter()
        sweep_s = (t1 - t0) - sort_s
        est = sort_s + sweep_s * (n / nrows)
        one = None
        if args.cpu_sample_1t > 0:
            nr1 = max(1, int(n * args.cpu_sample_1t))
            t0 = time.perf_counter()
            _, sort1 = oracle.sweep(boxes, axis=0, nthreads=1, nrows=nr1, return_sort_time=True)
            sw1 = time.perf_counter() - t0 - sort1
            one = {"value": n / (sort1 + sw1 * (n / nr1)), "unit": "boxes/s",
                   "sample": f"sort of all boxes ({sort1:.2f} s) + sweep of the first {nr1} sorted rows "
                             f"({sw1:.2f} s), scaled to all rows"}
        progress(f"cloth cpu baseline: {n / est:.3g} boxes/s on {threads} threads")
        cpu = {"value": n / est, "unit": "boxes/s", "cores": threads, "kind": "port", "one_thread": one,
               "sample": (f"oracle restatement of src/cpu/sweep.cpp (FP64, axis x): parallel sort of all "
                          f"{n} boxes ({sort_s:.2f} s) + chain sweep of the first {nrows} sorted rows "
                          f"({sweep_s:.2f} s, {len(pairs)} pairs), scaled to all rows; "
                          f"CPU {cpu_model()}, {threads} threads")}
    return {
        "workload": f"swept cloth {args.grid}x{args.grid} (z=0.05 sin6x sin6y, step N(0,2e-3)), "
                    f"{n} FP64 boxes, device-resident in/out",
        "boxes": n, "pairs": P, "candidates_x": r["stats"]["n_candidates"],
        "exact_pair_set_vs_plain": exact,
        "forms": {k: {kk: vv for kk, vv in v.items() if kk != "stats"} | {"stats": v["stats"]}
                  for k, v in res.items()},
        "speedup_vs_plain_fp64": (res["chosen"]["boxes_per_s"] / res["plain"]["boxes_per_s"]
                                  if "chosen" in res and "plain" in res else None),
        "speedup_vs": {k: res["chosen"]["boxes_per_s"] / v["boxes_per_s"] for k, v in res.items()
                       if k != "chosen"} if "chosen" in res else None,
        "sizing": ({"training": training, "items": r["stats"]["n_tiles"],
                    "fallback_items": r["fallback_items"], "mae_pairs_per_item": r["stats"]["sizing_mae"]}
                   if training else None),
        "roofline": roofline(60.0 * n + 8.0 * P, r["fill_sweep_ms"],
                             "emitting sweep (k_qsweep learned pass for the chosen form)", traffic, valu),
        "cpu_baseline": cpu,
    }


def bench_host(args, H):
    """PCIe-inclusive: host boxes in, host pairs out through
    stq_find_pairs_host on the cloth (never the benchmark value)."""
    import numpy as np
    import stq_amd
    from stq_amd import scenes
    boxes = scenes.cloth_scene(g=args.grid, seed=args.seed)
    c2 = stq_amd.Context(0, precision=stq_amd.PRECISION_F32_CONFIRM, axis_policy=stq_amd.AXIS_CHAIN,
                         sizing=stq_amd.SIZING_COUNT_FILL, autotune=1)
    c2.find_pairs_host(boxes)
    ts = []
    for _ in range(3):
        t0 = time.perf_counter()
        c2.find_pairs_host(boxes)
        ts.append(time.perf_counter() - t0)
    c2.close()
    return {"ms": 1e3 * float(np.median(ts)),
            "what": "stq_find_pairs_host on the cloth, reduced precision, count then fill: H2D of the "
                    "boxes, the call, D2H of the pairs"}


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only (collectives, sharding and ranks "
                                   "are not part of it); run with --gpus 1",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    from stq_amd import _lib

    H = _lib.hip()
    if args.copy_probe_mib:
        nb = args.copy_probe_mib << 20
        a, b = _lib.DeviceBuffer(nb), _lib.DeviceBuffer(nb)
        for _ in range(5):
            _lib._check_hip(H.hipMemcpy(b.ptr, a.ptr, nb, _lib.HIP_D2D), "hipMemcpy D2D")
        H.hipDeviceSynchronize()
        print(json.dumps({"copy_probe_bytes": nb, "launches": 5}))
        return
    legs = [x for x in args.legs.split(",") if x]
    line = {"metric": METRIC}
    if "batched" in legs:
        bt = bench_batched(args, H)
        forms = bt["forms"]
        head = "chosen" if "chosen" in forms else next(iter(forms))
        r = forms[head]
        traffic, valu = load_traffic("sweep_traffic_batched512", head)
        line.update({
            "value": r["boxes_per_s"], "unit": "boxes/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": r["ms_per_step"], "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64" if head == "chosen" else "f64",
            "data": "synthetic",
            "config": {"workload": bt["workload"] + " (batched form: one call per step over all scenes)",
                       "boxes": bt["boxes"], "scenes": bt["scenes"], "pairs": r["pairs"], "form": head},
            "pairs_per_s": r["pairs_per_s"],
            "speedup_vs_plain_fp64": bt.get("speedup_vs_plain"),
            "speedup_vs_plain_fp64_tuned": bt.get("speedup_vs_plain_tuned"),
            "speedup_vs_fp64_grouped": bt.get("speedup_vs_fp64_grouped"),
            "work_split_by_predicted_cost_speedup": bt.get("speedup_vs_chosen_item_split"),
            "exact_pair_sets": {"chosen_vs_plain_forms": bt.get("per_scene_digests_equal_across_forms"),
                                "oracle_digest_mismatches": (bt["cpu_baseline"] or {}).get(
                                    "per_scene_digest_mismatches_vs_gpu")},
            "roofline": roofline(60.0 * bt["boxes"] + 8.0 * r["pairs"], r["sweep_kernel_ms"],
                                 "emitting sweep of the batch (k_qsweep learned pass)", traffic, valu)
                        | {"kernel_timing": r["sweep_kernel_timing"]},
            "cpu_baseline": bt["cpu_baseline"],
            "batched": bt,
        })
        progress("batched leg done")
    if "cloth" in legs:
        line["cloth"] = bench_cloth(args, H)
        progress("cloth leg done")
    if "host" in legs:
        line["end_to_end_host"] = bench_host(args, H)
        progress("host leg done")
    if "skewed" in legs:
        line["skewed"] = bench_skewed(args, H)
        progress("skewed leg done")
    if "stream" in legs and args.stream_scenes > 0:
        line["learned_stream"] = bench_stream(args, H)
        progress("learned stream leg done")
    if "handoff" in legs:
        line["handoff"] = bench_handoff(args, H)
        progress("hand-off leg done")
    if "small" in legs:
        line["small"] = bench_small(args, H)
        progress("small scene leg done")
    print(json.dumps(line))


if __name__ == "__main__":
    main()
