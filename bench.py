"""Benchmark of the gfx950 STQ broad phase on one MI355X.

One step = one broad-phase call over the 708 x 708 swept cloth (3,001,923
FP64 boxes, ~48.4M pairs), boxes already resident in HBM, pairs written to a
device buffer.  Two forms run on the same scene:

  plain  : FP64 sort keys and comparisons, sweep axis x, exact count pass +
           fill pass (the reference method, src/simulation.cu:555-782)
  chosen : FP32 outward-rounded keys and chain ends, 15-bit outward levels in
           the sweep's hot loop, FP64 confirmation of every candidate,
           chain-length axis policy, and learned sizing: one emitting pass
           into per-item slots predicted by the network (trained online on
           the exact counts of a stream of other scenes before timing)

The JSON line reports the chosen form's boxes/s as `value`, its pairs/s, the
plain form's numbers and the speedup, whether both pair sets are identical,
the roofline of the dominant kernel (the emitting sweep) and a CPU baseline
(the oracle's restatement of the reference CPU sweep on this host).
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


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--grid", type=int, default=708, help="cloth grid size g (g x g vertices)")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--forms", default="chosen,plain")
    ap.add_argument("--cpu-sample", type=float, default=0.25,
                    help="fraction of sorted rows the CPU baseline sweeps (0 = skip)")
    ap.add_argument("--no-check", action="store_true", help="skip the pair-set comparison")
    ap.add_argument("--train-scenes", type=int, default=80,
                    help="scenes of the online training stream run before timing the chosen form")
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


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    import numpy as np
    import stq_amd
    from stq_amd import _lib, scenes

    H = _lib.hip()
    if args.copy_probe_mib:
        nb = args.copy_probe_mib << 20
        a, b = _lib.DeviceBuffer(nb), _lib.DeviceBuffer(nb)
        for _ in range(5):
            _lib._check_hip(H.hipMemcpy(b.ptr, a.ptr, nb, _lib.HIP_D2D), "hipMemcpy D2D")
        H.hipDeviceSynchronize()
        print(json.dumps({"copy_probe_bytes": nb, "launches": 5}))
        return

    boxes = scenes.cloth_scene(g=args.grid, seed=args.seed)
    n = len(boxes)
    d_boxes = stq_amd.DeviceBuffer.from_array(boxes)

    forms = {
        "chosen": dict(precision=stq_amd.PRECISION_F32_CONFIRM, axis_policy=stq_amd.AXIS_CHAIN,
                       sizing=stq_amd.SIZING_LEARNED, train_scenes=max(args.train_scenes, 1)),
        "plain": dict(precision=stq_amd.PRECISION_F64, axis_policy=stq_amd.AXIS_X,
                      sizing=stq_amd.SIZING_COUNT_FILL),
    }
    want = [f for f in args.forms.split(",") if f]
    res = {}
    outputs = {}
    training = None
    for name in want:
        ctx = stq_amd.Context(0, **forms[name])
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
            training = {"scenes": len(maes), "seconds": time.perf_counter() - t0,
                        "mae_first": maes[0], "mae_last": maes[-1]}
        # size the output once (the count is part of every step anyway)
        try:
            ctx.find_pairs_device(d_boxes.ptr, n, None, 0)
        except stq_amd.StqError as e:
            if e.status != _lib.STQ_ERR_CAPACITY:
                raise
        cap = ctx.stats()["n_pairs"]
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
                         fallback_items=int(np.max(fallback)) if fallback else 0)
        if not args.no_check:
            outputs[name] = d_out.to_array(np.int32, 2 * cnt).reshape(-1, 2)
        d_out.close()
        ctx.close()

    exact = None
    if not args.no_check and len(outputs) == 2:
        def canon(p):
            key = (p[:, 0].astype(np.int64) << 32) | p[:, 1].astype(np.int64)
            key.sort()
            return key
        a, b = outputs["chosen"], outputs["plain"]
        exact = bool(len(a) == len(b) and np.array_equal(canon(a), canon(b)))

    main_form = "chosen" if "chosen" in res else want[0]
    r = res[main_form]
    P = r["pairs"]
    alg_bytes = 60.0 * n + 8.0 * P
    achieved = alg_bytes / (r["fill_sweep_ms"] * 1e-3)
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "sweep_traffic_cloth708.json")
    if os.path.exists(tpath):
        try:
            with open(tpath) as fh:
                traffic = json.load(fh).get(main_form, {}).get("hbm_bytes_per_launch")
        except (OSError, ValueError):
            traffic = None

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
        cpu = {"value": n / est, "unit": "boxes/s", "cores": threads, "kind": "port",
               "sample": (f"oracle restatement of src/cpu/sweep.cpp (FP64, axis x): parallel sort of all "
                          f"{n} boxes ({sort_s:.2f} s) + chain sweep of the first {nrows} sorted rows "
                          f"({sweep_s:.2f} s, {len(pairs)} pairs), scaled to all rows; "
                          f"CPU {cpu_model()}, {threads} threads")}

    line = {
        "metric": METRIC,
        "value": r["boxes_per_s"],
        "unit": "boxes/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64" if main_form == "chosen" else "f64",
        "data": "synthetic",
        "config": {
            "workload": f"swept cloth {args.grid}x{args.grid} (z=0.05 sin6x sin6y, step N(0,2e-3)), "
                        f"{n} FP64 boxes, device-resident in/out",
            "boxes": n,
            "pairs": P,
            "candidates_x": r["stats"]["n_candidates"],
            "form": main_form,
        },
        "pairs_per_s": r["pairs_per_s"],
        "exact_pair_set_vs_plain": exact,
        "forms": {k: {kk: vv for kk, vv in v.items() if kk != "stats"} | {"stats": v["stats"]}
                  for k, v in res.items()},
        "speedup_vs_plain_fp64": (res["chosen"]["boxes_per_s"] / res["plain"]["boxes_per_s"]
                                  if "chosen" in res and "plain" in res else None),
        "sizing": ({"training": training, "items": r["stats"]["n_tiles"],
                    "fallback_items": r["fallback_items"], "mae_pairs_per_item": r["stats"]["sizing_mae"]}
                   if training else None),
        "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK, "traffic": traffic,
                     "kernel": "emitting sweep (k_qsweep learned pass for the chosen form)",
                     "algorithmic_bytes_per_launch": alg_bytes,
                     "kernel_ms": r["fill_sweep_ms"]},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
