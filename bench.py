"""Benchmark: decoded MPix/s of the 4096 x 256x256 RGB8 HTJ2K lossless batch.

One step = one decode of the whole batch (HT block decode, inverse 5/3 DWT,
inverse RCT and RGB8 packing of 4096 tiles) from codestreams already
resident in device memory, replayed as a captured hipGraph.  Prints one
JSON line.  The plain one-launch-sequence-per-tile decode of the same batch
is timed beside it (the denominator of the batched speed-up), and the CPU
oracle is timed on a bounded sample on the host cores.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--tiles", type=int, default=4096)
    ap.add_argument("--plain-steps", type=int, default=1)
    ap.add_argument("--stage-reps", type=int, default=5)
    ap.add_argument("--cpu-sample", type=int, default=4096)
    ap.add_argument("--cpu-sample-1t", type=int, default=512)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-autotune", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build decodes on a single MI355X; "
              "multi-GPU runs are not part of it (tiles would shard by cost, with no "
              "collectives)", file=sys.stderr)
        sys.exit(2)

    import signinum_amd as S
    from signinum_amd import synth

    t0 = time.time()
    streams, cls = synth.lossless_batch(args.tiles, size=256)
    coded = sum(len(s) for s in streams)
    log(f"generated {args.tiles} tiles in {time.time() - t0:.1f}s, {coded / 1e6:.1f} MB coded")

    sess = S.HipSession(0)
    t0 = time.perf_counter()
    plan = S.TilePlan(sess, streams, S.PixelFormat.RGB8)
    sess.synchronize()
    plan_ms = (time.perf_counter() - t0) * 1e3
    st = plan.stats()
    pixels = st["output_pixels"]
    log(f"plan: {plan_ms:.1f} ms host parse + upload; {st['code_blocks']} code-blocks")

    tune = None
    if not args.no_autotune:
        tune = plan.autotune(reps=3)
        log("autotune (HT lane kernel, ms): " +
            ", ".join(f"{k} {v:.3f}" for k, v in tune.items()) +
            f" -> {S.TilePlan.lane_variants()[plan.lane_variant]}")

    def timed(fn, reps):
        # HIP events on the session stream, where the kernels are launched
        evs = [(S.HipEvent(), S.HipEvent()) for _ in range(reps)]
        for a, b in evs:
            a.record(sess)
            fn()
            b.record(sess)
        sess.synchronize()
        return [a.elapsed_ms(b) for a, b in evs]

    # ---- batched, graph-replayed ----
    for _ in range(args.warmup):
        plan.decode(S.MODE_GRAPH)
    sess.synchronize()
    t0 = time.perf_counter()
    step_ms = timed(lambda: plan.decode(S.MODE_GRAPH), args.steps)
    wall = time.perf_counter() - t0
    ms_per_step = float(np.mean(step_ms))
    mpix = pixels / (ms_per_step * 1e-3) / 1e6
    log(f"graph: {ms_per_step:.3f} ms/step (wall {wall * 1e3 / args.steps:.3f}), "
        f"{mpix:.0f} MPix/s")

    # ---- correctness of what was timed: sampled tiles equal their sources ----
    check = list(range(0, args.tiles, max(1, args.tiles // 16)))
    for i in check:
        src = synth.batch_tile(synth.SEED_BASE + 2, i, int(cls[i]), 256)
        if not np.array_equal(plan.surfaces[i].to_numpy(), src):
            raise SystemExit(f"bench output of tile {i} differs from its source pixels")

    # ---- per-stage kernel time ----
    stage_ms = {}
    for stage, name in ((0, "blocks"), (1, "levels"), (2, "final")):
        stage_ms[name] = float(np.mean(timed(lambda s=stage: plan.decode_stage(s),
                                             args.stage_reps)))
    log(f"stages ms: {stage_ms}")

    # ---- plain: one launch sequence per tile ----
    plain_ms = plain_mpix = None
    if args.plain_steps > 0:
        plain_ms = float(np.mean(timed(lambda: plan.decode(S.MODE_PLAIN), args.plain_steps)))
        plain_mpix = pixels / (plain_ms * 1e-3) / 1e6
        log(f"plain: {plain_ms:.1f} ms/step ({plain_mpix:.0f} MPix/s)")
    batched_ms = float(np.mean(timed(lambda: plan.decode(S.MODE_BATCHED), 3)))
    log(f"batched without graph: {batched_ms:.3f} ms/step")

    # ---- roofline of the dominant kernel ----
    dom = max(stage_ms, key=stage_ms.get)
    samples = pixels * 3
    alg_bytes = {
        # compressed bytes read + int32 coefficients written
        "blocks": st["coded_bytes"] + 4 * samples,
        # every band sample of the intermediate levels read once (4 B) and the
        # intermediate low-pass outputs written once
        "levels": 4 * samples // 4 + 4 * samples // 16,
        # final level: its four inputs read (4 B) and RGB8 written
        "final": 4 * samples + samples,
    }
    achieved = alg_bytes[dom] / (stage_ms[dom] * 1e-3) / 1e9
    kernel_name = {"blocks": "sgn_ht_rows", "levels": "sgn_idwt_level", "final": "sgn_final"}[dom]
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "hbm_traffic_4096x256_rgb8.json")
    if args.tiles == 4096 and os.path.exists(tpath):
        # HBM bytes per launch from rocprofv3 FETCH_SIZE / WRITE_SIZE runs of
        # this workload (scaled by a copy-kernel calibration)
        with open(tpath) as fh:
            tj = json.load(fh)
        lane = S.TilePlan.lane_variants()[plan.lane_variant]
        cands = [k for k in tj["kernels"].values() if k["kernel"].startswith(kernel_name) and
                 (dom != "blocks" or k.get("variant") == lane)]
        if cands:
            k = max(cands, key=lambda k: k["read_bytes"] + k["write_bytes"])
            traffic = int(k["read_bytes"] + k["write_bytes"])
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "kernel": dom, "algorithmic_bytes": int(alg_bytes[dom])}

    # ---- CPU baseline: the oracle on host cores, bounded sample ----
    cpu = None
    if not args.no_cpu_baseline:
        from tests import oracle_ref as O
        workers = min(16, os.cpu_count() or 1)
        n = min(args.cpu_sample, args.tiles)
        t0 = time.perf_counter()
        rc, _idx, _outs = O.decode_batch(streams[:n], 1, workers=workers)
        dt = time.perf_counter() - t0
        if rc != 0:
            raise SystemExit("oracle failed on the CPU sample")
        n1 = min(args.cpu_sample_1t, args.tiles)
        t0 = time.perf_counter()
        rc, _idx, _outs = O.decode_batch(streams[:n1], 1, workers=1)
        dt1 = time.perf_counter() - t0
        if rc != 0:
            raise SystemExit("oracle failed on the single-thread CPU sample")
        cpu = {"value": round(n * 256 * 256 / dt / 1e6, 1), "unit": "MPix/s", "cores": workers,
               "kind": "port",
               "sample": f"first {n} tiles of the batch, decode_tiles_into-style contiguous "
                         f"chunks over {workers} threads ({dt:.2f} s wall); single thread on "
                         f"the first {n1} tiles: {n1 * 256 * 256 / dt1 / 1e6:.1f} MPix/s "
                         f"({dt1:.2f} s)",
               "single_thread_mpix_s": round(n1 * 256 * 256 / dt1 / 1e6, 1)}
        log(f"cpu baseline: {cpu['value']} MPix/s on {workers} threads, "
            f"{cpu['single_thread_mpix_s']} MPix/s on one")

    out = {
        "metric": "decoded MPix/s, 4096x256^2 RGB8 HTJ2K batch vs plain GPU; p50/p99 latency",
        "value": round(mpix, 1),
        "unit": "MPix/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic (seeded near-white / tissue / texture tiles, encoded by the "
                "package's HTJ2K encoder)",
        "config": {"workload": f"{args.tiles} x 256x256 RGB8 HTJ2K lossless (5/3, RCT, 5 "
                               "levels, 64x64 code-blocks), batched decode replayed as a "
                               "hipGraph, codestreams resident in HBM",
                   "tiles": args.tiles, "coded_mb": round(coded / 1e6, 2)},
        "p50_ms": round(float(np.percentile(step_ms, 50)), 4),
        "p99_ms": round(float(np.percentile(step_ms, 99)), 4),
        "plain_mpix_s": None if plain_mpix is None else round(plain_mpix, 1),
        "plain_ms_per_step": None if plain_ms is None else round(plain_ms, 3),
        "speedup_vs_plain": None if plain_ms is None else round(plain_ms / ms_per_step, 2),
        "batched_no_graph_ms": round(batched_ms, 4),
        "stage_ms": {k: round(v, 4) for k, v in stage_ms.items()},
        "lane_kernel": S.TilePlan.lane_variants()[plan.lane_variant],
        "autotune_ms": None if tune is None else {k: round(v, 4) for k, v in tune.items()},
        "plan_ms": round(plan_ms, 1),
        "end_to_end_mpix_s": round(pixels / ((plan_ms + ms_per_step) * 1e-3) / 1e6, 1),
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    plan.close()
    sess.close()


if __name__ == "__main__":
    main()
