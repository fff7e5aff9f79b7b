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


def _graph_ms(S, sess, plan, reps=10):
    for _ in range(2):
        plan.decode(S.MODE_GRAPH)
    sess.synchronize()
    a, b = S.HipEvent(), S.HipEvent()
    a.record(sess)
    for _ in range(reps):
        plan.decode(S.MODE_GRAPH)
    b.record(sess)
    sess.synchronize()
    return a.elapsed_ms(b) / reps


def extra_workloads(S, synth, sess, args, c2_streams, c2_cls):
    """The other decode workloads: the mixed-geometry batch, the 9/7 batch and
    the serving loop under Poisson arrivals (reported beside the headline)."""
    from signinum_amd.serving import Server, run_poisson
    res = {}
    # mixed sizes / bit depths / code-block sizes / levels / scales
    t0 = time.time()
    items = synth.mixed_batch(args.mixed_tiles)
    plan = S.TilePlan(sess, [i[0] for i in items], [i[1] for i in items], [i[2] for i in items])
    plan.autotune_order(reps=2)
    ms = _graph_ms(S, sess, plan)
    px = plan.stats()["output_pixels"]
    for k in range(0, len(items), max(1, len(items) // 16)):
        cs, fmt, scale, src = items[k]
        if int(scale) == 1:
            got = plan.surfaces[k].to_numpy()
            want = src if src.ndim == 3 else src[..., None]
            if not np.array_equal(got, want):
                raise SystemExit(f"mixed batch tile {k} differs from its source")
    res["mixed_batch"] = {
        "workload": f"{len(items)} tiles of 240/256/512/1024^2, 75% RGB8 / 25% 12-bit gray, "
                    "32/64 code-blocks, 3-6 levels, scale 1, 1/2, 1/4 (one mixed plan)",
        "mpix_s": round(px / (ms * 1e-3) / 1e6, 1), "ms_per_step": round(ms, 3),
        "output_mpix": round(px / 1e6, 2), "coded_mb": round(sum(len(i[0]) for i in items) / 1e6, 2)}
    plan.close()
    log(f"mixed batch: {res['mixed_batch']} ({time.time() - t0:.1f}s)")
    # irreversible 9/7 + ICT
    t0 = time.time()
    lossy = synth.lossy_batch(args.lossy_tiles)
    plan = S.TilePlan(sess, [c for c, _ in lossy], S.PixelFormat.RGB8)
    plan.autotune_order(reps=2)
    ms = _graph_ms(S, sess, plan)
    px = plan.stats()["output_pixels"]
    err = []
    for k in range(0, len(lossy), max(1, len(lossy) // 8)):
        d = plan.surfaces[k].to_numpy().astype(np.float64) - lossy[k][1]
        err.append((np.abs(d).max(), np.mean(d * d)))
    mse = float(np.mean([e[1] for e in err]))
    res["irreversible_batch"] = {
        "workload": f"{len(lossy)} x 512x512 RGB8 tissue, 9/7 + ICT, 5 levels, 64x64, "
                    "base step 1/64, 2 guard bits (f32 lifting)",
        "mpix_s": round(px / (ms * 1e-3) / 1e6, 1), "ms_per_step": round(ms, 3),
        "max_abs_vs_source": int(max(e[0] for e in err)),
        "psnr_vs_source_db": round(10 * np.log10(255.0 ** 2 / mse), 2) if mse > 0 else None}
    plan.close()
    log(f"irreversible batch: {res['irreversible_batch']} ({time.time() - t0:.1f}s)")
    # serving: Poisson arrivals over a pool of the workloads' tiles
    t0 = time.time()
    pool = [(c2_streams[i], S.PixelFormat.RGB8, S.Downscale.NONE, 256 * 256)
            for i in range(0, len(c2_streams), max(1, len(c2_streams) // 64))]
    pool += [(c, f, sc, 0) for c, f, sc, _ in items[:64]]
    pool += [(c, S.PixelFormat.RGB8, S.Downscale.NONE, 0) for c, _ in lossy[:16]]
    srv = Server(window_us=2000, hbm_budget=64 << 30)
    recs, wall = run_poisson(srv, pool, rate_per_s=args.serve_rate,
                             duration_s=args.serve_seconds, seed=7, max_tiles=64)
    srv.close()
    lat = np.array([r["latency_ms"] for r in recs])
    served_px = sum(r["pixels"] for r in recs)
    # the plain per-tile decode of the same pool, one launch sequence per tile
    plain_px, plain_ms = 0, 0.0
    for c, f, sc, _ in pool[::4]:
        p1 = S.TilePlan(sess, [c], f, sc)
        a, b = S.HipEvent(), S.HipEvent()
        p1.decode(S.MODE_PLAIN)
        a.record(sess)
        p1.decode(S.MODE_PLAIN)
        b.record(sess)
        sess.synchronize()
        plain_ms += a.elapsed_ms(b)
        plain_px += p1.stats()["output_pixels"]
        p1.close()
    res["serving"] = {
        "workload": f"open-loop Poisson arrivals at {args.serve_rate:g} requests/s for "
                    f"{args.serve_seconds:g} s, 1-64 tiles per request from a pool of "
                    f"{len(pool)} tiles (lossless 256^2 batch, mixed geometry, 9/7), 2 ms "
                    "window, 64 GiB device budget, outputs kept on the device",
        "requests": len(recs), "failed": int(sum(r["status"] != 0 for r in recs)),
        "p50_ms": round(float(np.percentile(lat, 50)), 3),
        "p99_ms": round(float(np.percentile(lat, 99)), 3),
        "max_ms": round(float(lat.max()), 3),
        "sustained_mpix_s": round(served_px / wall / 1e6, 1),
        "plain_per_tile_mpix_s": round(plain_px / (plain_ms * 1e-3) / 1e6, 1)}
    log(f"serving: {res['serving']} ({time.time() - t0:.1f}s)")
    return res


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
    ap.add_argument("--no-extras", action="store_true")
    ap.add_argument("--mixed-tiles", type=int, default=512)
    ap.add_argument("--lossy-tiles", type=int, default=256)
    ap.add_argument("--serve-seconds", type=float, default=2.0)
    ap.add_argument("--serve-rate", type=float, default=150.0)
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

    tune = order_ms = None
    order_name = "length"
    if not args.no_autotune:
        order_ms, chosen = plan.autotune_order(reps=3)
        order_name = {0: "length", 1: "input", 3: "geometry", 4: "geometry-cost"}[chosen]
        tune = plan.autotune(reps=3)
        log("autotune (code-block order, best variant ms): " +
            ", ".join(f"{k} {v:.3f}" for k, v in order_ms.items()) + f" -> {order_name}")
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
    es = 2 if st["narrow_samples"] else 4  # bytes per code-block sample in HBM
    alg_bytes = {
        # compressed bytes read + code-block samples written
        "blocks": st["coded_bytes"] + es * samples,
        # every band sample of the intermediate levels read once and the
        # intermediate low-pass outputs (int32) written once
        "levels": es * samples // 4 + 4 * samples // 16,
        # final level: its low-pass input (int32) and three detail bands read,
        # RGB8 written
        "final": 4 * samples // 4 + es * 3 * samples // 4 + samples,
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

    extras = {}
    if not args.no_extras:
        extras = extra_workloads(S, synth, sess, args, streams, cls)

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
        "code_block_order": order_name,
        "order_ms": None if order_ms is None else {k: round(v, 4) for k, v in order_ms.items()},
        "plan_ms": round(plan_ms, 1),
        "end_to_end_mpix_s": round(pixels / ((plan_ms + ms_per_step) * 1e-3) / 1e6, 1),
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    out.update(extras)
    print(json.dumps(out), flush=True)
    plan.close()
    sess.close()


if __name__ == "__main__":
    main()
