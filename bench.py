"""Benchmark: decoded MPix/s of the 4096 x 256x256 RGB8 HTJ2K lossless batch.

One step = one decode of the whole batch (HT block decode, inverse 5/3 DWT,
inverse RCT and RGB8 packing of 4096 tiles) from codestreams resident in
device memory, replayed as a captured hipGraph; `value` is its decoded
MPix/s.  Beside it, in the same JSON line:

* end-to-end MPix/s: host codestream bytes (page-locked, and pageable) to
  device RGB8 surfaces, parse + upload + decode pipelined in chunks, the
  median of >= 10 timed repetitions after 3 warm-ups;
* the roofline of the dominant kernel by the algorithmic bytes of the
  decode (codestream bytes read + output bytes written per tile);
* the plain one-launch-sequence-per-tile decode of the same batch;
* the CPU oracle on the host cores (decode_tiles_into chunking over the
  available parallelism, and one thread);
* the single 512x512 tile (latency), the mixed-geometry batch (2048 tiles),
  the 9/7 batch (1024 tiles, with its error against the oracle's f32 and
  f64 decodes on a sample) and the continuous-batching server at several
  request rates.

Every tile of the headline batch and every full-resolution tile of the
mixed batch is compared with its source pixels.
"""
import argparse
import ctypes
import json
import math
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def available_parallelism():
    """Worker count of the reference's CPU tile batch
    (std::thread::available_parallelism, crates/signinum-j2k/src/batch.rs:198):
    the CPUs this process may run on, capped by a cgroup CPU quota."""
    n = len(os.sched_getaffinity(0))
    try:
        with open("/sys/fs/cgroup/cpu.max") as fh:
            quota, period = fh.read().split()[:2]
        if quota != "max":
            n = min(n, max(1, int(math.ceil(int(quota) / int(period)))))
    except (OSError, ValueError):
        pass
    return max(1, n)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def timed_events(S, sess, fn, reps):
    """Per-call milliseconds from HIP events on the session stream, where the
    kernels are launched."""
    evs = [(S.HipEvent(), S.HipEvent()) for _ in range(reps)]
    for a, b in evs:
        a.record(sess)
        fn()
        b.record(sess)
    sess.synchronize()
    return [a.elapsed_ms(b) for a, b in evs]


def graph_ms(S, sess, plan, reps=10):
    for _ in range(2):
        plan.decode(S.MODE_GRAPH)
    sess.synchronize()
    return float(np.median(timed_events(S, sess, lambda: plan.decode(S.MODE_GRAPH), reps)))


class PinnedBatch:
    """Codestreams back to back (16-byte aligned) in one page-locked host
    buffer: the input form whose bytes the library copies to the device by
    DMA without staging."""

    def __init__(self, streams):
        L = __import__("signinum_amd")._lib.load()
        self._L = L
        offs, total = [], 0
        for s in streams:
            offs.append(total)
            total += (len(s) + 15) & ~15
        p = ctypes.c_void_p()
        if L.sgn_host_alloc(max(1, total), ctypes.byref(p)) != 0:
            raise RuntimeError("sgn_host_alloc failed")
        self.ptr = p.value
        buf = (ctypes.c_uint8 * max(1, total)).from_address(self.ptr)
        view = np.frombuffer(buf, np.uint8)
        for s, o in zip(streams, offs):
            view[o:o + len(s)] = np.frombuffer(s, np.uint8)
        self.ptrs = [self.ptr + o for o in offs]
        self.lens = [len(s) for s in streams]

    def close(self):
        if self.ptr:
            self._L.sgn_host_free(self.ptr)
            self.ptr = None


def e2e_ms(S, sess, ptrs, lens, surfaces, reps, warmup):
    """Milliseconds of sgn_j2k_decode_tiles_device (host codestreams to
    device surfaces, synchronous) per repetition."""
    L = S._lib.load()
    n = len(ptrs)
    jobs = (S.TileJob * n)()
    for i, (p, ln, s) in enumerate(zip(ptrs, lens, surfaces)):
        jobs[i] = S.TileJob(p, ln, s.device_ptr, s.pitch_bytes, s.byte_len)
    ff = ctypes.c_size_t(0)
    out = []
    for r in range(warmup + reps):
        t0 = time.perf_counter()
        rc = L.sgn_j2k_decode_tiles_device(sess.handle, jobs, n, int(S.PixelFormat.RGB8), 1,
                                           S.MODE_BATCHED, ctypes.byref(ff))
        dt = (time.perf_counter() - t0) * 1e3
        if rc != 0:
            raise SystemExit(f"end-to-end decode failed at tile {ff.value}: {S._lib.last_error()}")
        if r >= warmup:
            out.append(dt)
    return out


def check_all_tiles(plan, sources, what):
    for i, src in enumerate(sources):
        if not np.array_equal(plan.surfaces[i].to_numpy(), src):
            raise SystemExit(f"{what}: tile {i} differs from its source pixels")


def config1(S, synth, sess, reps):
    """One 512x512 RGB8 tile (gradient + band-limited noise, 5/3 + RCT, 5
    levels, 64x64): latency of the graph replay (codestream in HBM), of the
    plain per-tile launch sequence and of the end-to-end host-to-surface
    decode."""
    rng = synth.tile_rng(synth.SEED_BASE + 1, 0)
    img = synth.gradient_noise_rgb(rng, 512, 512)
    cs = S.encode_j2k_lossless(img)
    plan = S.TilePlan(sess, [cs], S.PixelFormat.RGB8)
    g = graph_ms(S, sess, plan, reps)
    plain = float(np.median(timed_events(S, sess, lambda: plan.decode(S.MODE_PLAIN), reps)))
    if not np.array_equal(plan.surfaces[0].to_numpy(), img):
        raise SystemExit("512x512 tile differs from its source")
    surf = plan.surfaces
    keep = np.frombuffer(cs, np.uint8)
    e2e = e2e_ms(S, sess, [keep.ctypes.data], [len(cs)], surf, reps, 3)
    plan.close()
    px = 512 * 512
    return {"workload": "one 512x512 RGB8 tile, gradient + band-limited noise sigma 12, "
                        "HTJ2K 5/3 + RCT, 5 levels, 64x64 code-blocks",
            "coded_kb": round(len(cs) / 1e3, 1),
            "graph_ms": round(g, 4), "graph_mpix_s": round(px / (g * 1e-3) / 1e6, 1),
            "plain_ms": round(plain, 4), "plain_mpix_s": round(px / (plain * 1e-3) / 1e6, 1),
            "end_to_end_p50_ms": round(float(np.median(e2e)), 4),
            "end_to_end_p99_ms": round(float(np.percentile(e2e, 99)), 4)}


def mixed_config(S, synth, sess, n):
    items = synth.mixed_batch(n)
    plan = S.TilePlan(sess, [i[0] for i in items], [i[1] for i in items], [i[2] for i in items])
    plan.autotune_order(reps=2)
    ms = graph_ms(S, sess, plan)
    st = plan.stats()
    checked = 0
    for k, (cs, fmt, scale, src) in enumerate(items):
        if int(scale) == 1:
            got = plan.surfaces[k].to_numpy()
            want = src if src.ndim == 3 else src[..., None]
            if not np.array_equal(got, want):
                raise SystemExit(f"mixed batch tile {k} differs from its source")
            checked += 1
    plan.close()
    return {"workload": f"{n} tiles of 240/256/512/1024^2, 75% RGB8 / 25% 12-bit gray, "
                        "32/64 code-blocks, 3-6 levels, scale 1, 1/2, 1/4 (one mixed plan)",
            "mpix_s": round(st["output_pixels"] / (ms * 1e-3) / 1e6, 1),
            "ms_per_step": round(ms, 3), "output_mpix": round(st["output_pixels"] / 1e6, 2),
            "coded_mb": round(st["coded_bytes"] / 1e6, 2),
            "tiles_checked_vs_source": checked}, items


def lossy_config(S, synth, sess, n, sample):
    lossy = synth.lossy_batch(n)
    plan = S.TilePlan(sess, [c for c, _ in lossy], S.PixelFormat.RGB8)
    plan.autotune_order(reps=2)
    ms = graph_ms(S, sess, plan)
    st = plan.stats()
    mx, se, cnt = 0, 0.0, 0
    for k, (_c, src) in enumerate(lossy):
        d = plan.surfaces[k].to_numpy().astype(np.int32) - src
        mx = max(mx, int(np.abs(d).max()))
        se += float(np.sum(d.astype(np.float64) ** 2))
        cnt += d.size
    outs = [plan.surfaces[k].to_numpy() for k in range(min(sample, n))]
    plan.close()
    mse = se / cnt
    res = {"workload": f"{n} x 512x512 RGB8 tissue, 9/7 + ICT, 5 levels, 64x64, "
                       "base step 1/64, 2 guard bits (f32 lifting)",
           "mpix_s": round(st["output_pixels"] / (ms * 1e-3) / 1e6, 1),
           "ms_per_step": round(ms, 3),
           "max_abs_vs_source": mx,
           "psnr_vs_source_db": round(10 * math.log10(255.0 ** 2 / mse), 2) if mse > 0 else None}
    return res, lossy, outs


def classic_config(S, sess, sources, n, sample):
    """Classic (EBCOT / MQ) code-blocks: the first n tiles of the headline
    batch re-encoded by the package's classic encoder (one layer, every
    coding pass), decoded as one graph-replayed plan; every tile checked
    against its source; a sample timed through the oracle on one thread."""
    from concurrent.futures import ThreadPoolExecutor
    with ThreadPoolExecutor(16) as ex:
        streams = list(ex.map(S.encode_j2k_classic, sources[:n]))
    plan = S.TilePlan(sess, streams, S.PixelFormat.RGB8)
    ms = graph_ms(S, sess, plan)
    for k in range(n):
        if not np.array_equal(plan.surfaces[k].to_numpy(), sources[k]):
            raise SystemExit(f"classic batch tile {k} differs from its source")
    st = plan.stats()
    plan.close()
    res = {"workload": f"{n} x 256x256 RGB8 classic EBCOT/MQ lossless (5/3 + RCT, 5 levels, "
                       "64x64 code-blocks, one layer), the headline tiles re-encoded",
           "mpix_s": round(st["output_pixels"] / (ms * 1e-3) / 1e6, 1),
           "ms_per_step": round(ms, 3), "coded_mb": round(st["coded_bytes"] / 1e6, 2),
           "code_blocks": st["code_blocks"], "tiles_checked_vs_source": n}
    if sample:
        from tests import oracle_ref as O
        t0 = time.perf_counter()
        rc, _i, _o = O.decode_batch(streams[:sample], 1, workers=1)
        dt = time.perf_counter() - t0
        if rc != 0:
            raise SystemExit("oracle failed on the classic sample")
        res["oracle_single_thread_mpix_s"] = round(sample * 65536 / dt / 1e6, 2)
    return res


def refine_config(S, sess, sources, n):
    """HT code-blocks with refinement passes: the first n headline tiles
    re-encoded with a cleanup of the bits above bit 0 and SigProp / MagRef
    passes for bit 0 in every code-block (ht_refine=2), against the same
    tiles cleanup-only; the decode is checked against the oracle on a
    sample (the refined blocks drop the samples SigProp does not reach)."""
    from concurrent.futures import ThreadPoolExecutor
    with ThreadPoolExecutor(16) as ex:
        refined = list(ex.map(lambda t: S.encode_j2k_ht(t, ht_refine=2), sources[:n]))
        plain = list(ex.map(S.encode_j2k_lossless, sources[:n]))
    out = {}
    for name, streams in (("refined", refined), ("cleanup_only", plain)):
        plan = S.TilePlan(sess, streams, S.PixelFormat.RGB8)
        plan.autotune(reps=2)
        ms = graph_ms(S, sess, plan)
        st = plan.stats()
        if name == "refined":
            from tests import oracle_ref as O
            for k in range(0, n, max(1, n // 16)):
                if not np.array_equal(plan.surfaces[k].to_numpy(), O.decode(streams[k], 1)):
                    raise SystemExit(f"refined tile {k} differs from the oracle")
        out[name] = {"mpix_s": round(st["output_pixels"] / (ms * 1e-3) / 1e6, 1),
                     "ms_per_step": round(ms, 3), "coded_mb": round(st["coded_bytes"] / 1e6, 2),
                     "lane_kernel": S.TilePlan.lane_variants()[plan.lane_variant]}
        plan.close()
    return {"workload": f"{n} x 256x256 RGB8 HTJ2K 5/3, every code-block with cleanup + "
                        "SigProp + MagRef passes, beside the same tiles cleanup-only",
            **out}


def serving(S, sess, pool, rates, seconds):
    from signinum_amd.serving import Server, run_poisson
    out = []
    for rate in rates:
        srv = Server(window_us=2000, hbm_budget=64 << 30)
        recs, wall = run_poisson(srv, pool, rate_per_s=rate, duration_s=seconds, seed=7,
                                 max_tiles=64)
        srv.close()
        lat = np.array([r["latency_ms"] for r in recs])
        px = sum(r["pixels"] for r in recs)
        row = {"rate_per_s": rate, "requests": len(recs),
               "failed": int(sum(r["status"] != 0 for r in recs)),
               "p50_ms": round(float(np.percentile(lat, 50)), 3),
               "p99_ms": round(float(np.percentile(lat, 99)), 3),
               "sustained_mpix_s": round(px / wall / 1e6, 1)}
        out.append(row)
        log(f"serving at {rate:g} req/s: {row}")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--tiles", type=int, default=4096)
    ap.add_argument("--plain-steps", type=int, default=1)
    ap.add_argument("--stage-reps", type=int, default=5)
    ap.add_argument("--e2e-reps", type=int, default=10)
    ap.add_argument("--cpu-sample-1t", type=int, default=512)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--no-extras", action="store_true")
    ap.add_argument("--mixed-tiles", type=int, default=2048)
    ap.add_argument("--lossy-tiles", type=int, default=1024)
    ap.add_argument("--lossy-oracle-sample", type=int, default=16)
    ap.add_argument("--classic-tiles", type=int, default=1024)
    ap.add_argument("--classic-oracle-sample", type=int, default=32)
    ap.add_argument("--serve-seconds", type=float, default=2.0)
    ap.add_argument("--serve-rates", default="150,600,2000,6000")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only copy a MIB MiB device buffer three times (the plain copy "
                         "kernel whose FETCH_SIZE / WRITE_SIZE calibrate the HBM counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build decodes on a single MI355X; "
              "multi-GPU runs are not part of it (tiles would shard by cost, with no "
              "collectives)", file=sys.stderr)
        sys.exit(2)

    import signinum_amd as S
    from signinum_amd import synth

    if args.copy_calibration:
        sess = S.HipSession(0)
        side = 1024
        rows = args.copy_calibration * (1 << 20) // (side * 4)
        a = S.DeviceSurface(sess, side, rows, S.PixelFormat.RGBA8)
        b = S.DeviceSurface(sess, side, rows, S.PixelFormat.RGBA8)
        a.fill(1)
        for _ in range(3):
            b.copy_from(a)
        sess.synchronize()
        log(f"copied {args.copy_calibration} MiB three times")
        sess.close()
        return

    t0 = time.time()
    streams, sources, cls = synth.lossless_batch_with_pixels(args.tiles, size=256)
    coded = sum(len(s) for s in streams)
    log(f"generated {args.tiles} tiles in {time.time() - t0:.1f}s, {coded / 1e6:.1f} MB coded")

    sess = S.HipSession(0)
    t0 = time.perf_counter()
    plan = S.TilePlan(sess, streams, S.PixelFormat.RGB8)
    sess.synchronize()
    plan_ms = (time.perf_counter() - t0) * 1e3
    st = plan.stats()
    pixels = st["output_pixels"]
    out_bytes = st["output_bytes"]
    log(f"plan: {plan_ms:.1f} ms (parse {st['host_parse_ms']:.1f}, merge "
        f"{st['host_merge_ms']:.1f}, stage {st['host_stage_ms']:.1f}, copy "
        f"{st['host_copy_ms']:.1f}); {st['code_blocks']} code-blocks")

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
        dwt = plan.autotune_dwt(reps=5)
        log("autotune (synthesis-level shape, ms): " +
            ", ".join(f"{k} {v:.3f}" for k, v in dwt.items()) +
            f" -> {S.TilePlan.dwt_shapes()[plan.dwt_shape]}")

    # ---- batched, graph-replayed, codestreams resident in HBM ----
    for _ in range(args.warmup):
        plan.decode(S.MODE_GRAPH)
    sess.synchronize()
    t0 = time.perf_counter()
    step_ms = timed_events(S, sess, lambda: plan.decode(S.MODE_GRAPH), args.steps)
    wall = time.perf_counter() - t0
    ms_per_step = float(np.mean(step_ms))
    mpix = pixels / (ms_per_step * 1e-3) / 1e6
    log(f"graph: {ms_per_step:.3f} ms/step (wall {wall * 1e3 / args.steps:.3f}), "
        f"{mpix:.0f} MPix/s")

    # ---- correctness of what was timed: every tile equals its source ----
    t0 = time.time()
    check_all_tiles(plan, sources, "headline batch")
    log(f"all {args.tiles} tiles equal their sources ({time.time() - t0:.1f}s)")

    # ---- per-stage kernel time ----
    stage_ms = {}
    for stage, name in ((0, "blocks"), (1, "levels"), (2, "final")):
        stage_ms[name] = float(np.median(timed_events(S, sess, lambda s=stage: plan.decode_stage(s),
                                                      args.stage_reps)))
    log(f"stages ms: {stage_ms}")

    # ---- plain: one launch sequence per tile ----
    plain_ms = plain_mpix = None
    if args.plain_steps > 0:
        plain_ms = float(np.mean(timed_events(S, sess, lambda: plan.decode(S.MODE_PLAIN),
                                              args.plain_steps)))
        plain_mpix = pixels / (plain_ms * 1e-3) / 1e6
        log(f"plain: {plain_ms:.1f} ms/step ({plain_mpix:.0f} MPix/s)")
    batched_ms = float(np.median(timed_events(S, sess, lambda: plan.decode(S.MODE_BATCHED), 3)))
    log(f"batched without graph: {batched_ms:.3f} ms/step")

    # ---- end to end: host codestreams -> device surfaces ----
    pinned = PinnedBatch(streams)
    e2e_pin = e2e_ms(S, sess, pinned.ptrs, pinned.lens, plan.surfaces, args.e2e_reps, 3)
    keep = [np.frombuffer(s, np.uint8) for s in streams]
    e2e_page = e2e_ms(S, sess, [k.ctypes.data for k in keep], [len(s) for s in streams],
                      plan.surfaces, args.e2e_reps, 3)
    pinned.close()
    for i in range(0, args.tiles, max(1, args.tiles // 64)):
        if not np.array_equal(plan.surfaces[i].to_numpy(), sources[i]):
            raise SystemExit(f"end-to-end decode: tile {i} differs from its source")
    e2e = {"pinned_p50_ms": round(float(np.median(e2e_pin)), 3),
           "pinned_mpix_s": round(pixels / (np.median(e2e_pin) * 1e-3) / 1e6, 1),
           "pageable_p50_ms": round(float(np.median(e2e_page)), 3),
           "pageable_mpix_s": round(pixels / (np.median(e2e_page) * 1e-3) / 1e6, 1),
           "reps": args.e2e_reps,
           "pinned_ms": [round(x, 3) for x in e2e_pin]}
    log(f"end to end: {e2e}")

    # ---- roofline of the dominant kernel: codestream bytes + output bytes ----
    dom = max(stage_ms, key=stage_ms.get)
    alg = st["coded_bytes"] + out_bytes
    achieved = alg / (stage_ms[dom] * 1e-3) / 1e9
    samples = pixels * 3
    es = 2 if st["narrow_samples"] else 4
    stage_bytes = {
        "blocks": st["coded_bytes"] + es * samples,
        "levels": es * samples // 4 + 4 * samples // 16,
        "final": 4 * samples // 4 + es * 3 * samples // 4 + samples,
    }
    kernel_name = {"blocks": "sgn_ht_rows", "levels": "sgn_idwt_level", "final": "sgn_final"}[dom]
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "hbm_traffic_4096x256_rgb8.json")
    if args.tiles == 4096 and os.path.exists(tpath):
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
                "kernel": dom, "kernel_ms": round(stage_ms[dom], 4),
                "algorithmic_bytes": int(alg),
                "frac_step": round(alg / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                "stage_bytes": {k: int(v) for k, v in stage_bytes.items()},
                "stage_gbs": {k: round(stage_bytes[k] / (stage_ms[k] * 1e-3) / 1e9, 1)
                              for k in stage_ms}}

    # ---- CPU baseline: the oracle on the host cores ----
    cpu = None
    if not args.no_cpu_baseline:
        from tests import oracle_ref as O
        workers = available_parallelism()
        t0 = time.perf_counter()
        rc, _idx, _outs = O.decode_batch(streams, 1, workers=workers)
        dt = time.perf_counter() - t0
        if rc != 0:
            raise SystemExit("oracle failed on the CPU batch")
        n1 = min(args.cpu_sample_1t, args.tiles)
        t0 = time.perf_counter()
        rc, _idx, _outs = O.decode_batch(streams[:n1], 1, workers=1)
        dt1 = time.perf_counter() - t0
        if rc != 0:
            raise SystemExit("oracle failed on the single-thread CPU sample")
        cpu = {"value": round(pixels / dt / 1e6, 1), "unit": "MPix/s", "cores": workers,
               "kind": "port",
               "cpu_model": cpu_model(), "os_cpu_count": os.cpu_count(),
               "sample": f"all {args.tiles} tiles of the batch, decode_tiles_into-style "
                         f"contiguous chunks over {workers} threads (the available "
                         f"parallelism; {dt:.2f} s wall); single thread on the first {n1} "
                         f"tiles ({dt1:.2f} s)",
               "single_thread_mpix_s": round(n1 * 256 * 256 / dt1 / 1e6, 1)}
        log(f"cpu baseline: {cpu['value']} MPix/s on {workers} threads "
            f"({cpu['cpu_model']}), {cpu['single_thread_mpix_s']} MPix/s on one")

    extras = {}
    if not args.no_extras:
        t0 = time.time()
        extras["single_tile"] = config1(S, synth, sess, 20)
        log(f"single tile: {extras['single_tile']} ({time.time() - t0:.1f}s)")
        t0 = time.time()
        extras["mixed_batch"], items = mixed_config(S, synth, sess, args.mixed_tiles)
        log(f"mixed batch: {extras['mixed_batch']} ({time.time() - t0:.1f}s)")
        t0 = time.time()
        extras["irreversible_batch"], lossy, outs = lossy_config(S, synth, sess, args.lossy_tiles,
                                                                 args.lossy_oracle_sample)
        if not args.no_cpu_baseline and outs:
            # the oracle's f32 and f64 decodes of a sample, timed as the CPU
            # baseline of this workload and compared with the GPU output
            from tests import oracle_ref as O
            k = len(outs)
            t0c = time.perf_counter()
            rc, _i, ref32 = O.decode_batch([c for c, _ in lossy[:k]], 1, workers=1)
            dt32 = time.perf_counter() - t0c
            if rc != 0:
                raise SystemExit("oracle failed on the 9/7 sample")
            lsb, mx64, se64, cnt = 0, 0.0, 0.0, 0
            for j in range(k):
                r32 = ref32[j].reshape(outs[j].shape)
                lsb = max(lsb, int(np.abs(outs[j].astype(np.int32) -
                                          r32.astype(np.int32)).max()))
                f64 = np.clip(O.decode_planes(lossy[j][0], use_f64=True), 0, 255)
                d = outs[j].astype(np.float64) - f64.transpose(1, 2, 0)
                mx64 = max(mx64, float(np.abs(d).max()))
                se64 += float(np.sum(d * d))
                cnt += d.size
            mse64 = se64 / cnt
            extras["irreversible_batch"]["vs_oracle"] = {
                "sample_tiles": k, "max_lsb_vs_f32_oracle": lsb,
                "max_abs_vs_f64_oracle": round(mx64, 4),
                "psnr_vs_f64_oracle_db": round(10 * math.log10(255.0 ** 2 / mse64), 2)
                if mse64 > 0 else None,
                "oracle_single_thread_mpix_s": round(k * 512 * 512 / dt32 / 1e6, 2)}
            if lsb > 1:
                raise SystemExit(f"9/7 output differs from the oracle by {lsb} LSB")
        log(f"irreversible batch: {extras['irreversible_batch']} ({time.time() - t0:.1f}s)")
        t0 = time.time()
        extras["classic_batch"] = classic_config(
            S, sess, sources, min(args.classic_tiles, len(sources)),
            0 if args.no_cpu_baseline else args.classic_oracle_sample)
        log(f"classic batch: {extras['classic_batch']} ({time.time() - t0:.1f}s)")
        t0 = time.time()
        extras["refinement_batch"] = refine_config(S, sess, sources,
                                                   min(args.classic_tiles, len(sources)))
        log(f"refinement batch: {extras['refinement_batch']} ({time.time() - t0:.1f}s)")
        t0 = time.time()
        pool = [(streams[i], S.PixelFormat.RGB8, S.Downscale.NONE, 256 * 256)
                for i in range(0, len(streams), max(1, len(streams) // 64))]
        pool += [(c, f, sc, 0) for c, f, sc, _ in items[:64]]
        pool += [(c, S.PixelFormat.RGB8, S.Downscale.NONE, 0) for c, _ in lossy[:16]]
        rates = [float(r) for r in args.serve_rates.split(",") if r]
        sweep = serving(S, sess, pool, rates, args.serve_seconds)
        # plain per-tile decode of the same pool, one launch sequence per tile
        plain_px, plain_t = 0, 0.0
        for c, f, sc, _ in pool[::4]:
            p1 = S.TilePlan(sess, [c], f, sc)
            p1.decode(S.MODE_PLAIN)
            plain_t += float(np.median(timed_events(S, sess, lambda: p1.decode(S.MODE_PLAIN), 3)))
            plain_px += p1.stats()["output_pixels"]
            p1.close()
        extras["serving"] = {
            "workload": f"open-loop Poisson arrivals for {args.serve_seconds:g} s per rate, "
                        f"1-64 tiles per request from a pool of {len(pool)} tiles (lossless "
                        "256^2 batch, mixed geometry, 9/7), 2 ms window, 64 GiB device budget, "
                        "outputs kept on the device",
            "rates": sweep,
            "saturation_mpix_s": max(r["sustained_mpix_s"] for r in sweep),
            "plain_per_tile_mpix_s": round(plain_px / (plain_t * 1e-3) / 1e6, 1)}
        log(f"serving done ({time.time() - t0:.1f}s)")

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
        "device_only_mpix_s": round(mpix, 1),
        "end_to_end_mpix_s": e2e["pinned_mpix_s"],
        "end_to_end": e2e,
        "plan_ms": round(plan_ms, 1),
        "plain_mpix_s": None if plain_mpix is None else round(plain_mpix, 1),
        "plain_ms_per_step": None if plain_ms is None else round(plain_ms, 3),
        "speedup_vs_plain": None if plain_ms is None else round(plain_ms / ms_per_step, 2),
        "batched_no_graph_ms": round(batched_ms, 4),
        "stage_ms": {k: round(v, 4) for k, v in stage_ms.items()},
        "lane_kernel": S.TilePlan.lane_variants()[plan.lane_variant],
        "autotune_ms": None if tune is None else {k: round(v, 4) for k, v in tune.items()},
        "dwt_shape": S.TilePlan.dwt_shapes()[plan.dwt_shape],
        "code_block_order": order_name,
        "order_ms": None if order_ms is None else {k: round(v, 4) for k, v in order_ms.items()},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    out.update(extras)
    print(json.dumps(out), flush=True)
    plan.close()
    sess.close()


if __name__ == "__main__":
    main()
