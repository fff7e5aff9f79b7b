"""Continuous-batching server: every request decodes the same alone and in
a mix of sizes, bit depths, formats, scales and transforms admitted while
others run; a corrupt tile fails only its own request; a small device
budget splits batches without losing requests."""
import numpy as np
import pytest

import signinum_amd as S
from signinum_amd import synth
from signinum_amd.serving import Server, run_poisson
from tests import oracle_ref as O

pytestmark = pytest.mark.gpu


def _pool():
    rng = np.random.default_rng(91)
    pool = []
    for k in range(14):
        if k % 4 == 3:
            img = synth.spots_gray12(rng, 96 + 8 * k, 80)
            cs = S.encode_j2k_ht(img, bit_depth=12, use_mct=0, cb_width_log2=5, cb_height_log2=5,
                                 levels=3 + k % 3)
            fmt = S.PixelFormat.GRAY16
        elif k % 4 == 2:
            img = synth.tissue_rgb(rng, 128, 96)
            cs = S.encode_j2k_ht(img, bit_depth=8, reversible=0, use_mct=1, levels=4)
            fmt = S.PixelFormat.RGB8
        else:
            img = synth.batch_tile(synth.SEED_BASE + 2, k, k % 3, 128 + 64 * (k % 2))
            cs = S.encode_j2k_lossless(img, levels=5)
            fmt = S.PixelFormat.RGB8 if k % 2 else S.PixelFormat.RGBA8
        pool.append((cs, fmt, S.Downscale((1, 2, 1, 4)[k % 4])))
    return pool


def _decode_alone(pool, tiles):
    srv = Server(window_us=0, max_inflight=1)
    rid = srv.submit([pool[i] for i in tiles])
    rec = srv.wait([rid])[rid]
    assert rec["status"] == 0, rec
    outs = [srv.download(rid, t, pool[i][1]) for t, i in enumerate(tiles)]
    srv.close()
    return outs


def test_solo_equals_mix(oracle):
    pool = _pool()
    rng = np.random.default_rng(5)
    requests = [list(rng.integers(0, len(pool), int(rng.integers(1, 9)))) for _ in range(24)]
    srv = Server(window_us=300, max_inflight=3, max_batch_tiles=24)
    ids = []
    for r in requests:
        ids.append(srv.submit([pool[i] for i in r]))
        srv.pump()  # requests arrive while earlier batches run
    recs = srv.wait(ids)
    st = srv.stats()
    assert st["batches"] >= 3, st
    for rid, r in zip(ids, requests[:8]):
        assert recs[rid]["status"] == 0
        solo = _decode_alone(pool, r)
        for t, i in enumerate(r):
            mixed = srv.download(rid, t, pool[i][1])
            assert np.array_equal(mixed, solo[t]), (rid, t)
            cs, fmt, den = pool[i]
            assert np.array_equal(mixed, O.decode(cs, int(fmt), int(den))), (rid, t)
    srv.close()


def test_corrupt_tile_fails_only_its_request(oracle):
    pool = _pool()
    good = pool[0]
    bad = bytearray(good[0])
    # the suffix-length field of the longest cleanup segment: Scup > Lcup
    j = max(S.code_block_jobs(bytes(bad)), key=lambda q: q.cleanup_length)
    bad[j.data_offset + j.cleanup_length - 1] = 0xFF
    srv = Server(window_us=5000, max_inflight=1)
    a = srv.submit([good, good])
    b = srv.submit([good, (bytes(bad), good[1], good[2]), good])
    c = srv.submit([pool[1]])
    recs = srv.wait([a, b, c])
    assert recs[a]["status"] == 0 and recs[c]["status"] == 0
    assert recs[b]["status"] != 0 and recs[b]["first_fail"] == 1
    ref = O.decode(good[0], int(good[1]), int(good[2]))
    assert np.array_equal(srv.download(a, 1, good[1]), ref)
    assert np.array_equal(srv.download(b, 2, good[1]), ref)
    srv.close()


def test_small_budget_splits_batches(oracle):
    pool = _pool()
    srv = Server(window_us=0, max_inflight=2, hbm_budget=8 << 20)
    ids = [srv.submit([pool[i % len(pool)] for i in range(k, k + 6)]) for k in range(10)]
    recs = srv.wait(ids)
    assert all(r["status"] == 0 for r in recs.values())
    assert srv.stats()["batches"] >= 3
    for rid in ids:
        srv.release(rid)
    assert srv.stats()["hbm_in_use"] == 0
    srv.close()


def test_poisson_load_reports_latency(oracle):
    pool = [(cs, fmt, den, 0) for cs, fmt, den in _pool()]
    srv = Server(window_us=2000)
    recs, wall = run_poisson(srv, pool, rate_per_s=200, duration_s=1.0, seed=3, max_tiles=16)
    assert recs and all(r["status"] == 0 for r in recs)
    lat = np.array([r["latency_ms"] for r in recs])
    print(f"{len(recs)} requests, p50 {np.percentile(lat, 50):.2f} ms, "
          f"p99 {np.percentile(lat, 99):.2f} ms, wall {wall:.2f} s")
    assert srv.stats()["hbm_in_use"] == 0
    srv.close()
