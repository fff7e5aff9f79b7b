"""Batch decodes on the device at the workloads' real shapes: every tile of
a graph-replayed batch against its source, the pipelined host-to-device
batch path (pageable and page-locked inputs, chunk boundaries, the first
failing tile), large tiles at reduced resolution against the oracle, and
the >8-bit to Gray8 conversion against its closed form."""
import ctypes
import os

import numpy as np
import pytest

import signinum_amd as S
from signinum_amd import _lib, synth
from tests import oracle_ref as O

pytestmark = pytest.mark.gpu


def _gray8_closed_form(v, bit_depth):
    """lib.rs:2282-2293: f32 v / (2^bd - 1) * 255, rounded half away from
    zero, saturating to u8."""
    x = (v.astype(np.float32) / np.float32((1 << bit_depth) - 1)) * np.float32(255.0)
    r = np.where(x >= 0, np.floor(x + np.float32(0.5)), np.ceil(x - np.float32(0.5)))
    return np.clip(r, 0, 255).astype(np.uint8)


def test_gray8_from_12bit_matches_closed_form(session, oracle):
    rng = np.random.default_rng(1212)
    img = synth.spots_gray12(rng, 133, 90)
    img[0, :3] = [0, 4095, 2048]
    cs = S.encode_j2k_ht(img, bit_depth=12, use_mct=0, levels=4, cb_width_log2=5,
                         cb_height_log2=5)
    out = S.J2kDecoder(cs, session).decode_to_numpy(S.PixelFormat.GRAY8)
    assert np.array_equal(out[..., 0], _gray8_closed_form(img, 12))


def test_graph_batch_every_tile_equals_source(session, oracle):
    """256 tiles of the 256x256 lossless batch content in one plan, replayed
    twice as a graph: all 256 outputs equal their sources each time."""
    streams, pix, _cls = synth.lossless_batch_with_pixels(256, size=256)
    plan = S.TilePlan(session, streams, S.PixelFormat.RGB8)
    for _ in range(2):
        for s in plan.surfaces:
            s.fill(0)
        plan.decode(S.MODE_GRAPH)
        session.synchronize()
        for i, (s, img) in enumerate(zip(plan.surfaces, pix)):
            assert np.array_equal(s.to_numpy(), img), i
    plan.close()


def _jobs(ptrs, lens, surfaces):
    jobs = (S.TileJob * len(ptrs))()
    for i, (p, n, s) in enumerate(zip(ptrs, lens, surfaces)):
        jobs[i] = S.TileJob(p, n, s.device_ptr, s.pitch_bytes, s.byte_len)
    return jobs


def _decode_tiles(session, jobs, n):
    ff = ctypes.c_size_t(0)
    rc = _lib.load().sgn_j2k_decode_tiles_device(session.handle, jobs, n,
                                                 int(S.PixelFormat.RGB8), 1, S.MODE_BATCHED,
                                                 ctypes.byref(ff))
    return rc, ff.value


@pytest.fixture
def small_chunks():
    old = os.environ.get("SGN_PIPELINE_TILES")
    os.environ["SGN_PIPELINE_TILES"] = "7"
    yield
    if old is None:
        os.environ.pop("SGN_PIPELINE_TILES", None)
    else:
        os.environ["SGN_PIPELINE_TILES"] = old


def test_pipelined_batch_pageable_and_pinned(session, oracle, small_chunks):
    """60 tiles in chunks of 7 (9 chunks over both decode streams): from
    separate pageable buffers (staged), and back to back in page-locked
    memory (one direct copy per chunk)."""
    streams, pix, _cls = synth.lossless_batch_with_pixels(60, size=128)
    surfs = S._alloc_surfaces(session, [(128, 128)] * 60, [S.PixelFormat.RGB8] * 60)
    keep = [np.frombuffer(s, np.uint8) for s in streams]
    rc, _ = _decode_tiles(session, _jobs([k.ctypes.data for k in keep],
                                         [len(s) for s in streams], surfs), 60)
    assert rc == 0, _lib.last_error()
    for i, (s, img) in enumerate(zip(surfs, pix)):
        assert np.array_equal(s.to_numpy(), img), i
        s.fill(0)
    L = _lib.load()
    offs, total = [], 0
    for s in streams:
        offs.append(total)
        total += (len(s) + 15) & ~15
    p = ctypes.c_void_p()
    assert L.sgn_host_alloc(total, ctypes.byref(p)) == 0
    try:
        view = np.frombuffer((ctypes.c_uint8 * total).from_address(p.value), np.uint8)
        for s, o in zip(streams, offs):
            view[o:o + len(s)] = np.frombuffer(s, np.uint8)
        rc, _ = _decode_tiles(session, _jobs([p.value + o for o in offs],
                                             [len(s) for s in streams], surfs), 60)
        assert rc == 0, _lib.last_error()
        for i, (s, img) in enumerate(zip(surfs, pix)):
            assert np.array_equal(s.to_numpy(), img), i
    finally:
        L.sgn_host_free(p.value)


@pytest.mark.parametrize("bad", [3, 25])
def test_pipelined_batch_first_failing_tile(session, oracle, small_chunks, bad):
    """A truncated tile in the first or a later chunk: the batch fails with
    that tile's index; the tiles of earlier chunks were decoded."""
    streams, pix, _cls = synth.lossless_batch_with_pixels(30, size=64)
    streams[bad] = streams[bad][:60]
    surfs = S._alloc_surfaces(session, [(64, 64)] * 30, [S.PixelFormat.RGB8] * 30)
    keep = [np.frombuffer(s, np.uint8) for s in streams]
    rc, ff = _decode_tiles(session, _jobs([k.ctypes.data for k in keep],
                                          [len(s) for s in streams], surfs), 30)
    assert rc != 0 and ff == bad
    for i in range(0, (bad // 7) * 7):
        assert np.array_equal(surfs[i].to_numpy(), pix[i]), i


def test_large_tiles_reduced_resolution(session, oracle):
    """512x512 and 1024x1024 tiles of the mixed-geometry workload (RGB8 and
    12-bit gray, 32x32 and 64x64 code-blocks, 5 and 6 levels) at 1/2 and
    1/4 in one plan, against the oracle."""
    rng = np.random.default_rng(5121024)
    cases = []
    for k, (size, cb, levels) in enumerate([(512, 5, 5), (512, 6, 6), (1024, 5, 6),
                                            (1024, 6, 5)]):
        for den in (2, 4):
            if k % 2 == 0:
                img = synth.batch_tile(synth.SEED_BASE + 3, k, k % 3, size)
                cs = S.encode_j2k_ht(img, bit_depth=8, levels=levels, cb_width_log2=cb,
                                     cb_height_log2=cb, reversible=1, use_mct=1)
                fmt = S.PixelFormat.RGB8
            else:
                img = synth.spots_gray12(rng, size, size)
                cs = S.encode_j2k_ht(img, bit_depth=12, levels=levels, cb_width_log2=cb,
                                     cb_height_log2=cb, reversible=1, use_mct=0)
                fmt = S.PixelFormat.GRAY16
            cases.append((cs, fmt, S.Downscale(den)))
    plan = S.TilePlan(session, [c[0] for c in cases], [c[1] for c in cases],
                      [c[2] for c in cases])
    plan.decode(S.MODE_GRAPH)
    session.synchronize()
    for j, ((cs, fmt, den), surf) in enumerate(zip(cases, plan.surfaces)):
        assert np.array_equal(surf.to_numpy(), O.decode(cs, int(fmt), int(den))), j
    plan.close()


def _packet_cases():
    """HT codestreams over the packet-header features: five progressions,
    precincts, SOP/EPH, tiles, origins, refinement passes, 9/7, 12-bit gray,
    elongated code-blocks, and the OpenHTJ2K fixtures."""
    rng = np.random.default_rng(4242)
    cases = [dict(), dict(progression=1), dict(progression=2, precinct_log2=5),
             dict(progression=3, precinct_log2=6), dict(progression=4, precinct_log2=5),
             dict(sop=1, eph=1, progression=2, precinct_log2=5), dict(tile_width=64, tile_height=48),
             dict(ht_refine=2), dict(ht_refine=1, cb_width_log2=5, cb_height_log2=4),
             dict(reversible=0, base_step=1 / 128), dict(levels=7, tile_width=64, tile_height=64),
             dict(cb_width_log2=4, cb_height_log2=6, levels=3)]
    out = []
    for k, kw in enumerate(cases):
        img = synth.tissue_rgb(rng, 150 + 7 * k, 100 + 3 * k)
        out.append((S.encode_j2k_ht(img, **kw), S.PixelFormat.RGB8))
    g = synth.spots_gray12(rng, 130, 90)
    out.append((S.encode_j2k_ht(g, bit_depth=12, use_mct=0, progression=2, precinct_log2=5),
                S.PixelFormat.GRAY16))
    here = os.path.dirname(os.path.abspath(__file__))
    for name in ("openhtj2k_ds0_ht_09_b11.j2k", "openhtj2k_ds0_ht_12_b11.j2k"):
        with open(os.path.join(here, "golden", name), "rb") as f:
            out.append((f.read(), S.PixelFormat.GRAY8))
    return out


def test_device_packet_headers(session, oracle, monkeypatch):
    """Packet headers parsed on the device (SGN_DEVICE_PACKETS=1) give the
    oracle's output on every case, in one batch per format, graph and
    per-tile modes, full and half resolution; the plan's code-block order is
    fixed, and its lane variants still autotune."""
    monkeypatch.setenv("SGN_DEVICE_PACKETS", "1")
    cases = _packet_cases()
    for fmt in (S.PixelFormat.RGB8, S.PixelFormat.GRAY16, S.PixelFormat.GRAY8):
        streams = [cs for cs, f in cases if f == fmt]
        for den in (1, 2):
            plan = S.TilePlan(session, streams, fmt, S.Downscale(den))
            for mode in (S.MODE_GRAPH, S.MODE_PLAIN):
                for s in plan.surfaces:
                    s.fill(0)
                plan.decode(mode)
                session.synchronize()
                for k, cs in enumerate(streams):
                    want = O.decode(cs, int(fmt), den)
                    got = plan.surfaces[k].to_numpy()
                    assert np.array_equal(got.reshape(want.shape), want), (fmt, den, mode, k)
            if fmt == S.PixelFormat.RGB8 and den == 1:
                plan.autotune(reps=1)
                with pytest.raises(S.UnsupportedError):
                    plan.autotune_order(reps=1)
            plan.close()


def test_device_packet_headers_pipelined_and_region(session, oracle, small_chunks, monkeypatch):
    """The pipelined host-buffer decode and a region decode with the packet
    headers parsed on the device."""
    monkeypatch.setenv("SGN_DEVICE_PACKETS", "1")
    streams, pix, _cls = synth.lossless_batch_with_pixels(40, size=128)
    surfs = S._alloc_surfaces(session, [(128, 128)] * 40, [S.PixelFormat.RGB8] * 40)
    keep = [np.frombuffer(s, np.uint8) for s in streams]
    rc, _ = _decode_tiles(session, _jobs([k.ctypes.data for k in keep],
                                         [len(s) for s in streams], surfs), 40)
    assert rc == 0, _lib.last_error()
    for i, (s, img) in enumerate(zip(surfs, pix)):
        assert np.array_equal(s.to_numpy(), img), i
    img = synth.tissue_rgb(np.random.default_rng(77), 300, 200)
    cs = S.encode_j2k_ht(img, tile_width=128, tile_height=96, progression=2, precinct_log2=5)
    got = S.J2kDecoder(cs, session).decode_region_to_numpy(S.PixelFormat.RGB8, (70, 33, 150, 90))
    assert np.array_equal(got, img[33:123, 70:220])
