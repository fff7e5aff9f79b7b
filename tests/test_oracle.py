"""The CPU oracle against golden vectors, closed forms and OpenJPEG.

Pins (each one fails for a plausible mistake in the restatement):
- OpenHTJ2K conformance codestreams with OpenJPH-decoded pixels
  (crates/signinum-j2k-native/fixtures/htj2k, reference test
  crates/signinum-j2k-native/src/lib.rs:3463-3545): exact pixels, block
  counts, refinement-job counts, maximum passes, and the number of samples
  that change when refinement bytes are zeroed;
- lossless round trips: the decode of a reversible codestream equals the
  source pixels (crates/signinum-j2k/tests/encode_lossless.rs:55-427);
- OpenJPEG 2.5.4 (bundled with Pillow) decoding the same codestreams;
- structure of the VLC tables (complete prefix codes per context).
"""
import io

import numpy as np
import pytest

import signinum_amd as S
from signinum_amd import synth
from tests import oracle_ref as O

GOLDEN = __import__("os").path.join(__import__("os").path.dirname(__file__), "golden")


def _fixture(name):
    cs = open(f"{GOLDEN}/{name}.j2k", "rb").read()
    exp = np.frombuffer(open(f"{GOLDEN}/{name}.gray", "rb").read(), np.uint8)
    return cs, exp


@pytest.mark.parametrize("name,w,h,blocks,ref_blocks,max_passes,zero_diffs", [
    ("openhtj2k_ds0_ht_12_b11", 3, 5, 8, 2, 3, 4),
    ("openhtj2k_ds0_ht_09_b11", 17, 37, 14, 14, 3, 629),
])
def test_openhtj2k_fixtures(oracle, name, w, h, blocks, ref_blocks, max_passes, zero_diffs):
    cs, exp = _fixture(name)
    out = O.decode(cs, 0)
    assert out.shape == (h, w, 1)
    assert np.array_equal(out.reshape(-1), exp)
    assert O.last_stats() == (blocks, ref_blocks, max_passes)
    planes = O.decode_planes(cs)
    O.set_zero_refinement(True)
    try:
        zeroed = O.decode_planes(cs)
    finally:
        O.set_zero_refinement(False)
    assert int((np.abs(planes - zeroed) > np.finfo(np.float32).eps).sum()) == zero_diffs


def test_vlc_tables_are_complete_prefix_codes(oracle):
    v0, v1, u0, u1 = O.tables()
    for table in (v0, v1):
        for c_q in range(8):
            entries = table[c_q * 128:(c_q + 1) * 128]
            lens = entries & 7
            assert (lens > 0).all()
            # every 7-bit window decodes; codewords cover the space exactly once
            codes = {}
            for i, e in enumerate(entries):
                ln = int(e & 7)
                codes.setdefault((i & ((1 << ln) - 1), ln), e)
            assert sum(2.0 ** -ln for (_c, ln) in codes) == pytest.approx(1.0)
    # U-VLC: a lone residual with prefix "1" is u=1 (no suffix), "000" needs 5 bits
    assert (u1[64 + 1] >> 10) & 7 == 1 and (u1[64 + 1] & 7) == 1
    assert (u1[64 + 0] >> 10) & 7 == 5 and ((u1[64 + 0] >> 3) & 0xF) == 5


def test_coefficient_to_i32_alignment(oracle):
    # ht_block_decode.rs:1695-1700
    aligned = 3 << (31 - 5)
    assert O.coefficient_to_i32(aligned, 5) == 3
    assert O.coefficient_to_i32(0x80000000 | aligned, 5) == -3


def _pillow(cs):
    from PIL import Image
    return np.array(Image.open(io.BytesIO(cs)))


@pytest.mark.parametrize("shape,levels,prog,cb", [
    ((64, 64), 5, 0, 6), ((37, 53), 3, 1, 5), ((129, 200), 5, 2, 6), ((5, 3), 1, 3, 6),
    ((256, 256), 5, 4, 5), ((1, 7), 2, 0, 4), ((100, 64), 0, 0, 6), ((80, 96), 4, 0, 4),
])
def test_lossless_roundtrip_rgb(oracle, shape, levels, prog, cb):
    rng = np.random.default_rng(sum(shape) + levels)
    img = synth.gradient_noise_rgb(rng, shape[1], shape[0])
    cs = S.encode_j2k_ht(img, levels=levels, progression=prog, cb_width_log2=cb,
                         cb_height_log2=cb)
    assert np.array_equal(O.decode(cs, 1), img)
    assert np.array_equal(_pillow(cs), img)


@pytest.mark.parametrize("kw", [dict(precinct_log2=5), dict(sop=1, eph=1),
                                dict(tile_width=96, tile_height=80),
                                dict(cb_width_log2=7, cb_height_log2=5)])
def test_lossless_roundtrip_options(oracle, kw):
    rng = np.random.default_rng(7)
    img = synth.tissue_rgb(rng, 200, 180)
    cs = S.encode_j2k_ht(img, **kw)
    assert np.array_equal(O.decode(cs, 1), img)
    assert np.array_equal(_pillow(cs), img)


def test_gray12_roundtrip(oracle):
    rng = np.random.default_rng(3)
    img = synth.spots_gray12(rng, 90, 70)
    cs = S.encode_j2k_ht(img, bit_depth=12, use_mct=0, cb_width_log2=5, cb_height_log2=5)
    out = O.decode(cs, 3)
    assert np.array_equal(out[..., 0], img)
    assert np.array_equal(_pillow(cs).astype(np.int64) >> 4, img.astype(np.int64)) or \
        np.array_equal(_pillow(cs), img)


@pytest.mark.parametrize("den", [2, 4, 8])
def test_reduced_resolution_matches_openjpeg(oracle, den):
    from PIL import Image
    rng = np.random.default_rng(den)
    img = synth.gradient_noise_rgb(rng, 160, 120)
    cs = S.encode_j2k_lossless(img)
    out = O.decode(cs, 1, den)
    pil = Image.open(io.BytesIO(cs))
    pil.reduce = {2: 1, 4: 2, 8: 3}[den]
    ref = np.array(pil)
    assert out.shape == ref.shape
    assert np.array_equal(out, ref)


def test_irreversible_close_to_openjpeg_and_f64(oracle):
    rng = np.random.default_rng(11)
    img = synth.tissue_rgb(rng, 128, 96)
    cs = S.encode_j2k_ht(img, reversible=0, base_step=1.0 / 64)
    out = O.decode(cs, 1).astype(np.int64)
    # OpenJPEG reconstructs quantised values at the bin midpoint; the
    # reference truncates (ht_block_decode.rs:225-234), so the two differ by
    # a fraction of a step, not more.
    diff = np.abs(out - _pillow(cs).astype(np.int64))
    assert diff.max() <= 6 and diff.mean() < 1.5
    p32 = O.decode_planes(cs)
    p64 = O.decode_planes(cs, use_f64=True)
    assert np.abs(p32 - p64).max() < 0.05
    # lossy but close to the source at this step size
    assert np.abs(out - img.astype(np.int64)).max() < 40


def test_batch_first_failing_index(oracle):
    rng = np.random.default_rng(5)
    streams = [S.encode_j2k_lossless(synth.texture_rgb(rng, 32, 32)) for _ in range(5)]
    bad = bytearray(streams[3])
    bad[:] = bad[:60]  # truncated main header
    streams[3] = bytes(bad)
    rc, idx, _outs = O.decode_batch(streams, 1, workers=2)
    assert rc != 0 and idx == 3


def _gray8_closed_form(v, bit_depth):
    """>8-bit sample to Gray8 as the reference's slow path computes it
    (crates/signinum-j2k-native/src/lib.rs:2282-2293): f32 division by
    2^bd - 1, times 255 in f32, rounded half away from zero, saturating."""
    x = (v.astype(np.float32) / np.float32((1 << bit_depth) - 1)) * np.float32(255.0)
    r = np.where(x >= 0, np.floor(x + np.float32(0.5)), np.ceil(x - np.float32(0.5)))
    return np.clip(r, 0, 255).astype(np.uint8)


@pytest.mark.parametrize("bit_depth", [10, 12, 16])
def test_gray8_from_high_bit_depth_closed_form(oracle, bit_depth):
    """Gray8 output of a >8-bit lossless codestream equals the closed form
    applied to its Gray16 output, which itself equals the source samples."""
    rng = np.random.default_rng(40 + bit_depth)
    img = rng.integers(0, 1 << bit_depth, size=(45, 61), dtype=np.int64).astype(np.uint16)
    img[0, :4] = [0, (1 << bit_depth) - 1, 1 << (bit_depth - 1), 8]
    cs = S.encode_j2k_ht(img, bit_depth=bit_depth, use_mct=0, levels=3)
    g16 = O.decode(cs, 3)[..., 0]
    assert np.array_equal(g16, img)
    assert np.array_equal(O.decode(cs, 0)[..., 0], _gray8_closed_form(img, bit_depth))


def test_irreversible_multitile_odd_origins_close_to_openjpeg(oracle):
    """9/7 + ICT over 100x70 tiles of a 300x200 image, 4 levels: the inner
    tiles' coarse levels start at odd coordinates, where the first lifting
    step and the K / 1/K scaling swap roles between even and odd samples.
    OpenJPEG is independent of the oracle's geometry code."""
    rng = np.random.default_rng(22)
    img = synth.tissue_rgb(rng, 300, 200)
    cs = S.encode_j2k_ht(img, tile_width=100, tile_height=70, levels=4, reversible=0,
                         base_step=1.0 / 64)
    out = O.decode(cs, 1).astype(np.int64)
    diff = np.abs(out - _pillow(cs).astype(np.int64))
    assert diff.max() <= 6 and diff.mean() < 1.5
    p32 = O.decode_planes(cs)
    p64 = O.decode_planes(cs, use_f64=True)
    assert np.abs(p32 - p64).max() < 0.05
    assert np.abs(out - img.astype(np.int64)).max() < 40


def test_lone_odd_origin_sample_round_trip(oracle):
    """A 193-pixel-wide image in 64-pixel tiles with 7 levels: the last tile
    column is one sample wide at an odd coordinate (x0 = 3) on the first
    synthesis level, where the 5/3 synthesis halves the sample
    (idwt.rs:726-731) that the encoder doubled.  The oracle and OpenJPEG
    both give the source."""
    img = synth.tissue_rgb(np.random.default_rng(193), 193, 70)
    for kw in (dict(), dict(use_mct=0)):
        cs = S.encode_j2k_ht(img, tile_width=64, tile_height=64, levels=7, **kw)
        assert np.array_equal(O.decode(cs, 1), img)
        assert np.array_equal(_pillow(cs), img)
