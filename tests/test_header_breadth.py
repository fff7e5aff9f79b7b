"""Codestream features beyond the common case (CPU): packed packet headers
(PPM in the main header, PPT in tile-part headers), non-zero image and
tile origins, signed components.

Pins: OpenJPEG 2.5.4 (through Pillow) decodes every codestream written
here to its source pixels, which checks the encoder's PPM/PPT/origin
syntax independently; the oracle, restating the reference's parsers
(codestream.rs:761-780 PPM with its 16-bit Nppm read, tile.rs:251-309 PPT
and the per-tile-part header readers, segment.rs:36-133), decodes them to
the same pixels; the host parser lists the same code-blocks at the same
codestream offsets as the oracle."""
import io
import struct

import numpy as np
import pytest

import signinum_amd as S
from signinum_amd import synth
from tests import oracle_ref as O

CASES = [
    dict(packed_headers=1),
    dict(packed_headers=2),
    dict(packed_headers=1, tile_width=64, tile_height=48, sop=1, eph=1),
    dict(packed_headers=2, tile_width=64, tile_height=48, progression=2, precinct_log2=5),
    dict(packed_headers=2, progression=4, precinct_log2=4, levels=3, sop=1),
    dict(image_x0=13, image_y0=7),
    dict(image_x0=100, image_y0=37, tile_x0=60, tile_y0=30, tile_width=64, tile_height=48),
    dict(image_x0=5, image_y0=3, tile_x0=1, tile_y0=2, tile_width=50, tile_height=40,
         packed_headers=2),
    dict(image_x0=64, image_y0=64, tile_x0=64, tile_y0=64, tile_width=64, tile_height=64,
         packed_headers=1, progression=1),
]


def _pillow(cs):
    from PIL import Image
    im = Image.open(io.BytesIO(cs))
    return np.asarray(im.convert("RGB") if im.mode != "L" else im)


def _img(seed=5, w=150, h=90):
    return synth.tissue_rgb(np.random.default_rng(seed), w, h)


@pytest.mark.parametrize("kw", CASES)
def test_openjpeg_and_oracle_decode_to_source(oracle, kw):
    img = _img(len(kw))
    cs = S.encode_j2k_ht(img, **kw)
    assert np.array_equal(_pillow(cs), img)
    assert np.array_equal(O.decode(cs, 1), img)
    info = S.J2kDecoder.inspect(cs)
    assert (info["width"], info["height"]) == (150, 90)


@pytest.mark.parametrize("kw", CASES)
def test_host_listing_matches_oracle(oracle, kw):
    img = _img(len(kw) + 1)
    cs = S.encode_j2k_ht(img, **kw)
    ours = [(j.width, j.height, j.cleanup_length, j.refinement_length,
             j.number_of_coding_passes, j.data_offset if j.number_of_coding_passes else None)
            for j in S.code_block_jobs(cs)]
    ref = [(c.x1 - c.x0, c.y1 - c.y0, c.lcup if c.passes else 0, c.lref if c.passes else 0,
            c.passes, c.offset if c.passes else None)
           for c in O.code_blocks(cs) if c.x1 > c.x0 and c.y1 > c.y0]
    assert ours == ref
    assert sum(1 for r in ref if r[4]) > 10


def test_ppt_split_over_segments(oracle):
    """A tile whose packet headers exceed one PPT segment (65532 bytes):
    several segments, Zppt 0, 1, ..., each holding whole packet headers."""
    img = synth.texture_rgb(np.random.default_rng(77), 512, 512)[..., 0].copy()
    cs = S.encode_j2k_ht(img, use_mct=0, levels=3, cb_width_log2=2, cb_height_log2=2,
                         packed_headers=2)
    assert cs.count(b"\xff\x61") >= 2
    assert np.array_equal(_pillow(cs), img)
    assert np.array_equal(O.decode(cs, 0)[..., 0], img)


def _set_signed(cs):
    """The same codestream with every component flagged signed (Ssiz bit 7)."""
    b = bytearray(cs)
    assert b[2:4] == b"\xff\x51"
    ncomp = struct.unpack(">H", b[40:42])[0]
    for c in range(ncomp):
        b[42 + 3 * c] |= 0x80
    return bytes(b)


def test_signed_components_decode_as_unsigned(oracle):
    """codestream.rs:647-649 ignores the signed flag; the decode is that of
    the unsigned codestream."""
    img = _img(9)
    cs = S.encode_j2k_lossless(img)
    sc = _set_signed(cs)
    assert np.array_equal(O.decode(sc, 1), img)
    a, b = S.J2kDecoder.inspect(sc), S.J2kDecoder.inspect(cs)
    assert a == b


def test_reduced_resolution_with_image_origin_is_refused(oracle):
    img = _img(11)
    cs = S.encode_j2k_ht(img, image_x0=13, image_y0=7)
    assert S.scaled_dims(cs, S.Downscale.NONE) == (150, 90)
    with pytest.raises(S.UnsupportedError):
        S.scaled_dims(cs, S.Downscale.HALF)
    with pytest.raises(AssertionError):
        O.scaled_dims(cs, 2)
    # without an origin the tile grid offset alone is not allowed (B-4)
    with pytest.raises(S.J2kError):
        S.encode_j2k_ht(img, tile_x0=3, tile_width=64, tile_height=64)


def test_ppm_and_ppt_together_are_malformed(oracle):
    """tile.rs:251-256: a PPT marker in a codestream with PPM."""
    img = _img(13)
    cs = bytearray(S.encode_j2k_ht(img, packed_headers=1))
    sot = cs.index(b"\xff\x90")
    sod = cs.index(b"\xff\x93", sot)
    ppt = b"\xff\x61\x00\x04\x00\x00"
    psot = struct.unpack(">I", cs[sot + 6:sot + 10])[0]
    cs[sot + 6:sot + 10] = struct.pack(">I", psot + len(ppt))
    bad = bytes(cs[:sod] + ppt + cs[sod:])
    with pytest.raises(S.MalformedError):
        S.code_block_jobs(bad)
    with pytest.raises(AssertionError):
        O.code_blocks(bad)
