"""Classic (EBCOT / MQ) code-blocks on the CPU: the oracle's restatement of
bitplane.rs and arithmetic_decoder.rs against OpenJPEG 2.5.4 (through
Pillow), which encodes and decodes these codestreams independently.

Pins: every lossless codestream (single- and multi-layer, 32x32, 64x64 and
elongated code-blocks, tiles, precincts, progressions, 8- and 16-bit) decodes
to OpenJPEG's pixels and to the source exactly; 9/7 within 1 LSB of
OpenJPEG (the two round the f32 lifting differently).  Codestreams whose last
layer leaves bit-planes out are decoded without OpenJPEG's mid-point
reconstruction (the reference's Coefficient::get, bitplane.rs:346-352), so
those are compared with the source within the loss the layers allow."""
import io

import numpy as np
import pytest

import signinum_amd as S
from signinum_amd import synth
from tests import oracle_ref as O

PIL = pytest.importorskip("PIL.Image")

LOSSLESS = [
    dict(),
    dict(codeblock_size=(32, 32), num_resolutions=3),
    dict(codeblock_size=(256, 16)),
    dict(codeblock_size=(16, 256)),
    dict(codeblock_size=(1024, 4), num_resolutions=2),
    dict(tile_size=(64, 64)),
    dict(progression="RPCL", precinct_size=(64, 64)),
    dict(progression="PCRL", precinct_size=(32, 32), codeblock_size=(16, 16)),
    dict(quality_layers=[40, 10, 0]),
    dict(quality_layers=[80, 40, 20, 0], progression="LRCP"),
]


def pil_encode(img, **kw):
    b = io.BytesIO()
    PIL.fromarray(img).save(b, "JPEG2000", no_jp2=True, **kw)
    return b.getvalue()


def pil_decode(cs):
    return np.asarray(PIL.open(io.BytesIO(cs)))


@pytest.mark.parametrize("kw", LOSSLESS)
def test_lossless_classic_equals_openjpeg_and_source(oracle, kw):
    img = synth.tissue_rgb(np.random.default_rng(len(kw) + 3), 150, 100)
    cs = pil_encode(img, **kw)
    assert not S.J2kDecoder.inspect(cs)["high_throughput"]
    out = O.decode(cs, 1)
    assert np.array_equal(out, pil_decode(cs))
    assert np.array_equal(out, img)


def test_sixteen_bit_gray_classic(oracle):
    g = synth.spots_gray12(np.random.default_rng(8), 120, 90).astype(np.uint16) * 13
    b = io.BytesIO()
    PIL.fromarray(g).save(b, "JPEG2000", no_jp2=True)
    cs = b.getvalue()
    out = O.decode(cs, int(S.PixelFormat.GRAY16)).view("<u2").reshape(90, 120)
    assert np.array_equal(out, g)


@pytest.mark.parametrize("kw", [dict(irreversible=True),
                                dict(irreversible=True, quality_layers=[30, 10, 0])])
def test_irreversible_classic_within_one_lsb_of_openjpeg(oracle, kw):
    img = synth.tissue_rgb(np.random.default_rng(21), 150, 100)
    cs = pil_encode(img, **kw)
    d = np.abs(O.decode(cs, 1).astype(int) - pil_decode(cs).astype(int)).max()
    assert d <= 1  # tolerance: 1 LSB of 8-bit output


def test_truncated_layers_decode_without_midpoint(oracle):
    """Layers at rates 40:1 .. 10:1 drop the last bit-planes of many blocks;
    the decode stays within the loss those rates allow."""
    img = synth.tissue_rgb(np.random.default_rng(22), 150, 100)
    cs = pil_encode(img, quality_layers=[40, 20, 10])
    out = O.decode(cs, 1).astype(int)
    ref = pil_decode(cs).astype(int)
    assert np.abs(out - img).max() <= 8
    # OpenJPEG adds half a step to every truncated coefficient; the reference
    # does not, so the two differ, but by little
    assert 0 < np.abs(out - ref).max() <= 4


def test_classic_has_no_ht_block_jobs(oracle):
    cs = pil_encode(synth.tissue_rgb(np.random.default_rng(5), 64, 64))
    with pytest.raises(S.UnsupportedError):
        S.code_block_jobs(cs)


STYLES = [0, 0x01, 0x02, 0x04, 0x08, 0x20, 0x05, 0x0D, 0x2F]


@pytest.mark.parametrize("style", STYLES)
def test_classic_encoder_styles_read_by_openjpeg_and_oracle(oracle, style):
    """The package's classic encoder with every code-block style: OpenJPEG
    decodes it to the source (pins the encoder's MQ coder, raw segments,
    terminations and segment lengths), and so does the oracle (pins its
    bypass, reset, termination, causal and segmentation-symbol handling)."""
    img = synth.tissue_rgb(np.random.default_rng(40 + style), 150, 100)
    cs = S.encode_j2k_classic(img, cb_style=style)
    assert np.array_equal(pil_decode(cs), img)
    assert np.array_equal(O.decode(cs, 1), img)


@pytest.mark.parametrize("style", [0, 0x05, 0x2F])
def test_classic_encoder_irreversible_and_gray12(oracle, style):
    img = synth.tissue_rgb(np.random.default_rng(50 + style), 130, 90)
    cs = S.encode_j2k_classic(img, cb_style=style, reversible=0, base_step=1 / 256)
    out = O.decode(cs, 1).astype(int)
    assert np.abs(out - img).max() <= 3
    # OpenJPEG reconstructs at the mid-point, the reference does not
    assert np.abs(out - pil_decode(cs).astype(int)).max() <= 2
    g = synth.spots_gray12(np.random.default_rng(60 + style), 100, 70)
    cs = S.encode_j2k_classic(g, bit_depth=12, use_mct=0, cb_style=style, cb_width_log2=5,
                              cb_height_log2=5, levels=4)
    assert np.array_equal(O.decode(cs, int(S.PixelFormat.GRAY16)).view("<u2").reshape(70, 100), g)
