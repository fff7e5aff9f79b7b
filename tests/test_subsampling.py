"""Sub-sampled components (XRsiz / YRsiz > 1) on the CPU.

Pins: OpenJPEG 2.5.4 decodes the package's codestreams whose components
share one sub-sampling to the decimated source (it upsamples to the full
grid; every second sample is the source's), in every progression order,
which checks the B-12 component rectangles and the position-driven packet
order with XRsiz / YRsiz steps.  Components on different grids are pinned
by the closed form of the reference's store (decode.rs:1858-1898): sample
u of a component of sub-sampling dx covers grid positions [u dx, u dx +
dx), so a lossless decode equals the source with each component's samples
replicated; OpenJPEG converts such images from YCbCr and cannot serve."""
import io

import numpy as np
import pytest

import signinum_amd as S
from signinum_amd import synth
from tests import oracle_ref as O

PIL = pytest.importorskip("PIL.Image")


def replicated(img, sub, xo=0, yo=0):
    h, w, _ = img.shape
    out = np.zeros_like(img)
    for c, (dx, dy) in enumerate(sub):
        u = (np.arange(w) + xo) // dx * dx - xo
        v = (np.arange(h) + yo) // dy * dy - yo
        out[..., c] = img[v[:, None], u[None, :], c]
    return out


MIXED = [[(1, 1), (2, 2), (2, 2)], [(1, 1), (2, 1), (2, 1)], [(1, 1), (3, 2), (1, 4)],
         [(2, 1), (1, 2), (1, 1)]]


@pytest.mark.parametrize("sub", MIXED)
@pytest.mark.parametrize("classic", [0, 1])
def test_mixed_sampling_is_the_replicated_source(oracle, sub, classic):
    img = synth.tissue_rgb(np.random.default_rng(len(sub[1]) + 3), 150, 100)
    cs = S.encode_j2k_ht(img, subsampling=sub, use_mct=0, block_coder=classic)
    assert S.J2kDecoder.inspect(cs)["width"] == 150
    assert np.array_equal(O.decode(cs, 1), replicated(img, sub))


@pytest.mark.parametrize("prog", [0, 1, 2, 3, 4])
def test_shared_sampling_read_by_openjpeg(oracle, prog):
    img = synth.texture_rgb(np.random.default_rng(prog), 150, 100)
    cs = S.encode_j2k_ht(img, subsampling=[(2, 2)] * 3, progression=prog, precinct_log2=4,
                         tile_width=64, tile_height=48, levels=3)
    info = S.J2kDecoder.inspect(cs)
    assert (info["width"], info["height"]) == (75, 50)
    pil = np.asarray(PIL.open(io.BytesIO(cs)))
    assert np.array_equal(pil[::2, ::2], img[::2, ::2])
    # the output grid is the components' own: the decimated source
    assert np.array_equal(O.decode(cs, 1), img[::2, ::2])


@pytest.mark.parametrize("prog", [2, 3, 4])
def test_mixed_sampling_position_orders_host_equals_oracle(oracle, prog):
    img = synth.texture_rgb(np.random.default_rng(10 + prog), 130, 90)
    sub = [(1, 1), (2, 2), (2, 1)]
    cs = S.encode_j2k_ht(img, subsampling=sub, progression=prog, precinct_log2=4,
                         tile_width=64, tile_height=64, image_x0=6, image_y0=4, tile_x0=0,
                         tile_y0=0, levels=3)
    assert np.array_equal(O.decode(cs, 1), replicated(img, sub, 6, 4))
    ours = [(j.width, j.height, j.cleanup_length, j.number_of_coding_passes)
            for j in S.code_block_jobs(cs)]
    ref = [(c.x1 - c.x0, c.y1 - c.y0, c.lcup if c.passes else 0, c.passes)
           for c in O.code_blocks(cs) if c.x1 > c.x0 and c.y1 > c.y0]
    assert ours == ref


def test_subsampled_refusals(oracle):
    img = synth.tissue_rgb(np.random.default_rng(4), 96, 64)
    cs = S.encode_j2k_ht(img, subsampling=[(1, 1), (2, 2), (2, 2)], use_mct=0)
    # reduced resolution of sub-sampled components
    with pytest.raises(S.UnsupportedError):
        S.scaled_dims(cs, S.Downscale.HALF)
    with pytest.raises(AssertionError):
        O.scaled_dims(cs, 2)
    # an image origin off a component's grid: the reference leaves the first
    # column without that component's samples (zeros); refused here
    # (the device refusal is in the GPU tests)
    odd = S.encode_j2k_ht(img, subsampling=[(1, 1), (2, 2), (2, 2)], use_mct=0, image_x0=1)
    assert not O.decode(odd, 1)[:, 0, 1:].any()
