"""Sub-sampled components on the device against the oracle: components on
different grids (per tile-component store jobs), a shared sub-sampling
(the fused final stage on the components' own grid), HT and classic,
multi-tile, regions (decoded whole and cropped, as the reference does),
RGBA8 / RGB16 outputs and the refusals."""
import numpy as np
import pytest

import signinum_amd as S
from signinum_amd import synth
from tests import oracle_ref as O
from tests.test_subsampling import MIXED, replicated

pytestmark = pytest.mark.gpu


def test_mixed_sampling_batch(session, oracle):
    rng = np.random.default_rng(81)
    streams, want = [], []
    for k, sub in enumerate(MIXED * 2):
        img = synth.tissue_rgb(rng, 120 + 8 * k, 88)
        kw = dict(subsampling=sub, use_mct=0, block_coder=k % 2)
        if k >= len(MIXED):
            kw.update(tile_width=64, tile_height=48, progression=k % 5, precinct_log2=5,
                      image_x0=12, image_y0=4)
        cs = S.encode_j2k_ht(img, **kw)
        streams.append(cs)
        ref = O.decode(cs, 1)
        want.append(ref)
        assert np.array_equal(ref, replicated(img, sub, kw.get("image_x0", 0),
                                              kw.get("image_y0", 0)))
    plan = S.TilePlan(session, streams, S.PixelFormat.RGB8)
    for mode in (S.MODE_GRAPH, S.MODE_PLAIN):
        for s in plan.surfaces:
            s.fill(7)
        plan.decode(mode)
        session.synchronize()
        for k, s in enumerate(plan.surfaces):
            assert np.array_equal(s.to_numpy(), want[k]), (mode, k)
    plan.close()


def test_shared_sampling_and_formats(session, oracle):
    img = synth.texture_rgb(np.random.default_rng(82), 150, 100)
    cs = S.encode_j2k_ht(img, subsampling=[(2, 2)] * 3, tile_width=64, tile_height=64)
    d = S.J2kDecoder(cs, session)
    assert np.array_equal(d.decode_to_numpy(S.PixelFormat.RGB8), img[::2, ::2])
    mixed = S.encode_j2k_ht(img, subsampling=[(1, 1), (2, 2), (2, 2)], use_mct=0)
    dm = S.J2kDecoder(mixed, session)
    for fmt in (S.PixelFormat.RGBA8, S.PixelFormat.RGB16):
        assert np.array_equal(dm.decode_to_numpy(fmt), O.decode(mixed, int(fmt)).reshape(
            100, 150, fmt.channels)), fmt
    # a region: decoded whole, written cropped
    roi = (17, 9, 60, 41)
    got = dm.decode_region_to_numpy(S.PixelFormat.RGB8, roi)
    assert np.array_equal(got, O.decode(mixed, 1)[9:50, 17:77])


def test_subsampled_refusals_on_device(session):
    img = synth.tissue_rgb(np.random.default_rng(83), 96, 64)
    odd = S.encode_j2k_ht(img, subsampling=[(1, 1), (2, 2), (2, 2)], use_mct=0, image_x0=1)
    with pytest.raises(S.UnsupportedError):
        S.J2kDecoder(odd, session).decode_to_numpy(S.PixelFormat.RGB8)
    cs = S.encode_j2k_ht(img, subsampling=[(1, 1), (2, 2), (2, 2)], use_mct=0)
    with pytest.raises(S.UnsupportedError):
        S.J2kDecoder(cs, session).decode_to_numpy(S.PixelFormat.RGB8, S.Downscale.HALF)
