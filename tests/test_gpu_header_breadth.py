"""Device decodes of codestreams with packed packet headers (PPM, PPT),
non-zero image and tile origins and signed components, against the oracle
and the source pixels, in one batch plan and through the per-tile
decoder."""
import numpy as np
import pytest

import signinum_amd as S
from signinum_amd import synth
from tests import oracle_ref as O
from tests.test_header_breadth import CASES, _set_signed

pytestmark = pytest.mark.gpu


def test_header_breadth_batch(session, oracle):
    imgs, streams = [], []
    for k, kw in enumerate(CASES):
        img = synth.tissue_rgb(np.random.default_rng(100 + k), 150, 90)
        imgs.append(img)
        streams.append(S.encode_j2k_ht(img, **kw))
    img = synth.texture_rgb(np.random.default_rng(7), 96, 64)
    imgs.append(img)
    streams.append(_set_signed(S.encode_j2k_lossless(img)))
    plan = S.TilePlan(session, streams, S.PixelFormat.RGB8)
    for variant in (None, "wave-coop"):
        if variant:
            plan.lane_variant = S.TilePlan.lane_variants().index(variant)
        for s in plan.surfaces:
            s.fill(0)
        plan.decode(S.MODE_GRAPH)
        session.synchronize()
        for k, (cs, src) in enumerate(zip(streams, imgs)):
            out = plan.surfaces[k].to_numpy()
            assert np.array_equal(out, src), (k, variant)
            assert np.array_equal(out, O.decode(cs, 1)), (k, variant)
    plan.close()


def test_origin_gray12_multitile(session, oracle):
    img = synth.spots_gray12(np.random.default_rng(12), 200, 120)
    cs = S.encode_j2k_ht(img, bit_depth=12, use_mct=0, levels=4, image_x0=37, image_y0=21,
                         tile_x0=30, tile_y0=20, tile_width=64, tile_height=64, packed_headers=2)
    out = S.J2kDecoder(cs, session).decode_to_numpy(S.PixelFormat.GRAY16)
    assert np.array_equal(out[..., 0], img)


def test_reduced_resolution_with_origin_refused_on_device(session):
    img = synth.tissue_rgb(np.random.default_rng(3), 120, 80)
    cs = S.encode_j2k_ht(img, image_x0=8, image_y0=8)
    d = S.J2kDecoder(cs, session)
    assert np.array_equal(d.decode_to_numpy(S.PixelFormat.RGB8), img)
    with pytest.raises(S.UnsupportedError):
        d.decode_to_numpy(S.PixelFormat.RGB8, S.Downscale.HALF)
