"""Region and region+scaled decode (row f2): the output geometry on the CPU,
and on the device the reference's own contract
(crates/signinum-j2k/tests/decode.rs:374-611): a region decode equals the
crop of the full decode, and a region+scaled decode equals the crop of the
scaled decode at roi.scaled_covering(scale).  The full decodes come from the
oracle."""
import numpy as np
import pytest

import signinum_amd as S
from signinum_amd import synth
from tests import oracle_ref as O


def _covering(roi, den):
    x, y, w, h = roi
    x0, y0 = x // den, y // den
    return x0, y0, -(-(x + w) // den) - x0, -(-(y + h) // den) - y0


@pytest.mark.parametrize("roi,den", [((0, 0, 150, 90), 1), ((13, 7, 50, 33), 1),
                                     ((1, 0, 2, 3), 2), ((13, 7, 50, 33), 4),
                                     ((149, 89, 1, 1), 8), ((20, 10, 0, 5), 2)])
def test_region_dims_are_the_scaled_covering(roi, den):
    cs = S.encode_j2k_lossless(synth.tissue_rgb(np.random.default_rng(1), 150, 90))
    assert S.region_dims(cs, roi, S.Downscale(den)) == _covering(roi, den)


def test_region_outside_the_image_is_refused():
    cs = S.encode_j2k_lossless(synth.tissue_rgb(np.random.default_rng(2), 150, 90))
    for roi in ((140, 0, 11, 5), (0, 80, 10, 11), (0xFFFFFFF0, 0, 0x20, 1)):
        with pytest.raises(S.InvalidRegionError):
            S.region_dims(cs, roi)


# ---- device ---------------------------------------------------------------

def _crop(full, roi, den):
    x, y, w, h = _covering(roi, den)
    return full[y:y + h, x:x + w]


@pytest.mark.gpu
def test_region_equals_crop_of_full_decode(session, oracle):
    rng = np.random.default_rng(3)
    img = synth.tissue_rgb(rng, 300, 220)
    cases = [S.encode_j2k_lossless(img),
             S.encode_j2k_ht(img, tile_width=100, tile_height=64, levels=4),
             S.encode_j2k_ht(img, reversible=0, base_step=1 / 64, cb_width_log2=5,
                             cb_height_log2=5)]
    rois = [(0, 0, 300, 220), (0, 0, 1, 1), (299, 219, 1, 1), (37, 51, 64, 64),
            (100, 64, 100, 64), (5, 150, 290, 3), (171, 3, 2, 210), (63, 63, 130, 90)]
    for k, cs in enumerate(cases):
        d = S.J2kDecoder(cs, session)
        for den in (1, 2, 4, 8):
            full = O.decode(cs, 1, den)
            for roi in rois:
                got = d.decode_region_to_numpy(S.PixelFormat.RGB8, roi, S.Downscale(den))
                assert np.array_equal(got, _crop(full, roi, den)), (k, den, roi)


@pytest.mark.gpu
def test_region_formats(session, oracle):
    rng = np.random.default_rng(4)
    g12 = synth.spots_gray12(rng, 180, 140)
    cs12 = S.encode_j2k_ht(g12, bit_depth=12, use_mct=0, levels=4, cb_width_log2=5,
                           cb_height_log2=5)
    rgb = synth.texture_rgb(rng, 130, 110)
    cs8 = S.encode_j2k_lossless(rgb)
    roi = (17, 29, 77, 53)
    for cs, fmt in ((cs12, S.PixelFormat.GRAY16), (cs12, S.PixelFormat.GRAY8),
                    (cs8, S.PixelFormat.RGBA8), (cs8, S.PixelFormat.RGB16)):
        for den in (1, 2):
            full = O.decode(cs, int(fmt), den)
            got = S.J2kDecoder(cs, session).decode_region_to_numpy(fmt, roi, S.Downscale(den))
            assert np.array_equal(got, _crop(full, roi, den)), (fmt, den)


@pytest.mark.gpu
def test_region_batch_prunes_code_blocks(session, oracle):
    """A batch of regions (decode_tiles_region_scaled_into's request form):
    small regions of large tiles skip most code-blocks, and a region inside
    one tile of a multi-tile image skips the other tiles."""
    rng = np.random.default_rng(5)
    big = synth.tissue_rgb(rng, 1024, 1024)
    cs_big = S.encode_j2k_lossless(big)
    tiled = S.encode_j2k_ht(big[:512, :512].copy(), tile_width=128, tile_height=128, levels=4)
    reqs = [(cs_big, (500, 500, 40, 40), 1), (cs_big, (0, 0, 1024, 1024), 4),
            (tiled, (130, 260, 100, 100), 1), (cs_big, (1000, 3, 24, 700), 2)]
    plan = S.TilePlan(session, [r[0] for r in reqs], S.PixelFormat.RGB8,
                      [S.Downscale(r[2]) for r in reqs], rois=[r[1] for r in reqs])
    plan.decode(S.MODE_GRAPH)
    session.synchronize()
    for k, (cs, roi, den) in enumerate(reqs):
        want = _crop(O.decode(cs, 1, den), roi, den)
        assert np.array_equal(plan.surfaces[k].to_numpy(), want), k
    st = plan.stats()
    full = S.TilePlan(session, [r[0] for r in reqs], S.PixelFormat.RGB8,
                      [S.Downscale(r[2]) for r in reqs])
    full_blocks = full.stats()["code_blocks"]
    full.close()
    assert st["pruned_code_blocks"] > full_blocks // 3
    plan.close()
    # the host-buffer batch form, with an empty region
    outs = []
    jobs = []
    for cs, roi, den in reqs + [(cs_big, (10, 10, 0, 0), 1)]:
        _x, _y, w, h = S.region_dims(cs, roi, S.Downscale(den))
        out = np.zeros((h, w, 3), np.uint8)
        outs.append(out)
        jobs.append((cs, out, w * 3, roi, S.Downscale(den)))
    S.decode_tiles_region_scaled_into(jobs, S.PixelFormat.RGB8, session=session)
    for k, (cs, roi, den) in enumerate(reqs):
        assert np.array_equal(outs[k], _crop(O.decode(cs, 1, den), roi, den)), k
    assert outs[-1].size == 0


@pytest.mark.gpu
def test_region_with_image_origin_and_packed_headers(session, oracle):
    rng = np.random.default_rng(6)
    img = synth.texture_rgb(rng, 200, 150)
    cs = S.encode_j2k_ht(img, image_x0=37, image_y0=21, tile_x0=30, tile_y0=20,
                         tile_width=64, tile_height=64, packed_headers=2, levels=3)
    roi = (50, 40, 90, 70)
    got = S.J2kDecoder(cs, session).decode_region_to_numpy(S.PixelFormat.RGB8, roi)
    assert np.array_equal(got, img[40:110, 50:140])
