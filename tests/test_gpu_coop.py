"""The cooperative HT kernel (one code-block per wavefront) against the
oracle: as the kernel of the whole lane list of mixed-geometry and 9/7
plans, on code-block shapes that leave lanes idle or span several
64-column chunks, and on blocks of 16-bit samples."""
import numpy as np
import pytest

import signinum_amd as S
from signinum_amd import synth
from tests import oracle_ref as O

pytestmark = pytest.mark.gpu


def _coop_index():
    names = S.TilePlan.lane_variants()
    return names.index("wave-coop")


def test_coop_mixed_plan_matches_oracle(session, oracle):
    items = synth.mixed_batch(24)
    plan = S.TilePlan(session, [i[0] for i in items], [i[1] for i in items],
                      [i[2] for i in items])
    plan.lane_variant = _coop_index()
    for _ in range(2):
        for s in plan.surfaces:
            s.fill(0x5A)
        plan.decode(S.MODE_GRAPH)
        session.synchronize()
        for k, (cs, fmt, den, _src) in enumerate(items):
            assert np.array_equal(plan.surfaces[k].to_numpy(),
                                  O.decode(cs, int(fmt), int(den))), k
    plan.close()


def test_coop_irreversible_equals_lane_kernel(session, oracle):
    lossy = synth.lossy_batch(6, size=256)
    plan = S.TilePlan(session, [c for c, _ in lossy], S.PixelFormat.RGB8)
    plan.decode(S.MODE_BATCHED)
    session.synchronize()
    lane = [s.to_numpy().copy() for s in plan.surfaces]
    plan.lane_variant = _coop_index()
    for s in plan.surfaces:
        s.fill(0)
    plan.decode(S.MODE_BATCHED)
    session.synchronize()
    for k, (cs, _src) in enumerate(lossy):
        got = plan.surfaces[k].to_numpy()
        assert np.array_equal(got, lane[k]), k
        assert np.abs(got.astype(np.int32) - O.decode(cs, 1).astype(np.int32)).max() <= 1
    plan.close()


@pytest.mark.parametrize("w,h,cbw,cbh,levels", [
    (37, 19, 5, 5, 2),      # odd widths and heights, partial quads
    (150, 70, 4, 6, 3),     # 16x64 blocks: 48 idle lanes
    (200, 90, 6, 4, 3),     # 64x16 blocks
    (300, 40, 7, 5, 2),     # 128x32 blocks: two 64-column chunks per quad row
    (260, 12, 10, 2, 1),    # 1024x4 blocks (capped to the band width)
    (8, 300, 2, 10, 1),     # 4x1024 blocks: one quad per row
])
def test_coop_block_shapes(session, oracle, w, h, cbw, cbh, levels):
    rng = np.random.default_rng(w * 1000 + h)
    img = synth.texture_rgb(rng, w, h)
    cs = S.encode_j2k_ht(img, levels=levels, cb_width_log2=cbw, cb_height_log2=cbh)
    plan = S.TilePlan(session, [cs], S.PixelFormat.RGB8)
    plan.lane_variant = _coop_index()
    plan.decode(S.MODE_BATCHED)
    session.synchronize()
    assert np.array_equal(plan.surfaces[0].to_numpy(), img)
    plan.close()


def test_coop_sixteen_bit_samples(session, oracle):
    """16-bit gray: more magnitude bit-planes (wide quads, MagSgn rows of
    several ring fills) and int32 coefficient samples."""
    rng = np.random.default_rng(16)
    img = rng.integers(0, 65536, size=(96, 80), dtype=np.int64).astype(np.uint16)
    cs = S.encode_j2k_ht(img, bit_depth=16, use_mct=0, levels=3)
    plan = S.TilePlan(session, [cs], S.PixelFormat.GRAY16)
    plan.lane_variant = _coop_index()
    plan.decode(S.MODE_BATCHED)
    session.synchronize()
    assert np.array_equal(plan.surfaces[0].to_numpy()[..., 0], img)
    plan.close()


def test_coop_graph_batch_every_tile(session, oracle):
    streams, pix, _cls = synth.lossless_batch_with_pixels(128, size=256)
    plan = S.TilePlan(session, streams, S.PixelFormat.RGB8)
    plan.lane_variant = _coop_index()
    plan.decode(S.MODE_GRAPH)
    session.synchronize()
    for i, (s, img) in enumerate(zip(plan.surfaces, pix)):
        assert np.array_equal(s.to_numpy(), img), i
    plan.close()
