"""HT code-blocks with SigProp / MagRef passes on the device, against the
oracle (the reference's decoder, including its SigProp context table):
every lane-kernel variant and the cooperative kernel, 5/3 and 9/7, blocks
with refinement only where lossless and in every block."""
import numpy as np
import pytest

import signinum_amd as S
from signinum_amd import synth
from tests import oracle_ref as O

pytestmark = pytest.mark.gpu


def test_refinement_passes_every_kernel(session, oracle):
    rng = np.random.default_rng(71)
    streams, want = [], []
    for k in range(12):
        gen = synth.tissue_rgb if k % 2 else synth.texture_rgb
        img = gen(rng, 96 + 16 * (k % 3), 80 + 8 * (k % 4))
        rev = 1 if k % 3 else 0
        cs = S.encode_j2k_ht(img, ht_refine=1 + k % 2, reversible=rev, base_step=1 / 256,
                             cb_width_log2=5 + (k % 2), cb_height_log2=6 - (k % 2))
        streams.append(cs)
        want.append((O.decode(cs, 1).astype(np.int64), rev))
    plan = S.TilePlan(session, streams, S.PixelFormat.RGB8)
    names = S.TilePlan.lane_variants()
    for v in range(len(names)):
        plan.lane_variant = v
        for s in plan.surfaces:
            s.fill(0)
        plan.decode(S.MODE_GRAPH)
        session.synchronize()
        for k, (ref, rev) in enumerate(want):
            d = np.abs(plan.surfaces[k].to_numpy().astype(np.int64) - ref).max()
            assert d <= (0 if rev else 1), (names[v], k, d)  # 9/7: 1 LSB of 8-bit
    plan.close()


def test_refined_single_tile_plain_path(session, oracle):
    img = synth.texture_rgb(np.random.default_rng(72), 300, 200)
    cs = S.encode_j2k_ht(img, ht_refine=2)
    got = S.J2kDecoder(cs, session).decode_to_numpy(S.PixelFormat.RGB8)
    assert np.array_equal(got, O.decode(cs, 1))
