"""HT code-blocks with SigProp / MagRef passes on the device, against the
oracle (the reference's decoder, including its SigProp context table):
every lane-kernel variant (cleanup in the lane kernel, then the passes in
the refinement kernel) and the cooperative kernel, graph and per-tile
modes, int16 and int32 sample planes, 5/3 and 9/7, blocks with refinement
only where lossless and in every block."""
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
    runs = [(v, S.MODE_GRAPH) for v in range(len(names))] + [(0, S.MODE_PLAIN)]
    for v, mode in runs:
        plan.lane_variant = v
        for s in plan.surfaces:
            s.fill(0)
        plan.decode(mode)
        session.synchronize()
        for k, (ref, rev) in enumerate(want):
            d = np.abs(plan.surfaces[k].to_numpy().astype(np.int64) - ref).max()
            assert d <= (0 if rev else 1), (names[v], mode, k, d)  # 9/7: 1 LSB of 8-bit
    plan.close()


def test_refinement_int32_samples(session, oracle, monkeypatch):
    """16-bit samples (more than 15 bit-planes: int32 code-block samples) and
    8-bit ones with the int16 planes switched off."""
    rng = np.random.default_rng(73)
    g = (synth.spots_gray12(rng, 150, 110).astype(np.uint32) * 16 + 7).astype(np.uint16)
    cs16 = S.encode_j2k_ht(g, bit_depth=16, use_mct=0, ht_refine=1)
    img = synth.tissue_rgb(rng, 140, 100)
    cs8 = S.encode_j2k_ht(img, ht_refine=2)
    names = S.TilePlan.lane_variants()
    want16 = O.decode(cs16, int(S.PixelFormat.GRAY16)).view("<u2").reshape(110, 150)
    for v in range(len(names)):
        plan = S.TilePlan(session, [cs16], S.PixelFormat.GRAY16)
        plan.lane_variant = v
        plan.decode(S.MODE_GRAPH)
        session.synchronize()
        assert np.array_equal(plan.surfaces[0].to_numpy()[..., 0], want16), names[v]
        plan.close()
    monkeypatch.setenv("SGN_COEF_NARROW", "0")
    plan = S.TilePlan(session, [cs8], S.PixelFormat.RGB8)
    for v in range(len(names)):
        plan.lane_variant = v
        plan.decode(S.MODE_GRAPH)
        session.synchronize()
        assert np.array_equal(plan.surfaces[0].to_numpy(), O.decode(cs8, 1)), names[v]
    plan.close()


def test_refined_single_tile_plain_path(session, oracle):
    img = synth.texture_rgb(np.random.default_rng(72), 300, 200)
    cs = S.encode_j2k_ht(img, ht_refine=2)
    got = S.J2kDecoder(cs, session).decode_to_numpy(S.PixelFormat.RGB8)
    assert np.array_equal(got, O.decode(cs, 1))
