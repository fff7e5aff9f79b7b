"""Classic (EBCOT / MQ) code-blocks on the device against the oracle:
OpenJPEG-encoded codestreams of every code-block kernel shape (32x32,
64x64, elongated up to 1024x4), several layers, tiles and precincts, 9/7,
reduced resolution, regions, and plans that mix classic and HT tiles."""
import numpy as np
import pytest

import signinum_amd as S
from signinum_amd import synth
from tests import oracle_ref as O
from tests.test_classic import LOSSLESS, pil_encode

pytestmark = pytest.mark.gpu


def test_classic_configs_equal_oracle(session, oracle):
    img = synth.tissue_rgb(np.random.default_rng(31), 150, 100)
    for kw in LOSSLESS + [dict(irreversible=True), dict(quality_layers=[40, 20, 10]),
                          dict(irreversible=True, quality_layers=[60, 30, 15, 5])]:
        cs = pil_encode(img, **kw)
        got = S.J2kDecoder(cs, session).decode_to_numpy(S.PixelFormat.RGB8).astype(np.int64)
        d = np.abs(got - O.decode(cs, 1).astype(np.int64)).max()
        assert d <= (1 if kw.get("irreversible") else 0), (kw, d)  # 9/7: 1 LSB of 8-bit


def test_classic_reduced_resolution_and_region(session, oracle):
    img = synth.texture_rgb(np.random.default_rng(32), 300, 200)
    cs = pil_encode(img, quality_layers=[20, 0], tile_size=(128, 128))
    d = S.J2kDecoder(cs, session)
    for den in (2, 4, 8):
        got = d.decode_to_numpy(S.PixelFormat.RGB8, S.Downscale(den))
        assert np.array_equal(got, O.decode(cs, 1, den)), den
    roi = (70, 33, 150, 90)
    got = d.decode_region_to_numpy(S.PixelFormat.RGB8, roi)
    assert np.array_equal(got, img[33:123, 70:220])


def test_classic_and_ht_in_one_graph_plan(session, oracle):
    rng = np.random.default_rng(33)
    streams, want = [], []
    for k in range(24):
        img = synth.tissue_rgb(rng, 96 + 8 * (k % 5), 80 + 4 * (k % 3))
        if k % 3 == 0:
            cs = S.encode_j2k_lossless(img)
        elif k % 3 == 1:
            cs = pil_encode(img, codeblock_size=(32, 32) if k % 2 else (64, 64))
        else:
            cs = pil_encode(img, codeblock_size=(128, 8), quality_layers=[30, 0])
        streams.append(cs)
        want.append(img)
    plan = S.TilePlan(session, streams, S.PixelFormat.RGB8)
    for _ in range(2):
        for s in plan.surfaces:
            s.fill(0)
        plan.decode(S.MODE_GRAPH)
        session.synchronize()
        for k, s in enumerate(plan.surfaces):
            assert np.array_equal(s.to_numpy(), want[k]), k
    for s in plan.surfaces:
        s.fill(0)
    plan.decode(S.MODE_PLAIN)
    session.synchronize()
    assert all(np.array_equal(s.to_numpy(), w) for s, w in zip(plan.surfaces, want))
    plan.close()


def test_classic_gray16(session, oracle):
    g = synth.spots_gray12(np.random.default_rng(34), 130, 70).astype(np.uint16) * 11
    cs = pil_encode(g)
    got = S.J2kDecoder(cs, session).decode_to_numpy(S.PixelFormat.GRAY16)
    assert np.array_equal(got[..., 0], g)


def test_classic_styles_on_device(session, oracle):
    """Every code-block style of the package's classic encoder (raw bypass
    segments, context reset, termination on each pass, vertically causal
    context, segmentation symbols) in one batch, 5/3 and 9/7."""
    from tests.test_classic import STYLES
    rng = np.random.default_rng(35)
    streams, want = [], []
    for k, style in enumerate(STYLES):
        img = synth.tissue_rgb(rng, 120 + 4 * k, 90)
        rev = k % 2
        cs = S.encode_j2k_classic(img, cb_style=style, reversible=rev, base_step=1 / 256,
                                  cb_width_log2=5 + (k % 2), cb_height_log2=6 - (k % 2))
        streams.append(cs)
        want.append((O.decode(cs, 1).astype(np.int64), rev))
    plan = S.TilePlan(session, streams, S.PixelFormat.RGB8)
    plan.decode(S.MODE_GRAPH)
    session.synchronize()
    for k, (ref, rev) in enumerate(want):
        d = np.abs(plan.surfaces[k].to_numpy().astype(np.int64) - ref).max()
        assert d <= (0 if rev else 1), (STYLES[k], d)  # 9/7: 1 LSB of 8-bit
    plan.close()


@pytest.mark.parametrize("lanes", ["64", "32", "16"])
def test_classic_lanes_per_wave(session, oracle, monkeypatch, lanes):
    """Every active-lane count of the classic kernels (64, 32 or 16 blocks
    per wave) on blocks of every width class, against the oracle."""
    monkeypatch.setenv("SGN_CLASSIC_LANES", lanes)
    rng = np.random.default_rng(36)
    streams, want = [], []
    for k, kw in enumerate([dict(), dict(codeblock_size=(32, 32)), dict(codeblock_size=(256, 16)),
                            dict(codeblock_size=(1024, 4), num_resolutions=2)]):
        img = synth.tissue_rgb(rng, 140 + 8 * k, 100)
        cs = pil_encode(img, **kw)
        streams.append(cs)
        want.append(img)
    for _ in range(3):
        streams += streams[:4]
        want += want[:4]
    plan = S.TilePlan(session, streams, S.PixelFormat.RGB8)
    plan.decode(S.MODE_GRAPH)
    session.synchronize()
    for k, s in enumerate(plan.surfaces):
        assert np.array_equal(s.to_numpy(), want[k]), (lanes, k)
    plan.close()
