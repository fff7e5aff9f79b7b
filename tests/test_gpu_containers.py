"""Device decodes of codestreams inside JP2 files and DICOM encapsulated
Pixel Data: the same pixels as the bare codestream (and the oracle), and
a refusal of the JP2 files whose samples change after the decode."""
import numpy as np
import pytest

import signinum_amd as S
from signinum_amd import synth
from tests import oracle_ref as O
from tests.test_containers import (GREY, SYCC, cdef, explicit_ob_pixel_data, jp2, pclr,
                                   with_extended)

pytestmark = pytest.mark.gpu


def test_jp2_decode_equals_codestream_decode(session, oracle):
    img = synth.tissue_rgb(np.random.default_rng(41), 150, 97)
    cs = S.encode_j2k_lossless(img)
    for f in (jp2(cs, 150, 97, 3), jp2(cs, 150, 97, 3, cs_form="xl"),
              jp2(cs, 150, 97, 3, extra_h=cdef([(0, 0, 1), (1, 0, 2), (2, 0, 3)]))):
        d = S.J2kDecoder(f, session)
        assert np.array_equal(d.decode_to_numpy(S.PixelFormat.RGB8), img)
        half = d.decode_to_numpy(S.PixelFormat.RGB8, S.Downscale.HALF)
        assert np.array_equal(half, O.decode(cs, 1, 2))


def test_jp2_gray_batch(session, oracle):
    rng = np.random.default_rng(42)
    imgs = [synth.spots_gray12(rng, 64 + 8 * k, 48) for k in range(4)]
    files = [jp2(S.encode_j2k_ht(im, bit_depth=12, use_mct=0, levels=3), im.shape[1],
                 im.shape[0], 1, enumcs=GREY) for im in imgs]
    plan = S.TilePlan(session, files, S.PixelFormat.GRAY16)
    plan.decode(S.MODE_BATCHED)
    session.synchronize()
    for k, im in enumerate(imgs):
        assert np.array_equal(plan.surfaces[k].to_numpy()[..., 0], im), k
    plan.close()


@pytest.mark.parametrize("extra,enumcs", [
    (pclr(), 16),
    (cdef([(0, 0, 3), (1, 0, 2), (2, 0, 1)]), 16),
    (b"", SYCC),
])
def test_jp2_post_decode_transforms_are_refused(session, extra, enumcs):
    img = synth.tissue_rgb(np.random.default_rng(43), 40, 30)
    f = jp2(S.encode_j2k_lossless(img), 40, 30, 3, enumcs=enumcs, extra_h=extra)
    S.J2kDecoder.inspect(f)  # inspection still works
    with pytest.raises(S.UnsupportedError):
        S.J2kDecoder(f, session).decode_to_numpy(S.PixelFormat.RGB8)


def test_dicom_frames_decode_in_one_batch(session):
    rng = np.random.default_rng(44)
    imgs = [synth.tissue_rgb(rng, 96, 80) for _ in range(3)]
    streams = [bytes(S.encode_j2k_lossless(im)) for im in imgs]
    # frame 0 split over two fragments (Basic Offset Table grouping), frames
    # 1 and 2 one fragment each
    a, b = streams[0][:1000], streams[0][1000:]
    frags = [a, b, streams[1], streams[2]]
    offs, o = [], 0
    for k, f in enumerate(frags):
        if k != 1:
            offs.append(o)
        o += 8 + len(f)
    bot = b"".join(int(x).to_bytes(4, "little") for x in offs)
    frames = S.extract_dicom_encapsulated_frames(explicit_ob_pixel_data(bot, frags))
    assert frames == streams
    # padded frames trimmed by the Extended Offset Table lengths
    padded = [s + b"\0" if len(s) % 2 else s for s in streams]
    eoffs, o = [], 0
    for p in padded:
        eoffs.append(o)
        o += 8 + len(p)
    frames2 = S.extract_dicom_encapsulated_frames(
        with_extended(eoffs, [len(s) for s in streams], b"", padded))
    assert frames2 == streams
    outs = S.decode_tiles_to_device(session, frames, S.PixelFormat.RGB8)
    for k, im in enumerate(imgs):
        assert np.array_equal(outs[k].to_numpy(), im), k
