"""JP2 files and DICOM encapsulated Pixel Data around a codestream (CPU).

JP2 box rules follow the reference facade's parser
(crates/signinum-j2k/src/parse/boxes.rs:14-160); the files written here are
also read by OpenJPEG (through Pillow), which pins the box layout
independently.  The DICOM cases restate the reference's own unit tests
(crates/signinum-j2k-metal/src/dicom.rs:460-598) plus its error paths.
"""
import io
import struct

import numpy as np
import pytest

import signinum_amd as S
from signinum_amd import synth

SRGB, GREY, SYCC = 16, 17, 18


def box(kind, payload, form="short"):
    if form == "xl":
        return struct.pack(">I", 1) + kind + struct.pack(">Q", 16 + len(payload)) + payload
    if form == "open":
        return struct.pack(">I", 0) + kind + payload
    return struct.pack(">I", 8 + len(payload)) + kind + payload


def jp2(cs, width, height, ncomp, enumcs=SRGB, icc=None, extra_h=b"", cs_form="short",
        order=("sig", "ftyp", "jp2h", "jp2c")):
    """A JP2 file (T.800 Annex I) around codestream `cs`."""
    ihdr = struct.pack(">IIHBBBB", height, width, ncomp, 7, 7, 0, 0)
    if icc is not None:
        colr = bytes([2, 0, 0]) + icc
    else:
        colr = bytes([1, 0, 0]) + struct.pack(">I", enumcs)
    parts = {
        "sig": box(b"jP  ", b"\r\n\x87\n"),
        "ftyp": box(b"ftyp", b"jp2 " + struct.pack(">I", 0) + b"jp2 "),
        "jp2h": box(b"jp2h", box(b"ihdr", ihdr) + box(b"colr", colr) + extra_h),
        "jp2c": box(b"jp2c", bytes(cs), cs_form),
    }
    return b"".join(parts[k] for k in order)


def cdef(entries):
    return box(b"cdef", struct.pack(">H", len(entries)) +
               b"".join(struct.pack(">HHH", *e) for e in entries))


def pclr():
    # two entries, three 8-bit columns
    return box(b"pclr", struct.pack(">HB", 2, 3) + bytes([7, 7, 7]) + bytes([0, 0, 0, 255, 255, 255]))


def _rgb(seed=3, w=70, h=45):
    img = synth.gradient_noise_rgb(np.random.default_rng(seed), w, h)
    return img, S.encode_j2k_lossless(img)


@pytest.mark.parametrize("form", ["short", "xl", "open"])
def test_jp2_inspect_matches_codestream(form):
    img, cs = _rgb()
    raw = S.J2kDecoder.inspect(cs)
    info = S.J2kDecoder.inspect(jp2(cs, 70, 45, 3, cs_form=form))
    assert info["container"] == 1 and raw["container"] == 0
    assert raw["colorspace"] == 8  # Rct, inferred from the codestream
    assert info["colorspace"] == 5  # SRgb, from colr
    for k in raw:
        if k not in ("colorspace", "container"):
            assert info[k] == raw[k], k
    assert S.scaled_dims(jp2(cs, 70, 45, 3), S.Downscale.HALF) == S.scaled_dims(cs, S.Downscale.HALF)


def test_jp2_colour_specifications():
    _img, cs = _rgb()
    assert S.J2kDecoder.inspect(jp2(cs, 70, 45, 3, enumcs=GREY))["colorspace"] == 6
    assert S.J2kDecoder.inspect(jp2(cs, 70, 45, 3, enumcs=SYCC))["colorspace"] == 1
    assert S.J2kDecoder.inspect(jp2(cs, 70, 45, 3, enumcs=12))["colorspace"] == 7
    assert S.J2kDecoder.inspect(jp2(cs, 70, 45, 3, icc=b"\0" * 40))["colorspace"] == 7
    # the last colr box that names a colour space sets Info's (boxes.rs:118-120)
    two = jp2(cs, 70, 45, 3, extra_h=box(b"colr", bytes([1, 0, 0]) + struct.pack(">I", GREY)))
    assert S.J2kDecoder.inspect(two)["colorspace"] == 6


def test_jp2_written_here_is_read_by_openjpeg():
    """The box writer above against an independent JP2 reader."""
    from PIL import Image
    img, cs = _rgb(5, 96, 64)
    out = np.asarray(Image.open(io.BytesIO(jp2(cs, 96, 64, 3))).convert("RGB"))
    assert np.array_equal(out, img)


def test_jp2_written_by_openjpeg_is_inspected():
    from PIL import Image
    a = synth.gradient_noise_rgb(np.random.default_rng(9), 90, 33)
    b = io.BytesIO()
    Image.fromarray(a).save(b, format="JPEG2000", irreversible=False, num_resolutions=4)
    info = S.J2kDecoder.inspect(b.getvalue())
    assert (info["width"], info["height"], info["num_components"]) == (90, 33, 3)
    assert info["container"] == 1 and info["num_levels"] == 3 and info["high_throughput"] == 0
    assert info["colorspace"] == 5


@pytest.mark.parametrize("order,err", [
    # without the signature first the input is neither JP2 nor a codestream
    (("ftyp", "sig", "jp2h", "jp2c"), S.UnsupportedError),
    (("sig", "jp2h", "jp2c"), S.MalformedError),           # no ftyp
    (("sig", "ftyp", "jp2c"), S.MalformedError),           # no jp2h
    (("sig", "ftyp", "jp2c", "jp2h"), S.MalformedError),   # jp2h after jp2c
    (("sig", "ftyp", "jp2h"), S.MalformedError),           # no codestream
    (("sig", "sig", "ftyp", "jp2h", "jp2c"), S.MalformedError),
    (("sig", "ftyp", "jp2h", "jp2c", "jp2c"), S.MalformedError),
])
def test_jp2_box_order_is_enforced(order, err):
    _img, cs = _rgb()
    with pytest.raises(err):
        S.J2kDecoder.inspect(jp2(cs, 70, 45, 3, order=order))


def test_jp2_truncated_and_foreign_inputs():
    _img, cs = _rgb()
    f = jp2(cs, 70, 45, 3)
    with pytest.raises(S.MalformedError):
        S.J2kDecoder.inspect(f[:-10])
    with pytest.raises(S.UnsupportedError):
        S.J2kDecoder.inspect(b"GIF89a" + bytes(40))


def test_jp2_code_block_offsets_index_the_file():
    _img, cs = _rgb()
    f = jp2(cs, 70, 45, 3)
    at = f.index(bytes(cs))
    a = S.code_block_jobs(cs)
    b = S.code_block_jobs(f)
    assert len(a) == len(b) > 0
    for x, y in zip(a, b):
        if x.cleanup_length:
            assert y.data_offset == x.data_offset + at
            n = x.cleanup_length + x.refinement_length
            assert f[y.data_offset:y.data_offset + n] == cs[x.data_offset:x.data_offset + n]


# ---- DICOM (dicom.rs:460-598) ------------------------------------------------

def push_item(payload):
    return bytes([0xFE, 0xFF, 0x00, 0xE0]) + struct.pack("<I", len(payload)) + payload


def explicit_ob_pixel_data(bot, fragments, prefix=b""):
    out = bytes(128) + b"DICM" + prefix
    out += bytes([0xE0, 0x7F, 0x10, 0x00]) + b"OB" + bytes(2) + struct.pack("<I", 0xFFFFFFFF)
    out += push_item(bot)
    for f in fragments:
        out += push_item(f)
    return out + bytes([0xFE, 0xFF, 0xDD, 0xE0]) + struct.pack("<I", 0)


def ov_element(tag, payload):
    return bytes(tag) + b"OV" + bytes(2) + struct.pack("<I", len(payload)) + payload


def with_extended(offsets, lengths, bot, fragments):
    pre = ov_element([0xE0, 0x7F, 0x01, 0x00], b"".join(struct.pack("<Q", o) for o in offsets))
    pre += ov_element([0xE0, 0x7F, 0x02, 0x00], b"".join(struct.pack("<Q", n) for n in lengths))
    return explicit_ob_pixel_data(bot, fragments, pre)


def test_dicom_empty_basic_offset_table_extracts_each_fragment_as_frame():
    assert S.extract_dicom_encapsulated_frames(
        explicit_ob_pixel_data(b"", [b"first", b"second"])) == [b"first", b"second"]


def test_dicom_basic_offset_table_groups_fragments_into_frames():
    second = 8 + 2 + 8 + 2
    bot = struct.pack("<II", 0, second)
    data = explicit_ob_pixel_data(bot, [b"aa", b"bb", b"cc"])
    assert S.extract_dicom_encapsulated_frames(data) == [b"aabb", b"cc"]


def test_dicom_extended_offset_table_lengths_trim_padded_frame_payloads():
    data = with_extended([0, 8 + 4], [3, 2], b"", [b"abc\0", b"de"])
    assert S.extract_dicom_encapsulated_frames(data) == [b"abc", b"de"]


def test_dicom_extract_limit_stops_after_requested_empty_bot_frames():
    data = explicit_ob_pixel_data(b"", [b"first", b"second", b"third"])
    assert S.extract_dicom_encapsulated_frames(data, 2) == [b"first", b"second"]
    assert S.extract_dicom_encapsulated_frames(data, 0) == []


def test_dicom_extract_limit_stops_after_requested_basic_offset_frames():
    second = 8 + 2 + 8 + 2
    third = second + 8 + 2
    bot = struct.pack("<III", 0, second, third)
    data = explicit_ob_pixel_data(bot, [b"aa", b"bb", b"cc", b"dd"])
    assert S.extract_dicom_encapsulated_frames(data, 2) == [b"aabb", b"cc"]
    assert S.extract_dicom_encapsulated_frames(data) == [b"aabb", b"cc", b"dd"]


def test_dicom_errors():
    with pytest.raises(S.MalformedError, match="not found"):
        S.extract_dicom_encapsulated_frames(bytes(200))
    native = bytes(132) + bytes([0xE0, 0x7F, 0x10, 0x00]) + b"OB" + bytes(2) + \
        struct.pack("<I", 4) + bytes(4)
    with pytest.raises(S.UnsupportedError, match="not encapsulated"):
        S.extract_dicom_encapsulated_frames(native)
    with pytest.raises(S.MalformedError, match="multiple of four"):
        S.extract_dicom_encapsulated_frames(explicit_ob_pixel_data(b"abc", [b"x"]))
    with pytest.raises(S.MalformedError, match="strictly increasing"):
        S.extract_dicom_encapsulated_frames(
            explicit_ob_pixel_data(struct.pack("<II", 10, 10), [b"x" * 10, b"y"]))
    ok = explicit_ob_pixel_data(b"", [b"first", b"second"])
    with pytest.raises(S.MalformedError, match="truncated"):
        S.extract_dicom_encapsulated_frames(ok[:-12])
    with pytest.raises(S.MalformedError, match="different entry counts"):
        S.extract_dicom_encapsulated_frames(with_extended([0, 12], [3], b"", [b"abc\0", b"de"]))


def test_damaged_dicom_and_jp2_inputs_return_or_raise():
    """DICOM encapsulated pixel data (basic and extended offset tables) and
    a JP2 file, each damaged 1500 times (flipped bytes, truncation, 0xFF
    runs): extraction and inspect return or raise J2kError, and extracted
    frames never hold more bytes than the input."""
    rng = np.random.default_rng(19)
    second = 8 + 2 + 8 + 2
    seeds = [
        explicit_ob_pixel_data(struct.pack("<II", 0, second), [b"aa", b"bb", b"cc"]),
        with_extended([0, 10], [2, 2], b"", [b"ab", b"cd"]),
        jp2(_rgb()[1], 70, 45, 3),
    ]
    for src in seeds:
        for _ in range(1500):
            d = bytearray(src)
            kind = rng.integers(3)
            if kind == 0:
                for _ in range(int(rng.integers(1, 5))):
                    d[int(rng.integers(len(d)))] ^= 1 << int(rng.integers(8))
            elif kind == 1:
                d = d[:int(rng.integers(1, len(d)))]
            else:
                at = int(rng.integers(len(d)))
                d[at:at + 8] = b"\xff" * len(d[at:at + 8])
            d = bytes(d)
            try:
                frames = S.extract_dicom_encapsulated_frames(d)
                assert sum(len(f) for f in frames) <= len(d)
            except S.J2kError:
                pass
            try:
                S.J2kDecoder.inspect(d)
            except S.J2kError:
                pass
