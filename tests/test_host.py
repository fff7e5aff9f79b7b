"""Host-side checks that need no GPU: the C ABI library loads and exports
every function declared in include/*.h; inspection, scaled geometry and the
code-block table agree with the oracle; error paths of the host parser."""
import ctypes
import glob
import os
import re

import numpy as np
import pytest

import signinum_amd as S
from signinum_amd import _lib, synth
from tests import oracle_ref as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(sgn_[a-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    declared = _declared_functions()
    assert len(declared) >= 24
    missing = [n for n in declared if not hasattr(lib, n)]
    assert not missing, missing
    # and the Python binding covers them all
    assert declared <= set(_lib.SIGNATURES), declared - set(_lib.SIGNATURES)


def test_inspect_matches_oracle(oracle):
    rng = np.random.default_rng(1)
    img = synth.tissue_rgb(rng, 77, 41)
    cs = S.encode_j2k_ht(img, levels=3, cb_width_log2=5, cb_height_log2=5, progression=2)
    a = S.J2kDecoder.inspect(cs)
    b = O.inspect(cs)
    assert {k: a[k] for k in b} == b
    # infer_colorspace (crates/signinum-j2k/src/parse/mod.rs:41-49): RCT
    # for three reversible components with MCT; raw codestream container
    assert (a["colorspace"], a["container"]) == (8, 0)
    ict = S.encode_j2k_ht(img, levels=3, reversible=0, base_step=1 / 64)
    assert S.J2kDecoder.inspect(ict)["colorspace"] == 9
    gray = S.encode_j2k_ht(img[..., 0].copy(), use_mct=0, levels=3)
    assert S.J2kDecoder.inspect(gray)["colorspace"] == 6


@pytest.mark.parametrize("size,den", [((240, 240), 4), ((1024, 256), 2), ((17, 37), 8),
                                      ((256, 256), 1)])
def test_scaled_dims_match_oracle(oracle, size, den):
    rng = np.random.default_rng(den)
    img = synth.texture_rgb(rng, size[0], size[1])
    cs = S.encode_j2k_lossless(img, levels=5)
    assert S.scaled_dims(cs, den) == O.scaled_dims(cs, den)


def test_code_block_table_matches_oracle_blocks(oracle):
    """Every listed HT job decodes on the oracle to the same coefficients
    the full oracle decode produces (via the refinement-zeroing counts)."""
    cs = open(os.path.join(ROOT, "tests", "golden", "openhtj2k_ds0_ht_09_b11.j2k"), "rb").read()
    jobs = S.code_block_jobs(cs)
    nonzero = [j for j in jobs if j.number_of_coding_passes > 0]
    assert len(nonzero) == 14
    assert sum(1 for j in nonzero if j.refinement_length > 0) == 14
    assert max(j.number_of_coding_passes for j in jobs) == 3
    for j in nonzero:
        data = cs[j.data_offset:j.data_offset + j.cleanup_length + j.refinement_length]
        rc, _blk = O.ht_decode_block(data, j.cleanup_length, j.refinement_length, j.width,
                                     j.height, j.missing_bit_planes, j.number_of_coding_passes)
        assert rc == 0


def test_malformed_inputs_raise():
    with pytest.raises(S.J2kError):
        S.J2kDecoder.inspect(b"\xff\x4f\xff\x51\x00")
    # neither JP2 nor a codestream: J2kError::Unsupported
    # (crates/signinum-j2k/src/parse/mod.rs:29-31)
    with pytest.raises(S.UnsupportedError):
        S.J2kDecoder.inspect(b"not a codestream")
    rng = np.random.default_rng(2)
    cs = S.encode_j2k_lossless(synth.texture_rgb(rng, 16, 16))
    with pytest.raises(S.J2kError):
        S.scaled_dims(cs, 3)


def test_code_block_exponents_beyond_a61_are_malformed():
    """COD code-block exponents (A.6.1): each at most 10 and a block at
    most 4096 samples; 256x64 and 2048x2 blocks are refused, 1024x4 is
    accepted."""
    rng = np.random.default_rng(5)
    cs = S.encode_j2k_ht(synth.texture_rgb(rng, 40, 24))
    at = cs.index(b"\xff\x52") + 10  # FF52, Lcod, Scod, SGcod (4 bytes), levels: xcb
    assert S.J2kDecoder.inspect(cs)["width"] == 40
    for xcb, ycb, ok in ((6, 4, False), (9, 0, False), (8, 0, True)):
        bad = bytearray(cs)
        bad[at], bad[at + 1] = xcb, ycb
        if ok:
            S.J2kDecoder.inspect(bytes(bad))
        else:
            with pytest.raises(S.J2kError):
                S.J2kDecoder.inspect(bytes(bad))


def test_encoder_rejects_bad_options():
    img = np.zeros((8, 8, 3), np.uint8)
    with pytest.raises(S.J2kError):
        S.encode_j2k_ht(img, cb_width_log2=11)
    with pytest.raises(TypeError):
        S.encode_j2k_ht(img, not_an_option=1)


def test_cpu_backend_request_is_rejected_before_any_work():
    rng = np.random.default_rng(4)
    cs = S.encode_j2k_lossless(synth.texture_rgb(rng, 16, 16))
    with pytest.raises(S.UnsupportedBackendError):
        S.decode_tiles_to_device(None, [cs], S.PixelFormat.RGB8, backend=S.BackendRequest.CPU)


@pytest.mark.parametrize("kw", [
    dict(), dict(progression=1), dict(progression=2, precinct_log2=5),
    dict(progression=3, precinct_log2=4), dict(progression=4, precinct_log2=6),
    dict(tile_width=100, tile_height=60, progression=2), dict(sop=1, eph=1),
    dict(levels=0), dict(cb_width_log2=4, cb_height_log2=7, levels=3),
])
def test_packet_parse_matches_oracle(oracle, kw):
    """The host packet-header parser yields the oracle's code-block table:
    same blocks in geometry order with equal lengths, Zblk and passes."""
    rng = np.random.default_rng(len(kw) * 13 + sum(map(int, kw.values())))
    img = synth.tissue_rgb(rng, 150, 90)
    cs = S.encode_j2k_ht(img, **kw)
    ours = [(j.width, j.height, j.cleanup_length, j.refinement_length, j.number_of_coding_passes,
             j.missing_bit_planes if j.number_of_coding_passes else None)
            for j in S.code_block_jobs(cs)]
    ref = [(c.x1 - c.x0, c.y1 - c.y0, c.lcup if c.passes else 0, c.lref if c.passes else 0,
            c.passes, c.zblk if c.passes else None)
           for c in O.code_blocks(cs) if c.x1 > c.x0 and c.y1 > c.y0]
    assert ours == ref


def test_packet_parse_matches_oracle_on_fixtures(oracle):
    for name in ("openhtj2k_ds0_ht_12_b11", "openhtj2k_ds0_ht_09_b11"):
        cs = open(os.path.join(ROOT, "tests", "golden", f"{name}.j2k"), "rb").read()
        ours = [(j.width, j.height, j.cleanup_length, j.refinement_length,
                 j.number_of_coding_passes) for j in S.code_block_jobs(cs)]
        ref = [(c.x1 - c.x0, c.y1 - c.y0, c.lcup if c.passes else 0, c.lref if c.passes else 0,
                c.passes) for c in O.code_blocks(cs) if c.x1 > c.x0 and c.y1 > c.y0]
        assert ours == ref


def test_device_packet_parser_matches_host_parser(tmp_path):
    """The device packet-header parser (packets.hpp, compiled for the host by
    packets_check.cpp with the host planner) against the host parser on HT
    codestreams over progressions, precincts, SOP/EPH, tiles, origins,
    refinement passes, 9/7, 12-bit gray and the OpenHTJ2K fixtures."""
    import shutil
    import subprocess
    if shutil.which("g++") is None:
        pytest.skip("needs g++")
    here = os.path.dirname(os.path.abspath(__file__))
    root = os.path.dirname(here)
    host = os.path.join(root, "signinum_amd", "csrc", "host")
    exe = tmp_path / "packets_check"
    subprocess.run(["g++", "-std=c++17", "-O1", "-w", "-o", str(exe),
                    os.path.join(here, "packets_check.cpp")] +
                   [os.path.join(host, f) for f in ("parse.cpp", "plan.cpp", "geometry.cpp",
                                                    "container.cpp")] + ["-lpthread"],
                   check=True)
    rng = np.random.default_rng(3)
    cases = [dict(), dict(progression=1), dict(progression=2, precinct_log2=5),
             dict(progression=3, precinct_log2=6), dict(progression=4, precinct_log2=5),
             dict(sop=1, eph=1), dict(sop=1, eph=1, progression=2, precinct_log2=5),
             dict(tile_width=64, tile_height=48), dict(ht_refine=2),
             dict(ht_refine=1, cb_width_log2=5, cb_height_log2=4),
             dict(reversible=0, base_step=1 / 128),
             dict(image_x0=5, image_y0=3, tile_width=64, tile_height=64),
             dict(levels=7, tile_width=64, tile_height=64),
             dict(cb_width_log2=4, cb_height_log2=6, levels=3)]
    files = []
    for k, kw in enumerate(cases):
        img = synth.tissue_rgb(rng, 150 + 7 * k, 100 + 3 * k)
        p = tmp_path / f"c{k}.j2k"
        p.write_bytes(S.encode_j2k_ht(img, **kw))
        files.append(str(p))
    g = synth.spots_gray12(rng, 130, 90)
    p = tmp_path / "g.j2k"
    p.write_bytes(S.encode_j2k_ht(g, bit_depth=12, use_mct=0, progression=2, precinct_log2=5))
    files.append(str(p))
    for name in ("openhtj2k_ds0_ht_09_b11.j2k", "openhtj2k_ds0_ht_12_b11.j2k"):
        files.append(os.path.join(here, "golden", name))
    out = subprocess.run([str(exe)] + files, capture_output=True, text=True, timeout=300)
    assert out.returncode == 0 and out.stdout.startswith("ok"), out.stdout + out.stderr


def test_planner_survives_damaged_codestreams(tmp_path):
    """plan_fuzz.cpp (the host planner under AddressSanitizer) on damaged
    HT, classic, multi-tile, PPM/PPT, JP2, region and OpenHTJ2K codestreams:
    every damaged stream is refused or planned with every job inside the
    buffers the kernels index."""
    import shutil
    import subprocess
    if shutil.which("g++") is None:
        pytest.skip("needs g++")
    here = os.path.dirname(os.path.abspath(__file__))
    host = os.path.join(os.path.dirname(here), "signinum_amd", "csrc", "host")
    exe = tmp_path / "plan_fuzz"
    subprocess.run(["g++", "-std=c++17", "-O1", "-g", "-w", "-fsanitize=address",
                    "-fno-omit-frame-pointer", "-o", str(exe), os.path.join(here, "plan_fuzz.cpp")] +
                   [os.path.join(host, f) for f in ("parse.cpp", "plan.cpp", "geometry.cpp",
                                                    "container.cpp")] + ["-lpthread"],
                   check=True)
    rng = np.random.default_rng(11)
    img = synth.tissue_rgb(rng, 150, 110)
    streams = [
        S.encode_j2k_ht(img),
        S.encode_j2k_ht(img, progression=2, precinct_log2=5, sop=1, eph=1, ht_refine=2),
        S.encode_j2k_ht(img, tile_width=64, tile_height=48, image_x0=5, image_y0=3),
        S.encode_j2k_ht(img, reversible=0, base_step=1 / 128, cb_width_log2=5, cb_height_log2=4),
        S.encode_j2k_classic(img, cb_style=0x3F & ~0x10),
        S.encode_j2k_classic(img, progression=1, precinct_log2=5, tile_width=64, tile_height=64),
    ]
    small = synth.tissue_rgb(rng, 120, 90)
    streams += [S.encode_j2k_ht(small, packed_headers=1, tile_width=64, tile_height=64),
                S.encode_j2k_classic(small, packed_headers=2, tile_width=64, tile_height=64),
                S.encode_j2k_ht(synth.spots_gray12(rng, 100, 70), bit_depth=12, use_mct=0)]
    from tests.test_containers import jp2
    streams.append(jp2(S.encode_j2k_ht(small), 120, 90, 3))  # JP2 boxes around it
    # 300 x 300 in 1 x 1 tiles: 90000 tiles, more than SOT can number (A.4.2)
    grid = bytearray(S.encode_j2k_ht(synth.texture_rgb(rng, 300, 300)))
    siz = grid.index(b"\xff\x51")
    grid[siz + 22:siz + 30] = (1).to_bytes(4, "big") * 2  # XTsiz, YTsiz
    streams.append(bytes(grid))
    files = []
    for k, cs in enumerate(streams):
        p = tmp_path / f"s{k}.j2k"
        p.write_bytes(cs)
        files.append(str(p))
    files += [os.path.join(here, "golden", n)
              for n in ("openhtj2k_ds0_ht_09_b11.j2k", "openhtj2k_ds0_ht_12_b11.j2k")]
    # no allocation above 2 GiB: a header must not size host tables by itself
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:max_allocation_size_mb=2048")
    out = subprocess.run([str(exe)] + files, capture_output=True, text=True, timeout=600, env=env)
    lines = out.stdout.splitlines()
    assert out.returncode == 0 and lines and lines[-1].startswith("ok"), \
        (out.stdout + out.stderr)[-3000:]
    status = {ln.split()[1]: int(ln.split()[2]) for ln in lines if ln.startswith("plan ")}
    assert status[files[len(streams) - 1]] == -3  # malformed
    assert all(status[f] == 0 for f in files if f != files[len(streams) - 1]), \
        {f: v for f, v in status.items() if v}
