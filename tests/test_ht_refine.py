"""HT code-blocks with refinement passes (SigProp + MagRef after a cleanup
of the bits above bit 0), written by the package's encoder.

Pins: OpenJPEG 2.5.4 decodes every lossless (ht_refine=1) codestream to
its source, which checks the encoder's SigProp / MagRef bit layout against
an independent decoder.  The oracle restates the reference's SigProp
decoder, which keeps the significance of the stripe above in a table whose
next entry it overwrites with the current stripe's (ht_block_decode.rs:
1360-1363): where that entry feeds a later group's context the reference,
and so the oracle and the device, decode some standard codestreams
differently from OpenJPEG.  Elsewhere the oracle equals OpenJPEG and the
source."""
import io

import numpy as np
import pytest

import signinum_amd as S
from signinum_amd import synth
from tests import oracle_ref as O

PIL = pytest.importorskip("PIL.Image")


def _pil(cs):
    return np.asarray(PIL.open(io.BytesIO(cs)))


@pytest.mark.parametrize("seed,kind", [(1, "tissue"), (2, "texture"), (5, "tissue"),
                                       (7, "texture")])
def test_refined_lossless_read_by_openjpeg(oracle, seed, kind):
    gen = synth.tissue_rgb if kind == "tissue" else synth.texture_rgb
    img = gen(np.random.default_rng(seed), 160, 120)
    cs = S.encode_j2k_ht(img, ht_refine=1)
    assert np.array_equal(_pil(cs), img)
    O.decode(cs, 1)
    blocks, refined, passes = O.last_stats()
    assert refined > 0 and passes == 3


def test_oracle_equals_openjpeg_where_the_context_table_agrees(oracle):
    # dense content: every stripe's significance is the same in the
    # overwritten entry as in the stripe above, so the quirk never shows
    img = synth.texture_rgb(np.random.default_rng(2), 130, 90)
    cs = S.encode_j2k_ht(img, ht_refine=2)
    assert np.array_equal(O.decode(cs, 1), _pil(cs))
    assert np.array_equal(O.decode(cs, 1), img)


def test_reference_sigprop_context_deviation_is_restated(oracle):
    """A codestream where the reference's SigProp context differs from the
    standard's: OpenJPEG gives the source, the oracle (the reference) does
    not.  The device is compared with the oracle in the GPU tests."""
    img = synth.tissue_rgb(np.random.default_rng(1), 150, 100)
    cs = S.encode_j2k_ht(img, ht_refine=1)
    assert np.array_equal(_pil(cs), img)
    d = np.abs(O.decode(cs, 1).astype(int) - img).max()
    assert 0 < d <= 8


def test_refined_irreversible_close_to_openjpeg(oracle):
    img = synth.tissue_rgb(np.random.default_rng(3), 140, 100)
    cs = S.encode_j2k_ht(img, ht_refine=2, reversible=0, base_step=1 / 256)
    ours = O.decode(cs, 1).astype(int)
    assert np.abs(ours - img).max() <= 6
    assert np.abs(ours - _pil(cs).astype(int)).max() <= 6


def test_lane_refinement_passes_match_oracle(tmp_path):
    """The device's per-lane SigProp / MagRef passes (ht_refine.hpp, compiled
    for the host by refine_check.cpp) on the cleanup samples of every block
    with refinement passes, against the oracle's full block decode."""
    import os
    import shutil
    import struct
    import subprocess
    if shutil.which("g++") is None:
        pytest.skip("needs g++")
    here = os.path.dirname(os.path.abspath(__file__))
    root = os.path.dirname(here)
    exe = tmp_path / "refine_check"
    subprocess.run(["g++", "-std=c++17", "-O1", "-w", "-I", os.path.join(root, "oracle"),
                    "-o", str(exe), os.path.join(here, "refine_check.cpp")], check=True)
    cases = [(synth.tissue_rgb, dict(ht_refine=1), 160, 120),
             (synth.texture_rgb, dict(ht_refine=2), 133, 77),
             (synth.tissue_rgb, dict(ht_refine=2, reversible=0, base_step=1 / 256), 150, 100),
             (synth.texture_rgb, dict(ht_refine=1, cb_width_log2=5, cb_height_log2=6), 140, 90),
             (synth.tissue_rgb, dict(ht_refine=2, cb_width_log2=6, cb_height_log2=4), 120, 100)]
    total = 0
    for k, (gen, kw, w, h) in enumerate(cases):
        img = gen(np.random.default_rng(90 + k), w, h)
        streams = [S.encode_j2k_ht(img, **kw)]
        if k == 0:
            g = synth.spots_gray12(np.random.default_rng(99), 100, 70).astype(np.uint16) * 16
            streams.append(S.encode_j2k_ht(g, bit_depth=16, use_mct=0, ht_refine=1))
        for cs in streams:
            jobs = [j for j in S.code_block_jobs(cs) if j.number_of_coding_passes > 1]
            assert jobs
            path = tmp_path / f"blocks{k}.bin"
            with open(path, "wb") as f:
                f.write(struct.pack("<QQ", len(cs), len(jobs)))
                f.write(cs)
                for j in jobs:
                    f.write(struct.pack("<Q8I", j.data_offset, j.cleanup_length,
                                        j.refinement_length, j.width, j.height,
                                        j.missing_bit_planes, j.number_of_coding_passes,
                                        j.num_bitplanes, j.stripe_causal))
            out = subprocess.run([str(exe), str(path)], capture_output=True, text=True,
                                 timeout=120)
            assert out.returncode == 0 and out.stdout.startswith("ok"), out.stdout + out.stderr
            total += len(jobs)
    assert total > 100
