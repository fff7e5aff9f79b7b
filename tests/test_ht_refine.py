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
