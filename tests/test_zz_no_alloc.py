"""sgn_j2k_decode_tiles_device_noalloc: the batch decode whose plan lives in
the library's fixed device scratch array.  Its outputs equal those of
sgn_j2k_decode_tiles_device on two calls in a row, the second with a
smaller batch (the scratch keeps the first call's contents), and a batch
larger than the scratch is rejected before any device work."""
import ctypes

import numpy as np
import pytest

import signinum_amd as S
from signinum_amd import _lib, synth

pytestmark = pytest.mark.gpu


def _jobs(keep, surfaces):
    jobs = (S.TileJob * len(keep))()
    for i, (k, s) in enumerate(zip(keep, surfaces)):
        jobs[i] = S.TileJob(k.ctypes.data, k.size, s.device_ptr, s.pitch_bytes, s.byte_len)
    return jobs


def _decode(session, keep, surfaces, fmt, noalloc):
    L = _lib.load()
    ff = ctypes.c_size_t(0)
    jobs = _jobs(keep, surfaces)
    if noalloc:
        rc = L.sgn_j2k_decode_tiles_device_noalloc(session.handle, jobs, len(keep), int(fmt), 1,
                                                  ctypes.byref(ff))
    else:
        rc = L.sgn_j2k_decode_tiles_device(session.handle, jobs, len(keep), int(fmt), 1,
                                           S.MODE_BATCHED, ctypes.byref(ff))
    return rc, ff.value


def _case(n, size, seed):
    streams, pix, _cls = synth.lossless_batch_with_pixels(n, size=size, seed=seed)
    return [np.frombuffer(s, np.uint8) for s in streams], pix


def test_noalloc_matches_allocating_entry_point_twice(session):
    fmt = S.PixelFormat.RGB8
    for n, size, seed in ((24, 256, 11), (5, 128, 12)):
        keep, pix = _case(n, size, seed)
        want = S._alloc_surfaces(session, [(size, size)] * n, [fmt] * n)
        got = S._alloc_surfaces(session, [(size, size)] * n, [fmt] * n)
        for s in got:
            s.fill(0xA5)
        rc, _ = _decode(session, keep, want, fmt, False)
        assert rc == 0, _lib.last_error()
        rc, _ = _decode(session, keep, got, fmt, True)
        assert rc == 0, _lib.last_error()
        for i in range(n):
            a, b = got[i].to_numpy(), want[i].to_numpy()
            assert np.array_equal(a, b), (n, i)
            assert np.array_equal(a, pix[i]), (n, i)


def test_noalloc_irreversible_matches(session):
    rng = np.random.default_rng(97)
    fmt = S.PixelFormat.RGB8
    keep = [np.frombuffer(S.encode_j2k_ht(synth.tissue_rgb(rng, 200, 150), reversible=0,
                                          base_step=1 / 64), np.uint8) for _ in range(3)]
    want = S._alloc_surfaces(session, [(200, 150)] * 3, [fmt] * 3)
    got = S._alloc_surfaces(session, [(200, 150)] * 3, [fmt] * 3)
    assert _decode(session, keep, want, fmt, False)[0] == 0, _lib.last_error()
    assert _decode(session, keep, got, fmt, True)[0] == 0, _lib.last_error()
    for i in range(3):
        assert np.array_equal(got[i].to_numpy(), want[i].to_numpy()), i


def test_noalloc_rejects_batch_beyond_scratch(session):
    cap = _lib.load().sgn_j2k_noalloc_scratch_bytes()
    assert cap == 64 << 20
    # 1024x1024 RGB tiles need about 17 MB of coefficients each: eight of
    # them exceed the scratch
    img = synth.batch_tile(synth.SEED_BASE + 9, 0, 0, 1024)
    cs = np.frombuffer(S.encode_j2k_lossless(img), np.uint8)
    keep = [cs] * 8
    fmt = S.PixelFormat.RGB8
    surfs = S._alloc_surfaces(session, [(1024, 1024)] * 8, [fmt] * 8)
    rc, _ = _decode(session, keep, surfs, fmt, True)
    assert rc == _lib.SGN_E_ARG
    assert "scratch" in _lib.last_error()
    # and the next call that fits still decodes correctly
    rc, _ = _decode(session, keep[:1], surfs[:1], fmt, True)
    assert rc == 0, _lib.last_error()
    assert np.array_equal(surfs[0].to_numpy(), img)
