"""Seeded synthetic tiles for the decode workloads.

The content classes follow the workload definitions: smooth gradients with
band-limited noise, near-white background, two-stain tissue, high texture,
and 12-bit Poisson spot images.  Every tile is reproducible from
(seed, index): tile i uses the sub-seed seed ^ (i * 0x9E3779B97F4A7C15).
"""
from concurrent.futures import ThreadPoolExecutor

import numpy as np

try:
    from scipy.ndimage import gaussian_filter as _gauss
except ImportError:  # pragma: no cover - scipy is present in the image
    _gauss = None

SEED_BASE = 20261202
_GOLDEN = 0x9E3779B97F4A7C15


def tile_rng(seed, index):
    return np.random.default_rng((seed ^ ((index * _GOLDEN) & 0xFFFFFFFFFFFFFFFF)) & 0xFFFFFFFFFFFFFFFF)


def _blur(a, sigma):
    if _gauss is not None:
        return _gauss(a, sigma, mode="reflect")
    # separable fallback
    r = int(3 * sigma + 0.5)
    k = np.exp(-0.5 * (np.arange(-r, r + 1) / sigma) ** 2)
    k /= k.sum()
    a = np.apply_along_axis(lambda v: np.convolve(np.pad(v, r, mode="reflect"), k, "valid"), 0, a)
    return np.apply_along_axis(lambda v: np.convolve(np.pad(v, r, mode="reflect"), k, "valid"), 1, a)


def gradient_noise_rgb(rng, w, h, sigma=12.0, blur=2.0):
    """Smooth 2-D gradient plus band-limited Gaussian noise, clamped to 0..255."""
    y, x = np.mgrid[0:h, 0:w].astype(np.float64)
    base = np.stack([255 * x / max(1, w - 1), 255 * y / max(1, h - 1),
                     255 * (x + y) / max(1, w + h - 2)], axis=-1)
    noise = np.stack([_blur(rng.standard_normal((h, w)), blur) for _ in range(3)], axis=-1)
    noise *= sigma / max(noise.std(), 1e-9)
    return np.clip(np.rint(base + noise), 0, 255).astype(np.uint8)


def near_white_rgb(rng, w, h):
    base = rng.uniform(235, 245, size=3)
    return np.clip(np.rint(base + rng.normal(0, 2, (h, w, 3))), 0, 255).astype(np.uint8)


def _lowpass01(rng, w, h, sigma):
    f = _blur(rng.standard_normal((h, w)), sigma)
    f -= f.min()
    return f / max(f.max(), 1e-9)


def tissue_rgb(rng, w, h):
    """Two low-pass fields mixing pink and purple stain vectors, intensity 60-220."""
    a = _lowpass01(rng, w, h, 6.0)
    b = _lowpass01(rng, w, h, 6.0)
    pink = np.array([235.0, 140.0, 190.0])
    purple = np.array([120.0, 70.0, 160.0])
    mix = a[..., None] * pink + (1 - a[..., None]) * purple
    inten = 60 + 160 * b
    img = mix * (inten[..., None] / 220.0)
    return np.clip(np.rint(img), 0, 255).astype(np.uint8)


def texture_rgb(rng, w, h):
    return np.clip(np.rint(rng.normal(128, 40, (h, w, 3))), 0, 255).astype(np.uint8)


def spots_gray12(rng, w, h):
    """Poisson(200) background with 1% bright spots at 3000-4000, blurred 1 px."""
    img = rng.poisson(200, (h, w)).astype(np.float64)
    spots = rng.random((h, w)) < 0.01
    img[spots] = rng.uniform(3000, 4000, spots.sum())
    img = _blur(img, 1.0)
    return np.clip(np.rint(img), 0, 4095).astype(np.uint16)


def batch_classes(n, seed):
    """Deterministic shuffle of 30% near-white, 50% tissue, 20% texture."""
    n_white = int(round(n * 0.3))
    n_tissue = int(round(n * 0.5))
    cls = np.array([0] * n_white + [1] * n_tissue + [2] * (n - n_white - n_tissue))
    np.random.default_rng(seed).shuffle(cls)
    return cls


def batch_tile(seed, i, cls, size=256):
    rng = tile_rng(seed, i)
    if cls == 0:
        return near_white_rgb(rng, size, size)
    if cls == 1:
        return tissue_rgb(rng, size, size)
    return texture_rgb(rng, size, size)


def lossless_batch(n, size=256, seed=SEED_BASE + 2, levels=5, workers=16):
    """The 256x256 RGB8 lossless batch: returns (codestreams, classes)."""
    from . import encode_j2k_lossless
    cls = batch_classes(n, seed)

    def make(i):
        img = batch_tile(seed, i, int(cls[i]), size)
        return encode_j2k_lossless(img, levels=levels)

    with ThreadPoolExecutor(max_workers=workers) as ex:
        streams = list(ex.map(make, range(n)))
    return streams, cls


def lossless_batch_with_pixels(n, size=256, seed=SEED_BASE + 2, levels=5, workers=16):
    from . import encode_j2k_lossless
    cls = batch_classes(n, seed)

    def make(i):
        img = batch_tile(seed, i, int(cls[i]), size)
        return img, encode_j2k_lossless(img, levels=levels)

    with ThreadPoolExecutor(max_workers=workers) as ex:
        pairs = list(ex.map(make, range(n)))
    return [p[1] for p in pairs], [p[0] for p in pairs], cls


def mixed_batch(n, seed=SEED_BASE + 3, workers=16):
    """Tiles of square sizes 240/256/512/1024 (equal shares): 75% RGB8 with the
    lossless-batch content classes, 25% 12-bit Poisson spot images in uint16;
    code-blocks 32x32 or 64x64, 3-6 levels, decoded at scale 1, 1/2 or 1/4.
    Returns [(codestream, PixelFormat, Downscale, source pixels)]."""
    from . import encode_j2k_ht, PixelFormat, Downscale
    sizes = (240, 256, 512, 1024)

    def make(i):
        rng = tile_rng(seed, i)
        size = sizes[i % 4]
        cb = 5 if rng.random() < 0.5 else 6
        levels = int(rng.integers(3, 7))
        scale = (Downscale.NONE, Downscale.HALF, Downscale.QUARTER)[int(rng.integers(0, 3))]
        if rng.random() < 0.75:
            img = batch_tile(seed, i, int(rng.integers(0, 3)), size)
            cs = encode_j2k_ht(img, bit_depth=8, levels=levels, cb_width_log2=cb,
                               cb_height_log2=cb, reversible=1, use_mct=1)
            return cs, PixelFormat.RGB8, scale, img
        img = spots_gray12(rng, size, size)
        cs = encode_j2k_ht(img, bit_depth=12, levels=levels, cb_width_log2=cb, cb_height_log2=cb,
                           reversible=1, use_mct=0)
        return cs, PixelFormat.GRAY16, scale, img

    with ThreadPoolExecutor(max_workers=workers) as ex:
        return list(ex.map(make, range(n)))


def lossy_batch(n, size=512, seed=SEED_BASE + 4, workers=16):
    """Tissue-like RGB8 tiles encoded with 9/7, ICT, 5 levels, 64x64
    code-blocks, base step 1/64 scaled by the sub-band gains, 2 guard bits.
    Returns [(codestream, source pixels)]."""
    from . import encode_j2k_ht

    def make(i):
        img = tissue_rgb(tile_rng(seed, i), size, size)
        return encode_j2k_ht(img, bit_depth=8, levels=5, reversible=0, use_mct=1,
                             base_step=1.0 / 64, guard_bits=2), img

    with ThreadPoolExecutor(max_workers=workers) as ex:
        return list(ex.map(make, range(n)))
