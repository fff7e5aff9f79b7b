// Inverse DWT kernels for gfx950, LDS-tiled.
//
// A work-group synthesises one 64x64 output block of one level.  It loads
// the block plus a halo (2 samples for 5/3, 4 for 9/7) straight from the
// four sub-bands, interleaving them by absolute parity (F.3.3) and applying
// whole-sample symmetric extension by reflection of the coordinates (F.3.7).
// Lifting on that window gives, in the block interior, exactly the values of
// the reference's in-place lifting over whole rows and columns
// (crates/signinum-j2k-native/src/j2c/idwt.rs:535-1143): rows first, then
// columns; 5/3 in integers (exact), 9/7 in f32 with fused multiply-add.
// Dequantisation is applied to the integer sub-band samples as they are
// loaded, so it is fused into the first synthesis level that reads them.
//
// sgn_final_kernel fuses the last synthesis level of every component with
// the inverse colour transform (mct.rs:71-156), the sign shift
// (decode.rs:1720-1752), rounding/saturation and interleaved packing
// (lib.rs:2202-2295, 1772-1835).
#include <hip/hip_runtime.h>

#include "../host/plan_types.h"

namespace sgn {

constexpr int kBlk = 64;
constexpr int kIdwtThreads = 256;
constexpr float kKappa = 1.2301741f;
constexpr float kNegAlpha = 1.5861343f;
constexpr float kNegBeta = 0.052980117f;
constexpr float kNegGamma = -0.8829111f;
constexpr float kNegDelta = -0.44350687f;

__device__ __forceinline__ int reflect(int i, int n) {
    if (n == 1) return 0;
    const int period = 2 * (n - 1);
    int m = i % period;
    if (m < 0) m += period;
    return m < n ? m : period - m;
}

template <bool REV>
struct Sample;
template <>
struct Sample<true> {
    using T = int32_t;
};
template <>
struct Sample<false> {
    using T = float;
};

// Value of level-local position (lx, ly) after extension.
template <bool REV>
__device__ __forceinline__ typename Sample<REV>::T load_level(const LevelJob &lj,
                                                              const uint32_t *__restrict__ coef,
                                                              int lx, int ly, int w, int h) {
    const uint32_t mx = (uint32_t)reflect(lx, w), my = (uint32_t)reflect(ly, h);
    uint32_t b, col, row;
    if (lj.levels == 0) {
        // no synthesis: the low-pass band is the level itself
        b = 0;
        col = mx;
        row = my;
    } else {
        const uint32_t ax = lj.x0 + mx, ay = lj.y0 + my;
        b = (ax & 1u) | ((ay & 1u) << 1);
        col = (ax & 1u) ? (ax >> 1) - (lj.x0 >> 1) : (ax >> 1) - ((lj.x0 + 1) >> 1);
        row = (ay & 1u) ? (ay >> 1) - (lj.y0 >> 1) : (ay >> 1) - ((lj.y0 + 1) >> 1);
    }
    const uint32_t v = coef[lj.band[b] + row * lj.bw[b] + col];
    const bool isf = (lj.float_in >> b) & 1u;
    if constexpr (REV) {
        return isf ? (int32_t)__uint_as_float(v) : (int32_t)v;
    } else {
        return isf ? __uint_as_float(v) : (float)(int32_t)v * lj.delta[b];
    }
}

// One-dimensional lifting over `count` lines of a window.  Element e of
// line l is at base[l * lstride + e * estride]; absolute parity of e is
// (origin + e) & 1.  Only elements 1..n-2 are updated.
template <bool REV, typename T>
__device__ __forceinline__ void lift_lines(T *base, int count, int lstride, int n, int estride,
                                           uint32_t origin, bool single, bool odd_single) {
    const int tid = threadIdx.x;
    if (single) {
        // length-1 signal: identity, or halved when it sits at an odd
        // coordinate (idwt.rs:726-731,942-957)
        if (!REV && odd_single)
            for (int k = tid; k < count * n; k += kIdwtThreads) {
                const int l = k / n, e = k - l * n;
                base[l * lstride + e * estride] = base[l * lstride + e * estride] * (T)0.5f;
            }
        __syncthreads();
        return;
    }
    const int half = n / 2 + 1;
    const int items = count * half;
    auto step = [&](uint32_t parity, auto f) {
        for (int k = tid; k < items; k += kIdwtThreads) {
            const int l = k / half;
            int e = 2 * (k - l * half);
            if (((origin + (uint32_t)e) & 1u) != parity) e += 1;
            if (e < 1 || e > n - 2) continue;
            T *p = base + l * lstride + e * estride;
            p[0] = f(p[0], p[-estride], p[estride]);
        }
        __syncthreads();
    };
    if constexpr (REV) {
        step(0u, [](int32_t s, int32_t a, int32_t b) { return s - ((a + b + 2) >> 2); });
        step(1u, [](int32_t s, int32_t a, int32_t b) { return s + ((a + b) >> 1); });
    } else {
        const float inv_k = 1.0f / kKappa;
        for (int k = tid; k < count * n; k += kIdwtThreads) {
            const int l = k / n, e = k - l * n;
            T *p = base + l * lstride + e * estride;
            p[0] = p[0] * (((origin + (uint32_t)e) & 1u) ? inv_k : kKappa);
        }
        __syncthreads();
        step(0u, [](float s, float a, float b) { return fmaf(a + b, kNegDelta, s); });
        step(1u, [](float s, float a, float b) { return fmaf(a + b, kNegGamma, s); });
        step(0u, [](float s, float a, float b) { return fmaf(a + b, kNegBeta, s); });
        step(1u, [](float s, float a, float b) { return fmaf(a + b, kNegAlpha, s); });
    }
}

// Synthesises block (bx, by) of one level into the window; returns the
// block's geometry.  Result element (x, y) of the block is at
// win[(y + H) * WN + x + H].
template <bool REV>
__device__ void synth_block(const LevelJob &lj, const uint32_t *__restrict__ coef, int bx, int by,
                            typename Sample<REV>::T *win, int &bw_out, int &bh_out) {
    constexpr int H = REV ? 2 : 4;
    constexpr int WN = kBlk + 2 * H;
    const int w = (int)(lj.x1 - lj.x0), h = (int)(lj.y1 - lj.y0);
    const int ox = bx * kBlk, oy = by * kBlk;
    const int bw = min(kBlk, w - ox), bh = min(kBlk, h - oy);
    const int nw = bw + 2 * H, nh = bh + 2 * H;
    for (int k = threadIdx.x; k < nw * nh; k += kIdwtThreads) {
        const int yy = k / nw, xx = k - yy * nw;
        win[yy * WN + xx] = load_level<REV>(lj, coef, ox - H + xx, oy - H + yy, w, h);
    }
    __syncthreads();
    if (lj.levels) {
        // rows (all window rows), origin = absolute x of window column 0
        lift_lines<REV>(win, nh, WN, nw, 1, lj.x0 + (uint32_t)(ox - H), w == 1, (lj.x0 & 1u) != 0);
        // columns of the block interior
        lift_lines<REV>(win + H, bw, 1, nh, WN, lj.y0 + (uint32_t)(oy - H), h == 1,
                        (lj.y0 & 1u) != 0);
    }
    bw_out = bw;
    bh_out = bh;
}

template <typename J>
__device__ __forceinline__ uint32_t find_job(const J *jobs, uint32_t njobs, uint32_t block) {
    uint32_t lo = 0, hi = njobs;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (jobs[mid].block_base <= block)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

template <bool REV>
__device__ void level_body(const LevelJob *jobs, uint32_t njobs, uint32_t block_offset,
                           uint32_t *coef) {
    constexpr int H = REV ? 2 : 4;
    constexpr int WN = kBlk + 2 * H;
    __shared__ typename Sample<REV>::T win[WN * WN];
    const uint32_t blk = blockIdx.x + block_offset;
    const LevelJob lj = jobs[find_job(jobs, njobs, blk)];
    const uint32_t local = blk - lj.block_base;
    const int bx = (int)(local % lj.nbx), by = (int)(local / lj.nbx);
    int bw, bh;
    synth_block<REV>(lj, coef, bx, by, win, bw, bh);
    const int w = (int)(lj.x1 - lj.x0);
    for (int k = threadIdx.x; k < bw * bh; k += kIdwtThreads) {
        const int y = k / bw, x = k - y * bw;
        const auto v = win[(y + H) * WN + x + H];
        uint32_t bits;
        if constexpr (REV)
            bits = (uint32_t)v;
        else
            bits = __float_as_uint(v);
        coef[lj.out + (uint32_t)(by * kBlk + y) * (uint32_t)w + (uint32_t)(bx * kBlk + x)] = bits;
    }
}

extern "C" __global__ void __launch_bounds__(kIdwtThreads)
sgn_idwt_level_rev(const LevelJob *jobs, uint32_t njobs, uint32_t block_offset, uint32_t *coef) {
    level_body<true>(jobs, njobs, block_offset, coef);
}

extern "C" __global__ void __launch_bounds__(kIdwtThreads)
sgn_idwt_level_irr(const LevelJob *jobs, uint32_t njobs, uint32_t block_offset, uint32_t *coef) {
    level_body<false>(jobs, njobs, block_offset, coef);
}

// ---- final stage -------------------------------------------------------

__device__ __forceinline__ uint8_t sat_u8(float v) {
    // Rust `as u8`: NaN -> 0, saturating
    if (!(v == v) || v <= 0.0f) return 0;
    if (v >= 255.0f) return 255;
    return (uint8_t)v;
}

template <bool REV>
__device__ void final_body(const FinalJob *jobs, uint32_t njobs, uint32_t block_offset,
                           const LevelJob *lv, const uint32_t *coef) {
    constexpr int H = REV ? 2 : 4;
    constexpr int WN = kBlk + 2 * H;
    using T = typename Sample<REV>::T;
    __shared__ T win[WN * WN];
    __shared__ T res[2][kBlk * kBlk];  // all but the last component; the last stays in win
    const uint32_t blk = blockIdx.x + block_offset;
    const FinalJob fj = jobs[find_job(jobs, njobs, blk)];
    const uint32_t local = blk - fj.block_base;
    const int bx = (int)(local % fj.nbx), by = (int)(local / fj.nbx);
    int bw = 0, bh = 0;
    const int nc = fj.ncomp > 3 ? 3 : fj.ncomp;
    for (int c = 0; c < nc; ++c) {
        const LevelJob lj = lv[fj.level[c]];
        synth_block<REV>(lj, coef, bx, by, win, bw, bh);
        if (c + 1 < nc) {
            for (int k = threadIdx.x; k < bw * bh; k += kIdwtThreads) {
                const int y = k / bw, x = k - y * bw;
                res[c][k] = win[(y + H) * WN + x + H];
            }
            __syncthreads();
        }
    }
    const int fmt = fj.fmt;
    const int ch = (fmt == 0 || fmt == 3) ? 1 : (fmt == 2 ? 4 : 3);
    const int bps = (fmt == 3 || fmt == 4) ? 2 : 1;
    uint8_t *dst = reinterpret_cast<uint8_t *>(fj.dst);
    for (int k = threadIdx.x; k < bw * bh; k += kIdwtThreads) {
        const int y = k / bw, x = k - y * bw;
        float v[4];
        if constexpr (REV) {
            int32_t s[4];
            for (int c = 0; c < nc; ++c)
                s[c] = c + 1 < nc ? res[c][k] : win[(y + H) * WN + x + H];
            if (fj.mct) {
                const int32_t i1 = s[0] - ((s[2] + s[1]) >> 2);
                const int32_t r = s[2] + i1, b = s[1] + i1;
                s[0] = r;
                s[1] = i1;
                s[2] = b;
            }
            for (int c = 0; c < nc; ++c) v[c] = (float)(s[c] + (1 << (fj.precision[c] - 1)));
        } else {
            for (int c = 0; c < nc; ++c)
                v[c] = c + 1 < nc ? res[c][k] : win[(y + H) * WN + x + H];
            if (fj.mct) {
                const float y0 = v[0], y1 = v[1], y2 = v[2];
                v[0] = fmaf(y2, 1.402f, y0);
                v[1] = fmaf(y2, -0.71414f, fmaf(y1, -0.34413f, y0));
                v[2] = fmaf(y1, 1.772f, y0);
            }
            for (int c = 0; c < nc; ++c) v[c] = v[c] + (float)(1u << (fj.precision[c] - 1));
        }
        const int gx = bx * kBlk + x, gy = by * kBlk + y;
        uint8_t *px = dst + (size_t)gy * fj.pitch + (size_t)gx * ch * bps;
        if (bps == 1) {
            for (int c = 0; c < nc && c < ch; ++c) {
                uint8_t o;
                if (fj.same8) {
                    o = sat_u8(roundf(v[c]));
                } else {
                    const float mx = (float)((1u << fj.precision[c]) - 1);
                    o = sat_u8(roundf((v[c] / mx) * 255.0f));
                }
                px[c] = o;
            }
            if (ch == 4) px[3] = 255;
        } else {
            const uint32_t bd = fj.precision[0];
            const float mx = (float)((1ull << bd) - 1);
            for (int c = 0; c < nc && c < ch; ++c) {
                const float r = roundf(v[c]);
                const float cl = r < 0.0f ? 0.0f : (r > mx ? mx : r);
                uint32_t s;
                if (bd <= 8) {
                    uint32_t m8 = (1u << bd) - 1;
                    if (m8 == 0) m8 = 1;
                    s = ((uint32_t)cl * 65535u + m8 / 2) / m8;
                } else {
                    s = (uint32_t)(cl > 65535.0f ? 65535.0f : cl);
                }
                px[2 * c] = (uint8_t)(s & 0xFF);
                px[2 * c + 1] = (uint8_t)(s >> 8);
            }
        }
    }
}

extern "C" __global__ void __launch_bounds__(kIdwtThreads)
sgn_final_rev(const FinalJob *jobs, uint32_t njobs, uint32_t block_offset, const LevelJob *lv,
              const uint32_t *coef) {
    final_body<true>(jobs, njobs, block_offset, lv, coef);
}

extern "C" __global__ void __launch_bounds__(kIdwtThreads)
sgn_final_irr(const FinalJob *jobs, uint32_t njobs, uint32_t block_offset, const LevelJob *lv,
              const uint32_t *coef) {
    final_body<false>(jobs, njobs, block_offset, lv, coef);
}

// Byte fill used instead of memset nodes (safe under graph capture).
extern "C" __global__ void sgn_fill_bytes(uint8_t *p, size_t bytes, size_t words, uint32_t v4) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    uint32_t *w = reinterpret_cast<uint32_t *>(p);
    const bool aligned = (reinterpret_cast<uintptr_t>(p) & 3) == 0;
    if (aligned) {
        for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < words; i += stride)
            w[i] = v4;
        for (size_t i = words * 4 + (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < bytes;
             i += stride)
            p[i] = (uint8_t)v4;
    } else {
        for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < bytes; i += stride)
            p[i] = (uint8_t)v4;
    }
}

// Inverse MCT on f32 planes (J2kInverseMctJob hook, lib.rs:785-800).
extern "C" __global__ void sgn_inverse_mct_planes(float *p0, float *p1, float *p2, uint32_t n,
                                                  int reversible, float a0, float a1, float a2) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float y0 = p0[i], y1 = p1[i], y2 = p2[i];
    float r, g, b;
    if (reversible) {
        const float i1 = y0 - floorf((y2 + y1) * 0.25f);
        r = y2 + i1;
        g = i1;
        b = y1 + i1;
    } else {
        r = fmaf(y2, 1.402f, y0);
        g = fmaf(y2, -0.71414f, fmaf(y1, -0.34413f, y0));
        b = fmaf(y1, 1.772f, y0);
    }
    p0[i] = r + a0;
    p1[i] = g + a1;
    p2[i] = b + a2;
}

}  // namespace sgn
