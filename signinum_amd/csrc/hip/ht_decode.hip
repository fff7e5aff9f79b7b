// HT code-block decode kernel for the general case, gfx950.
//
// ht_block_wave: one code-block per wavefront, any size and all three HT
//   passes (cleanup, SigProp, MagRef).  The symbol table, sigma map and the
//   block itself live in LDS; lane 0 walks the serial bit streams, then the
//   whole wave converts and stores the block with coalesced writes.  The
//   cleanup-only blocks of at most 64x64 samples (every block of the lossless
//   batches) take the lane kernel in ht_rows.hip instead.
//
// Restates crates/signinum-j2k-native/src/j2c/ht_block_decode.rs (decode_impl
// :1563-1680 and the phases it calls) and produces coefficient_to_i32
// integers in the sub-band arena.
#include <hip/hip_runtime.h>

#include "../host/plan_types.h"
#include "ht_common.hpp"

namespace sgn {

__device__ __forceinline__ void load_tables(const uint16_t *g, uint16_t *s, int n, int tid,
                                            int nthreads) {
    for (int i = tid; i < n; i += nthreads) s[i] = g[i];
}

// Magnitude/sign of one sample (ht_block_decode.rs:1010-1032); returns the
// u32 sign-magnitude sample and the decoded v_n.
__device__ __forceinline__ uint32_t ms_sample(FwdReader &ms, uint32_t inf, uint32_t bit,
                                              uint32_t uq, uint32_t p, uint32_t &vn) {
    if ((inf & (1u << (4 + bit))) == 0) {
        vn = 0;
        return 0;
    }
    const uint32_t ms_val = fwd_fetch(ms);
    const uint32_t m_n = uq - ((inf >> (12 + bit)) & 1u);
    fwd_advance(ms, m_n);
    uint32_t value = ms_val << 31;
    const uint32_t mask = m_n == 0 ? 0u : (m_n >= 32 ? 0xFFFFFFFFu : ((1u << m_n) - 1));
    uint32_t v = ms_val & mask;
    v |= ((inf >> (8 + bit)) & 1u) << m_n;
    v |= 1u;
    value |= (v + 2) << (p - 1);
    vn = v;
    return value;
}

// ---------------------------------------------------------------------------
// General path: one wavefront per code-block, all HT passes.

constexpr int kWaveMaxSamples = 4096;
constexpr int kWaveSymWords = 4800;   // u16 symbol table: sstr * (ceil(h/2) + 1)
constexpr int kWaveSigmaWords = 2200; // u16 sigma map: mstr * (ceil(h/4) + 1)
constexpr int kWaveVn = 520;          // u32 v_n: ceil(w/2) + 2
constexpr int kWavePrs = 272;         // u16 previous-row significance: ceil(w/4) + 8

__device__ uint32_t wave_cleanup_symbols(const uint8_t *d, uint32_t lcup, uint32_t scup,
                                         uint32_t w, uint32_t h, uint32_t sstr, uint16_t *sc,
                                         const uint16_t *vlc0, const uint16_t *vlc1,
                                         const uint16_t *uvlc0, const uint16_t *uvlc1) {
    MelState mel;
    mel.d = d;
    mel.pos = lcup - scup;
    mel.remaining = scup - 1;
    mel.unstuff = 0;
    mel.cur = 0;
    mel.bits_left = 0;
    mel.k = 0;
    mel.num_runs = 0;
    mel.runs = 0;
    RevReader vlc;
    rev_init_vlc(vlc, d, lcup, scup);
    int run = mel_run(mel);
    uint32_t c_q = 0, ro = 0;
    for (uint32_t x = 0; x < w;) {
        uint32_t vv = rev_fetch(vlc);
        uint32_t t0 = vlc0[c_q + (vv & 0x7F)];
        if (c_q == 0) {
            run -= 2;
            t0 = (run == -1) ? t0 : 0;
            if (run < 0) run = mel_run(mel);
        }
        sc[ro] = (uint16_t)t0;
        x += 2;
        c_q = ((t0 & 0x10) << 3) | ((t0 & 0xE0) << 2);
        vv = rev_advance(vlc, t0 & 7);
        uint32_t t1 = vlc0[c_q + (vv & 0x7F)];
        if (c_q == 0 && x < w) {
            run -= 2;
            t1 = (run == -1) ? t1 : 0;
            if (run < 0) run = mel_run(mel);
        }
        if (x >= w) t1 = 0;
        sc[ro + 2] = (uint16_t)t1;
        x += 2;
        c_q = ((t1 & 0x10) << 3) | ((t1 & 0xE0) << 2);
        vv = rev_advance(vlc, t1 & 7);
        uint32_t mode = ((t0 & 8) << 3) | ((t1 & 8) << 4);
        if (mode == 0xC0) {
            run -= 2;
            if (run == -1) mode += 0x40;
            if (run < 0) run = mel_run(mel);
        }
        uint32_t e = uvlc0[mode + (vv & 0x3F)];
        vv = rev_advance(vlc, e & 7);
        e >>= 3;
        uint32_t len = e & 0xF;
        const uint32_t tmp = vv & ((1u << len) - 1);
        rev_advance(vlc, len);
        e >>= 4;
        len = e & 7;
        e >>= 3;
        sc[ro + 1] = (uint16_t)(1 + (e & 7) + (tmp & ~(0xFFu << len)));
        sc[ro + 3] = (uint16_t)(1 + (e >> 3) + (tmp >> len));
        ro += 4;
    }
    sc[ro] = 0;
    sc[ro + 1] = 0;
    for (uint32_t y = 2; y < h; y += 2) {
        const uint32_t rb = (y >> 1) * sstr, pb = rb - sstr;
        c_q = 0;
        ro = rb;
        for (uint32_t x = 0; x < w;) {
            const uint32_t o = ro - rb;
            c_q |= ((uint32_t)sc[pb + o] & 0xA0) << 2;
            c_q |= ((uint32_t)sc[pb + o + 2] & 0x20) << 4;
            uint32_t vv = rev_fetch(vlc);
            uint32_t t0 = vlc1[c_q + (vv & 0x7F)];
            if (c_q == 0) {
                run -= 2;
                t0 = (run == -1) ? t0 : 0;
                if (run < 0) run = mel_run(mel);
            }
            sc[ro] = (uint16_t)t0;
            x += 2;
            c_q = ((t0 & 0x40) << 2) | ((t0 & 0x80) << 1);
            c_q |= (uint32_t)sc[pb + o] & 0x80;
            c_q |= ((uint32_t)sc[pb + o + 2] & 0xA0) << 2;
            c_q |= ((uint32_t)sc[pb + o + 4] & 0x20) << 4;
            vv = rev_advance(vlc, t0 & 7);
            uint32_t t1 = vlc1[c_q + (vv & 0x7F)];
            if (c_q == 0 && x < w) {
                run -= 2;
                t1 = (run == -1) ? t1 : 0;
                if (run < 0) run = mel_run(mel);
            }
            if (x >= w) t1 = 0;
            sc[ro + 2] = (uint16_t)t1;
            x += 2;
            c_q = ((t1 & 0x40) << 2) | ((t1 & 0x80) << 1);
            c_q |= (uint32_t)sc[pb + o + 2] & 0x80;
            vv = rev_advance(vlc, t1 & 7);
            const uint32_t mode = ((t0 & 8) << 3) | ((t1 & 8) << 4);
            uint32_t e = uvlc1[mode + (vv & 0x3F)];
            vv = rev_advance(vlc, e & 7);
            e >>= 3;
            uint32_t len = e & 0xF;
            const uint32_t tmp = vv & ((1u << len) - 1);
            rev_advance(vlc, len);
            e >>= 4;
            len = e & 7;
            e >>= 3;
            sc[ro + 1] = (uint16_t)((e & 7) + (tmp & ~(0xFFu << len)));
            sc[ro + 3] = (uint16_t)((e >> 3) + (tmp >> len));
            ro += 4;
        }
        sc[ro] = 0;
        sc[ro + 1] = 0;
    }
    return 0;
}

__device__ uint32_t wave_magsgn(const uint8_t *d, uint32_t lcup, uint32_t scup,
                                const uint16_t *sc, uint32_t *blk, uint32_t zblk, uint32_t w,
                                uint32_t h, uint32_t sstr, uint32_t *v_n) {
    const uint32_t vn_w = (w + 1) / 2 + 2;
    for (uint32_t k = 0; k < vn_w; ++k) v_n[k] = 0;
    const uint32_t p = 30 - zblk, mmsbp2 = zblk + 2;
    FwdReader ms;
    fwd_init(ms, d, lcup - scup, 0xFF);
    for (uint32_t y = 0; y < h; y += 2) {
        uint32_t sp = (y >> 1) * sstr, vp = 0, dp = y * w, prev_v = 0;
        const bool second = (y + 1 < h);
        for (uint32_t x = 0; x < w;) {
            const uint32_t inf = sc[sp];
            uint32_t uq = sc[sp + 1];
            if (y > 0) {
                uint32_t gamma = inf & 0xF0;
                gamma &= gamma - 0x10;
                uint32_t emax = v_n[vp] | v_n[vp + 1];
                emax = 31 - __clz(emax | 2);
                uq += gamma ? emax : 1u;
            }
            if (uq > mmsbp2) return 1;
            uint32_t vv, v1, v3;
            blk[dp] = ms_sample(ms, inf, 0, uq, p, vv);
            const uint32_t a1 = ms_sample(ms, inf, 1, uq, p, v1);
            if (second) blk[dp + w] = a1;
            v_n[vp] = prev_v | v1;
            prev_v = 0;
            ++dp;
            ++x;
            if (x >= w) {
                ++vp;
                break;
            }
            blk[dp] = ms_sample(ms, inf, 2, uq, p, vv);
            const uint32_t a3 = ms_sample(ms, inf, 3, uq, p, v3);
            if (second) blk[dp + w] = a3;
            prev_v = v3;
            ++dp;
            ++x;
            sp += 2;
            ++vp;
        }
        v_n[vp] = prev_v;
    }
    return 0;
}

__device__ void wave_sigma(const uint16_t *sc, uint16_t *sigma, uint32_t w, uint32_t h,
                           uint32_t sstr, uint32_t mstr) {
    for (uint32_t y = 0; y < h; y += 4) {
        uint32_t sp = (y >> 1) * sstr, dp = (y >> 2) * mstr;
        for (uint32_t x = 0; x < w; x += 4, sp += 4, ++dp) {
            uint32_t t0 = (((uint32_t)sc[sp] & 0x30) >> 4) | (((uint32_t)sc[sp] & 0xC0) >> 2);
            t0 |= (((uint32_t)sc[sp + 2] & 0x30) << 4) | (((uint32_t)sc[sp + 2] & 0xC0) << 6);
            uint32_t t1 = (((uint32_t)sc[sp + sstr] & 0x30) >> 2) | ((uint32_t)sc[sp + sstr] & 0xC0);
            t1 |= (((uint32_t)sc[sp + sstr + 2] & 0x30) << 6) |
                  (((uint32_t)sc[sp + sstr + 2] & 0xC0) << 8);
            sigma[dp] = (uint16_t)(t0 | t1);
        }
        sigma[dp] = 0;
    }
    const uint32_t base = ((h + 3) / 4) * mstr;
    for (uint32_t x = 0; x <= (w + 3) / 4; ++x) sigma[base + x] = 0;
}

__device__ __forceinline__ uint32_t pair16(const uint16_t *v, uint32_t i) {
    return (uint32_t)v[i] | ((uint32_t)v[i + 1] << 16);
}

__device__ void wave_sigprop(const uint8_t *ref, uint32_t lref, const uint16_t *sigma,
                             uint32_t *blk, uint32_t w, uint32_t h, uint32_t mstr,
                             bool stripe_causal, uint32_t p, uint16_t *prs) {
    const uint32_t spread[16] = {0x33,   0x76,   0xEC,   0xC8,   0x330,   0x760,   0xEC0,   0xC80,
                                 0x3300, 0x7600, 0xEC00, 0xC800, 0x33000, 0x76000, 0xEC000, 0xC8000};
    const uint32_t prs_len = (w + 3) / 4 + 8;
    for (uint32_t k = 0; k < prs_len; ++k) prs[k] = 0;
    FwdReader sp;
    fwd_init(sp, ref, lref, 0);
    for (uint32_t y = 0; y < h; y += 4) {
        uint32_t pattern = 0xFFFF;
        if (h - y < 4) {
            pattern = 0x7777;
            if (h - y < 3) {
                pattern = 0x3333;
                if (h - y < 2) pattern = 0x1111;
            }
        }
        uint32_t prev = 0;
        const uint32_t cur_row = (y >> 2) * mstr, next_row = cur_row + mstr;
        const uint32_t dpp = y * w;
        for (uint32_t x = 0; x < w; x += 4) {
            uint32_t col_pattern = pattern;
            int s = (int)x + 4 - (int)w;
            if (s < 0) s = 0;
            col_pattern >>= (uint32_t)(s * 4);
            const uint32_t idx = x >> 2;
            const uint32_t ps = (uint32_t)prs[idx] | ((uint32_t)prs[idx + 1] << 16);
            const uint32_t ns = pair16(sigma, next_row + idx);
            uint32_t u = (ps & 0x88888888u) >> 3;
            if (!stripe_causal) u |= (ns & 0x11111111u) << 3;
            const uint32_t cs = pair16(sigma, cur_row + idx);
            uint32_t mbr = cs;
            mbr |= (cs & 0x77777777u) << 1;
            mbr |= (cs & 0xEEEEEEEEu) >> 1;
            mbr |= u;
            const uint32_t t = mbr;
            mbr |= t << 4;
            mbr |= t >> 4;
            mbr |= prev >> 12;
            mbr &= col_pattern;
            mbr &= ~cs;
            uint32_t new_sig = 0;
            if (mbr) {
                uint32_t cwd = fwd_fetch(sp), cnt = 0;
                const uint32_t inv_sig = ~cs & col_pattern;
                uint32_t cand = mbr, processed = 0;
                while (cand) {
                    const uint32_t bit = __ffs(cand) - 1;
                    const uint32_t m = 1u << bit;
                    cand &= ~m;
                    processed |= m;
                    if (cwd & 1) {
                        new_sig |= m;
                        cand |= spread[bit] & inv_sig & ~processed;
                    }
                    cwd >>= 1;
                    ++cnt;
                }
                if (new_sig) {
                    const uint32_t value = 3u << (p - 2);
                    uint32_t sb = new_sig;
                    while (sb) {
                        const uint32_t bit = __ffs(sb) - 1;
                        sb &= ~(1u << bit);
                        const uint32_t off = (bit >> 2) + (bit & 3) * w;
                        blk[dpp + x + off] = ((cwd & 1) << 31) | value;
                        cwd >>= 1;
                        ++cnt;
                    }
                }
                fwd_advance(sp, cnt);
            }
            const uint32_t comb = new_sig | cs;
            prs[idx] = (uint16_t)comb;
            if (idx + 1 < prs_len) prs[idx + 1] = (uint16_t)(comb >> 16);
            uint32_t np = comb;
            np |= (comb & 0x7777) << 1;
            np |= (comb & 0xEEEE) >> 1;
            prev = (np | u) & 0xF000;
        }
    }
}

__device__ uint32_t wave_magref(const uint8_t *ref, uint32_t lref, const uint16_t *sigma,
                                uint32_t *blk, uint32_t w, uint32_t h, uint32_t mstr,
                                uint32_t p) {
    if (p < 2) return 1;
    RevReader mr;
    rev_init_mrp(mr, ref, lref);
    const uint32_t half = 1u << (p - 2);
    for (uint32_t y = 0; y < h; y += 4) {
        uint32_t si = (y >> 2) * mstr;
        const uint32_t dpp = y * w;
        for (uint32_t i = 0; i < w; i += 8) {
            uint32_t cwd = rev_fetch(mr);
            const uint32_t sig = pair16(sigma, si);
            si += 2;
            uint32_t col_mask = 0xF;
            if (sig) {
                for (uint32_t jj = 0; jj < 8; ++jj) {
                    if (sig & col_mask) {
                        uint32_t dp = dpp + i + jj;
                        uint32_t sm = 0x11111111u & col_mask;
                        for (int k = 0; k < 4; ++k) {
                            if (sig & sm) {
                                uint32_t sym = cwd & 1;
                                sym = (1 - sym) << (p - 1);
                                sym |= half;
                                blk[dp] ^= sym;
                                cwd >>= 1;
                            }
                            sm <<= 1;
                            dp += w;
                        }
                    }
                    col_mask <<= 4;
                }
            }
            rev_advance(mr, __popc(sig));
        }
    }
    return 0;
}

extern "C" __global__ void __launch_bounds__(64)
sgn_ht_block_wave(const CbJob *__restrict__ jobs, uint32_t n, const uint8_t *__restrict__ arena,
                  int32_t *__restrict__ coef, const uint16_t *__restrict__ tables,
                  uint8_t *__restrict__ status, uint32_t raw_out, uint32_t narrow) {
    __shared__ uint16_t s_tab[2624];
    __shared__ uint32_t s_blk[kWaveMaxSamples];
    __shared__ uint16_t s_sym[kWaveSymWords];
    __shared__ uint16_t s_sigma[kWaveSigmaWords];
    __shared__ uint32_t s_vn[kWaveVn];
    __shared__ uint16_t s_prs[kWavePrs];
    __shared__ uint32_t s_fail;
    const uint32_t i = blockIdx.x;
    if (i >= n) return;
    const int lane = threadIdx.x;
    load_tables(tables, s_tab, 2624, lane, 64);
    const CbJob j = jobs[i];
    const uint32_t w = j.w, h = j.h;
    const uint32_t sstr = (w + 2 + 7) & ~7u;
    const uint32_t mstr = (((w + 3) / 4) + 2 + 7) & ~7u;
    for (uint32_t k = lane; k < w * h && k < (uint32_t)kWaveMaxSamples; k += 64) s_blk[k] = 0;
    const uint32_t sym_words = sstr * ((h + 1) / 2 + 1);
    for (uint32_t k = lane; k < sym_words && k < (uint32_t)kWaveSymWords; k += 64) s_sym[k] = 0;
    const uint32_t sig_words = mstr * ((h + 3) / 4 + 1) + 1;
    for (uint32_t k = lane; k < sig_words && k < (uint32_t)kWaveSigmaWords; k += 64) s_sigma[k] = 0;
    if (lane == 0) s_fail = 0;
    __syncthreads();
    if (lane == 0) {
        uint32_t fail = 0;
        if (w * h > (uint32_t)kWaveMaxSamples || sym_words > (uint32_t)kWaveSymWords ||
            sig_words > (uint32_t)kWaveSigmaWords || (w + 1) / 2 + 2 > (uint32_t)kWaveVn ||
            (w + 3) / 4 + 8 > (uint32_t)kWavePrs) {
            fail = 1;
        } else if (!(j.flags & kCbZero)) {
            const uint8_t *d = arena + j.src;
            uint32_t passes = j.passes;
            const uint32_t zblk = j.zblk, lcup = j.lcup, lref = j.lref;
            if (passes > 1 && lref == 0) passes = 1;
            if (passes > 3 || zblk >= 30 || lcup < 2) {
                fail = 1;
            } else {
                if (zblk == 29 && passes > 1) passes = 1;
                const uint32_t scup = ((uint32_t)d[lcup - 1] << 4) + (d[lcup - 2] & 0x0F);
                if (scup < 2 || scup > lcup || scup > 4079) {
                    fail = 1;
                } else {
                    wave_cleanup_symbols(d, lcup, scup, w, h, sstr, s_sym, s_tab, s_tab + 1024,
                                         s_tab + 2048, s_tab + 2368);
                    fail = wave_magsgn(d, lcup, scup, s_sym, s_blk, zblk, w, h, sstr, s_vn);
                    if (!fail && passes > 1) {
                        const uint32_t p = 30 - zblk;
                        wave_sigma(s_sym, s_sigma, w, h, sstr, mstr);
                        wave_sigprop(d + lcup, lref, s_sigma, s_blk, w, h, mstr,
                                     (j.flags & kCbStripeCausal) != 0, p, s_prs);
                        if (passes > 2) fail = wave_magref(d + lcup, lref, s_sigma, s_blk, w, h, mstr, p);
                    }
                }
            }
        }
        s_fail = fail;
        status[i] = fail ? kCbBad : kCbOk;
    }
    __syncthreads();
    const bool bad = s_fail != 0;
    int32_t *out = coef + j.dst;
    int16_t *out16 = reinterpret_cast<int16_t *>(coef) + j.dst;  // 16-bit coefficient arena
    for (uint32_t k = lane; k < w * h; k += 64) {
        const uint32_t y = k / w, x = k - y * w;
        uint32_t v = (bad || k >= (uint32_t)kWaveMaxSamples) ? 0u : s_blk[k];
        const int32_t c = raw_out ? (int32_t)v : to_i32(v, j.kmax);
        if (narrow)
            out16[y * j.stride + x] = (int16_t)c;
        else
            out[y * j.stride + x] = c;
    }
}

}  // namespace sgn
