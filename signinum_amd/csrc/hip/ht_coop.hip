// HT code-block decode with one code-block per wavefront, the wave's 64
// lanes cooperating on every pass, gfx950.
//
// Cleanup (crates/signinum-j2k-native/src/j2c/ht_block_decode.rs,
// decode_cleanup_symbols :1056-1204 and decode_magnitude_sign_phase
// :1435-1560), one quad row at a time:
//   * the tail of the cleanup segment (MEL and VLC streams, at most 4079
//     bytes) is staged in LDS with LDS-DMA loads at the start; lane 0 walks
//     the MEL/VLC/UVLC chain of the quad row (a serial chain: each
//     codeword's context is its left neighbour's significance) from LDS and
//     writes the row's quad symbols (significance, e_1, e_k, u_q) to LDS;
//   * then every lane takes one column of the quad row (two samples): the
//     exponent bound U_q from the bottom samples of the row above (LDS), the
//     MagSgn bit count of its samples, and the column's bit offset in the
//     MagSgn stream as a wave prefix sum of the counts (the stream holds the
//     quads' samples in column order, so lane order is stream order);
//   * the MagSgn stream is unstuffed 64 bytes at a time by the whole wave
//     (each byte's bit count depends only on its predecessor, so positions
//     are a prefix sum too) into an LDS ring of words that any lane reads at
//     any bit offset.
// SigProp (:1254-1374) is serial by nature (a sample's context includes the
// samples decoded before it in the same pass) and runs on lane 0 over LDS
// significance maps.  MagRef (:1377-1432) takes one bit per sample that the
// cleanup made significant, in stripe-column order: its bits are unstuffed
// into LDS by the whole wave and each lane finds its column's offset by a
// prefix sum of significance counts.
//
// The kernel writes coefficient_to_i32 integers (int16 in narrow plans) or
// the raw sign-magnitude words into the coefficient arena.
#include <hip/hip_runtime.h>

#include "../host/plan_types.h"
#include "ht_common.hpp"
#include "ht_readers4.hpp"

namespace sgn {

constexpr uint32_t kCoopStageUnits = 264;   // 16-byte units of a cleanup segment's tail in LDS
constexpr uint32_t kCoopRingWords = 512;    // unstuffed MagSgn ring, 32-bit words
constexpr uint32_t kCoopMaxSamples = 4096;  // w * h of a code-block (T.800 B.7)
constexpr uint32_t kCoopRefWords = 520;     // unstuffed MagRef bits: Lref <= 2047 bytes
constexpr uint32_t kCoopRhoBytes = 2056;    // (w/2 + 2) * (h/2 + 1) quad significance nibbles

// Staged tail of the cleanup segment: arena byte addresses [lo, hi)
// (16-byte aligned) live in LDS at `lds`; addresses below lo come from
// memory.  Word source of the serial readers.
struct CoopSrc {
    const uint64_t *lds;
    uint32_t lo, hi;
    const uint8_t *arena;
    __device__ __forceinline__ uint64_t word(uint32_t a) const {
        if (a >= lo && a < hi) return lds[(a - lo) >> 3];
        return *reinterpret_cast<const uint64_t *>(arena + a);
    }
    __device__ __forceinline__ uint32_t byte(uint32_t a) const {
        if (a >= lo && a < hi) return reinterpret_cast<const uint8_t *>(lds)[a - lo];
        return arena[a];
    }
};

// DPP move of v from another lane (CTRL), 0 where the source lane is out
// of range or the row is not in ROWS.
template <int CTRL, int ROWS = 0xF>
__device__ __forceinline__ uint32_t dpp0(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, ROWS, 0xF, false);
}

// Exclusive prefix sum over the 64 lanes (lane order); *total gets the sum.
// Inclusive scan inside each 16-lane row by row shifts, then the rows'
// totals by row broadcasts (lane 15 into rows 1 and 3, lane 31 into rows 2
// and 3).  Called by all lanes.
__device__ __forceinline__ uint32_t wave_scan_excl(uint32_t v, uint32_t &total) {
    uint32_t s = v;
    s += dpp0<0x111>(s);        // row_shr:1
    s += dpp0<0x112>(s);        // row_shr:2
    s += dpp0<0x114>(s);        // row_shr:4
    s += dpp0<0x118>(s);        // row_shr:8
    s += dpp0<0x142, 0xA>(s);   // row_bcast:15
    s += dpp0<0x143, 0xC>(s);   // row_bcast:31
    total = (uint32_t)__builtin_amdgcn_readlane((int)s, 63);
    return s - v;
}

// Lanes below this one whose flag is set (ballot + mbcnt).
__device__ __forceinline__ uint32_t count_below(bool flag, uint64_t &mask) {
    mask = __ballot(flag);
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                     __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

// 32 bits of an unstuffed stream held in LDS words (LSB-first) from bit
// position pos; the buffer is a ring of MASK + 1 words.
template <uint32_t MASK>
__device__ __forceinline__ uint32_t lds_bits(const uint32_t *buf, uint32_t pos) {
    const uint32_t wi = pos >> 5;
    const uint32_t lo = buf[wi & MASK], hi = buf[(wi + 1) & MASK];
    return __builtin_amdgcn_alignbit(hi, lo, pos & 31u);
}

// MagSgn stream being unstuffed into the LDS ring (uniform across lanes).
struct MsRing {
    uint32_t next;     // arena address of the next raw byte
    uint32_t end;      // arena address past the MagSgn segment
    uint32_t fill;     // unstuffed bits written so far
    uint32_t prev_ff;  // the last raw byte was 0xFF
};

// Unstuffs MagSgn bytes, 64 per round, until `need` bits are in the ring.
// Forward stream (ht_block_decode.rs:881-929): LSB-first, a byte after 0xFF
// carries 7 bits, 0xFF past the segment.  Every byte is ORed in whole at its
// position, as the serial reader does (a 7-bit byte's top bit lands on the
// next byte's first bit).  Called by all lanes.
__device__ __forceinline__ void ms_ring_fill(MsRing &r, uint32_t need, uint32_t *ring,
                                             const CoopSrc &src) {
    const uint32_t lane = __lane_id();
    while (r.fill < need) {
        const uint32_t a = r.next + lane;
        const uint32_t b = a < r.end ? src.byte(a) : 0xFFu;
        const uint32_t up = dpp0<0x138>(b);  // wave_shr:1, the previous byte
        const bool pff = lane == 0 ? r.prev_ff != 0 : up == 0xFFu;
        // each byte carries 8 bits, 7 after a 0xFF
        uint64_t mask;
        const uint32_t off = 8u * lane - count_below(pff, mask);
        const uint32_t tot = 512u - (uint32_t)__builtin_popcountll(mask);
        // words past the partly written one that this 64-byte step reaches start at zero
        const uint32_t w0 = r.fill >> 5, wl = (r.fill + tot) >> 5;
        if (lane >= 1 && w0 + lane <= wl) ring[(w0 + lane) & (kCoopRingWords - 1)] = 0;
        __syncthreads();
        const uint32_t pos = r.fill + off, sh = pos & 31u, wi = pos >> 5;
        atomicOr(&ring[wi & (kCoopRingWords - 1)], b << sh);
        if (sh > 24) atomicOr(&ring[(wi + 1) & (kCoopRingWords - 1)], b >> (32 - sh));
        __syncthreads();
        r.prev_ff = __builtin_amdgcn_readlane((int)b, 63) == 0xFF ? 1u : 0u;
        r.next += 64;
        r.fill += tot;
    }
}

// Serial state of the MEL / VLC / UVLC chain (lane 0).
struct CoopChain {
    Fwd4 fmel;
    Rev4 rvlc;
    VlcBits vlc;
    Mel mel;
    int run;
};

// Quad symbols of quad row `row` (ht_block_decode.rs:1056-1204), lane 0
// only: sym[q] = t_q | u_q << 16 for the row's quads, two zero entries
// past the last pair; prev holds the row above.
__device__ __forceinline__ void coop_row_symbols(CoopChain &c, const CoopSrc &src, uint32_t row,
                                                 uint32_t w, const uint16_t *tab, uint32_t *sym,
                                                 const uint32_t *prev) {
    const uint16_t *vlc0 = tab, *vlc1 = tab + 1024, *uvlc0 = tab + 2048, *uvlc1 = tab + 2368;
    const uint32_t npairs = (w + 3) >> 2;
    uint32_t c_q = 0;
    for (uint32_t pr = 0; pr < npairs; ++pr) {
        const uint32_t x = 4 * pr, qj = 2 * pr;
        uint32_t t0, t1, u0, u1;
        vlc_fill(c.vlc, c.rvlc, src);
        if (row == 0) {
            t0 = vlc0[c_q + (c.vlc.tmp & 0x7F)];
            if (c_q == 0) {
                c.run -= 2;
                t0 = (c.run == -1) ? t0 : 0;
                if (c.run < 0) c.run = mel_run4(c.mel, c.fmel, src);
            }
            c_q = ((t0 & 0x10) << 3) | ((t0 & 0xE0) << 2);
            c.vlc.tmp >>= t0 & 7;
            c.vlc.bits -= t0 & 7;
            t1 = vlc0[c_q + (c.vlc.tmp & 0x7F)];
            if (c_q == 0 && x + 2 < w) {
                c.run -= 2;
                t1 = (c.run == -1) ? t1 : 0;
                if (c.run < 0) c.run = mel_run4(c.mel, c.fmel, src);
            }
            if (x + 2 >= w) t1 = 0;
            c_q = ((t1 & 0x10) << 3) | ((t1 & 0xE0) << 2);
            c.vlc.tmp >>= t1 & 7;
            c.vlc.bits -= t1 & 7;
            uint32_t mode = ((t0 & 8) << 3) | ((t1 & 8) << 4);
            if (mode == 0xC0) {
                c.run -= 2;
                if (c.run == -1) mode += 0x40;
                if (c.run < 0) c.run = mel_run4(c.mel, c.fmel, src);
            }
            uint32_t e = uvlc0[mode + (c.vlc.tmp & 0x3F)];
            c.vlc.tmp >>= e & 7;
            c.vlc.bits -= e & 7;
            e >>= 3;
            uint32_t len = e & 0xF;
            const uint32_t tmp = (uint32_t)c.vlc.tmp & ((1u << len) - 1);
            c.vlc.tmp >>= len;
            c.vlc.bits -= len;
            e >>= 4;
            len = e & 7;
            e >>= 3;
            u0 = 1 + (e & 7) + (tmp & ~(0xFFu << len));
            u1 = 1 + (e >> 3) + (tmp >> len);
        } else {
            const uint32_t pt0 = prev[qj], pt1 = prev[qj + 1], pt2 = prev[qj + 2];
            c_q |= (pt0 & 0xA0) << 2;
            c_q |= (pt1 & 0x20) << 4;
            t0 = vlc1[c_q + (c.vlc.tmp & 0x7F)];
            if (c_q == 0) {
                c.run -= 2;
                t0 = (c.run == -1) ? t0 : 0;
                if (c.run < 0) c.run = mel_run4(c.mel, c.fmel, src);
            }
            c_q = ((t0 & 0x40) << 2) | ((t0 & 0x80) << 1);
            c_q |= pt0 & 0x80;
            c_q |= (pt1 & 0xA0) << 2;
            c_q |= (pt2 & 0x20) << 4;
            c.vlc.tmp >>= t0 & 7;
            c.vlc.bits -= t0 & 7;
            t1 = vlc1[c_q + (c.vlc.tmp & 0x7F)];
            if (c_q == 0 && x + 2 < w) {
                c.run -= 2;
                t1 = (c.run == -1) ? t1 : 0;
                if (c.run < 0) c.run = mel_run4(c.mel, c.fmel, src);
            }
            if (x + 2 >= w) t1 = 0;
            c_q = ((t1 & 0x40) << 2) | ((t1 & 0x80) << 1);
            c_q |= pt1 & 0x80;
            c.vlc.tmp >>= t1 & 7;
            c.vlc.bits -= t1 & 7;
            const uint32_t mode = ((t0 & 8) << 3) | ((t1 & 8) << 4);
            uint32_t e = uvlc1[mode + (c.vlc.tmp & 0x3F)];
            c.vlc.tmp >>= e & 7;
            c.vlc.bits -= e & 7;
            e >>= 3;
            uint32_t len = e & 0xF;
            const uint32_t tmp = (uint32_t)c.vlc.tmp & ((1u << len) - 1);
            c.vlc.tmp >>= len;
            c.vlc.bits -= len;
            e >>= 4;
            len = e & 7;
            e >>= 3;
            u0 = (e & 7) + (tmp & ~(0xFFu << len));
            u1 = (e >> 3) + (tmp >> len);
        }
        sym[qj] = t0 | (u0 << 16);
        sym[qj + 1] = t1 | (u1 << 16);
    }
    sym[2 * npairs] = 0;
    sym[2 * npairs + 1] = 0;
}

// One MagSgn sample (ht_block_decode.rs:1010-1032) from the 32 stream bits
// at its offset: the sign-magnitude word and v_n.
__device__ __forceinline__ uint32_t coop_sample(uint32_t bits, uint32_t sig, uint32_t m,
                                                uint32_t ibit, uint32_t p, uint32_t &vn) {
    if (!sig) {
        vn = 0;
        return 0;
    }
    uint32_t v = bits & (m >= 32 ? 0xFFFFFFFFu : ((1u << m) - 1u));
    v |= ibit << m;
    v |= 1u;
    vn = v;
    return (bits << 31) | ((v + 2) << (p - 1));
}

// LDS-DMA of the 16-byte arena units [lo_u, hi_u) (at most
// kCoopStageUnits) into stage; the caller waits on vmcnt.
__device__ __forceinline__ void coop_stage(uint4 *stage, const uint8_t *arena, uint32_t lo_u,
                                           uint32_t hi_u) {
    const uint32_t lane = __lane_id();
    const uint4 *g = reinterpret_cast<const uint4 *>(arena);
    for (uint32_t k = 0; k < kCoopStageUnits; k += 64)
        if (lo_u + k + lane < hi_u)
            __builtin_amdgcn_global_load_lds(g + lo_u + k + lane, stage + k, 16, 0, 0);
}

template <typename C>
__device__ __forceinline__ C coop_out(uint32_t v, uint32_t kmax, bool raw) {
    if (raw) return (C)(int32_t)v;
    return (C)to_i32(v, kmax);
}

// LDS of one wave.  WMAX: widest code-block; REFINE: SigProp / MagRef
// state and the block itself (cleanup-only blocks go straight to memory).
template <uint32_t WMAX, bool REFINE>
struct CoopLds {
    uint16_t tab[2624];
    uint4 stage[kCoopStageUnits];
    uint32_t ring[kCoopRingWords];
    uint32_t sym[2][WMAX / 2 + 4];
    uint8_t erow[2][WMAX + 4];  // exponent bound (floor log2 v_n) of the bottom samples
    uint32_t fail;
    // REFINE only (sizes 1 otherwise)
    uint32_t blk[REFINE ? kCoopMaxSamples : 1];
    uint8_t rho[REFINE ? kCoopRhoBytes : 1];
    uint16_t sigma[REFINE ? 2200 : 1];
    uint16_t prs[REFINE ? 272 : 1];
    uint32_t ref[REFINE ? kCoopRefWords : 1];
};

// SigProp (lane 0) and MagRef (whole wave) over the cleanup result in
// s.blk; returns nonzero on a malformed block.  Called by all lanes.
template <uint32_t WMAX>
__device__ uint32_t coop_refine(CoopLds<WMAX, true> &s, const uint8_t *arena, const CbJob &j,
                                uint32_t passes, uint32_t w, uint32_t h) {
    const uint32_t lane = __lane_id();
    const uint32_t p = 30 - j.zblk;
    const uint32_t mstr = (((w + 3) / 4) + 2 + 7) & ~7u;
    const uint32_t nq = (w + 1) >> 1, nqr = (h + 1) >> 1;
    // significance map of the cleanup (ht_block_decode.rs:1207-1251):
    // 16 bits per 4x4 group, bit = column * 4 + row
    const uint32_t ngy = (h + 3) / 4;
    const uint32_t sig_words = mstr * (ngy + 1) + 1;
    for (uint32_t k = lane; k < sig_words; k += 64) s.sigma[k] = 0;
    __syncthreads();
    const uint32_t ngx = (w + 3) / 4, rs = nq + 2;
    for (uint32_t g = lane; g < ngx * ngy; g += 64) {
        const uint32_t gy = g / ngx, gx = g - gy * ngx;
        const uint32_t qr = 2 * gy, qc = 2 * gx;
        auto rho = [&](uint32_t r, uint32_t q) -> uint32_t {
            return (r < nqr + 1 && q < rs) ? (uint32_t)s.rho[r * rs + q] : 0u;
        };
        const uint32_t a = rho(qr, qc), b = rho(qr, qc + 1), c = rho(qr + 1, qc),
                       d = rho(qr + 1, qc + 1);
        // quads a | b over c | d; a quad's nibble is (top, bottom) of its
        // left column then of its right column
        uint32_t t = (a & 3) | ((a & 0xC) << 2) | ((b & 3) << 8) | ((b & 0xC) << 10);
        t |= ((c & 3) << 2) | ((c & 0xC) << 4) | ((d & 3) << 10) | ((d & 0xC) << 12);
        s.sigma[gy * mstr + gx] = (uint16_t)t;
    }
    __syncthreads();
    if (lane == 0)
        wave_sigprop(arena + j.src + j.lcup, j.lref, s.sigma, s.blk, w, h, mstr,
                     (j.flags & kCbStripeCausal) != 0, p, s.prs);
    __syncthreads();
    if (passes < 3) return 0;
    if (p < 2 || j.lref > 2047) return 1;
    // MagRef bits (ht_block_decode.rs:931-999, 1377-1432): backward from the
    // segment's last byte, a byte whose low 7 bits are ones after a byte
    // > 0x8F carries 7 bits, zeros past the start; unstuffed into s.ref
    const uint8_t *ref = arena + j.src + j.lcup;
    const uint32_t lref = j.lref;
    for (uint32_t k = lane; k < kCoopRefWords; k += 64) s.ref[k] = 0;
    __syncthreads();
    uint32_t fill = 0, prev_big = 1;
    for (uint32_t k0 = 0; k0 < lref; k0 += 64) {
        const uint32_t k = k0 + lane;
        const uint32_t b = k < lref ? (uint32_t)ref[lref - 1 - k] : 0u;
        const uint32_t up = dpp0<0x138>(b);  // wave_shr:1, the byte read before
        const bool unst = lane == 0 ? prev_big != 0 : up > 0x8Fu;
        // 8 bits, 7 for a byte with its low 7 bits set after a byte > 0x8F
        uint64_t mask;
        const uint32_t pos = fill + 8u * lane - count_below(unst && (b & 0x7F) == 0x7F, mask);
        const uint32_t tot = 512u - (uint32_t)__builtin_popcountll(mask);
        const uint32_t sh = pos & 31u, wi = pos >> 5;
        if (b) {
            atomicOr(&s.ref[wi], b << sh);
            if (sh > 24) atomicOr(&s.ref[wi + 1], b >> (32 - sh));
        }
        prev_big = (uint32_t)__builtin_amdgcn_readlane((int)b, 63) > 0x8Fu ? 1u : 0u;
        fill += tot;
    }
    __syncthreads();
    const uint32_t half = 1u << (p - 2);
    uint32_t mpos = 0;
    for (uint32_t y = 0; y < h; y += 4) {
        for (uint32_t c0 = 0; c0 < w; c0 += 64) {
            const uint32_t x = c0 + lane;
            uint32_t nib = 0;
            if (x < ((w + 7) & ~7u))
                nib = ((uint32_t)s.sigma[(y >> 2) * mstr + (x >> 2)] >> ((x & 3) * 4)) & 0xFu;
            uint32_t tot;
            const uint32_t off = wave_scan_excl((uint32_t)__builtin_popcount(nib), tot);
            // past the unstuffed bits the stream reads as zeros
            uint32_t bits = 0;
            if (nib) {
                const uint32_t at = mpos + off, wi = at >> 5;
                const uint32_t lo = wi < kCoopRefWords ? s.ref[wi] : 0u;
                const uint32_t hi = wi + 1 < kCoopRefWords ? s.ref[wi + 1] : 0u;
                bits = __builtin_amdgcn_alignbit(hi, lo, at & 31u);
            }
#pragma unroll
            for (uint32_t k = 0; k < 4; ++k) {
                if (nib & (1u << k)) {
                    const uint32_t sym = ((1u - (bits & 1u)) << (p - 1)) | half;
                    bits >>= 1;
                    if (x < w && y + k < h) s.blk[(y + k) * w + x] ^= sym;
                }
            }
            mpos += tot;
        }
    }
    __syncthreads();
    return 0;
}

// One code-block per wave: jobs[blockIdx.x].  C: int32 or int16 samples;
// raw_out: store the sign-magnitude words (C must be int32).
template <uint32_t WMAX, bool REFINE, typename C>
__global__ void __launch_bounds__(64)
sgn_ht_coop(const CbJob *__restrict__ jobs, uint32_t n, const uint8_t *__restrict__ arena,
            C *__restrict__ coef, const uint16_t *__restrict__ tables,
            uint8_t *__restrict__ status, uint32_t raw_out) {
    __shared__ CoopLds<WMAX, REFINE> s;
    const uint32_t i = blockIdx.x;
    if (i >= n) return;
    const uint32_t lane = __lane_id();
    const CbJob j = jobs[i];
    const uint32_t w = j.w, h = j.h, stride = j.stride;
    C *out = coef + j.dst;
    const bool raw = raw_out != 0;
    uint32_t passes = j.passes;
    uint32_t bad = 0, run_block = 1;
    const uint32_t lcup = j.lcup, zblk = j.zblk;
    if (w > WMAX || w * h > kCoopMaxSamples) {
        bad = 1;
    } else if ((j.flags & kCbZero) || w == 0 || h == 0) {
        run_block = 0;
    } else {
        // pass count rules of decode_impl (ht_block_decode.rs:1563-1680)
        if (passes > 1 && j.lref == 0) passes = 1;
        if (passes > 3 || zblk >= 30 || lcup < 2) {
            bad = 1;
        } else {
            if (zblk == 29 && passes > 1) passes = 1;
            // the cleanup-only kernel is never given refinement passes
            if (!REFINE && passes > 1) bad = 1;
        }
    }
    // stage the tail of the cleanup segment (up to kCoopStageUnits units)
    const uint32_t seg0 = j.src, seg1 = j.src + lcup;
    uint32_t lo_u = 0, hi_u = 0;
    if (!bad && run_block) {
        hi_u = ((seg1 - 1) >> 4) + 1;
        lo_u = seg0 >> 4;
        if (hi_u - lo_u > kCoopStageUnits) lo_u = hi_u - kCoopStageUnits;
        coop_stage(s.stage, arena, lo_u, hi_u);
    }
    for (uint32_t t = lane; t < 2624; t += 64) s.tab[t] = tables[t];
    for (uint32_t t = lane; t < 2 * (WMAX + 4); t += 64) (&s.erow[0][0])[t] = 0;
    if constexpr (REFINE) {
        for (uint32_t t = lane; t < kCoopRhoBytes; t += 64) s.rho[t] = 0;
        for (uint32_t t = lane; t < w * h && t < kCoopMaxSamples; t += 64) s.blk[t] = 0;
    }
    if (lane == 0) s.ring[0] = 0;
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    __syncthreads();
    const CoopSrc src{reinterpret_cast<const uint64_t *>(s.stage), lo_u << 4, hi_u << 4, arena};
    uint32_t scup = 0;
    if (!bad && run_block) {
        scup = (src.byte(seg1 - 1) << 4) + (src.byte(seg1 - 2) & 0x0F);
        if (scup < 2 || scup > lcup || scup > 4079) bad = 1;
    }
    if (!bad && run_block) {
        const uint32_t p = 30 - zblk, mmsbp2 = zblk + 2;
        const uint32_t kmax = j.kmax;
        CoopChain ch;
        if (lane == 0) {
            const uint32_t mel_off = seg1 - scup, mel_len = scup - 1;
            fwd4_init(ch.fmel, mel_off, mel_len, 0xFF, src);
            ch.mel = mel_make(mel_len);
            rev4_init(ch.rvlc, seg1 - 3, scup - 2, src);
            const uint32_t b0 = src.byte(seg1 - 2);
            ch.vlc.tmp = b0 >> 4;
            ch.vlc.bits = 4 - (((ch.vlc.tmp & 7) == 7) ? 1u : 0u);
            ch.vlc.unstuff = ((b0 | 0x0F) > 0x8F) ? 1u : 0u;
            ch.run = mel_run4(ch.mel, ch.fmel, src);
        }
        MsRing ms{seg0, seg0 + (lcup - scup), 0, 0};
        uint32_t pos = 0;
        const uint32_t nrows = (h + 1) >> 1;
        for (uint32_t r = 0; r < nrows; ++r) {
            uint32_t *sym = s.sym[r & 1];
            const uint32_t *prev = s.sym[(r + 1) & 1];
            if (lane == 0) coop_row_symbols(ch, src, r, w, s.tab, sym, prev);
            __syncthreads();
            const uint8_t *eprev = s.erow[(r + 1) & 1];
            uint8_t *ecur = s.erow[r & 1];
            const uint32_t y = 2 * r;
            const bool second = y + 1 < h;
            uint32_t row_bad = 0;
            for (uint32_t c0 = 0; c0 < w; c0 += 64) {
                const uint32_t x = c0 + lane, q = x >> 1;
                const bool col = x < w;
                const uint32_t sq = col ? sym[q] : 0u;
                const uint32_t inf = sq & 0xFFFFu;
                uint32_t uq = sq >> 16;
                if (r > 0) {
                    uint32_t gamma = inf & 0xF0;
                    gamma &= gamma - 0x10;
                    uint32_t emax = max(max((uint32_t)eprev[2 * q], (uint32_t)eprev[2 * q + 1]),
                                        max((uint32_t)eprev[2 * q + 2], (uint32_t)eprev[2 * q + 3]));
                    uq += gamma ? max(emax, 1u) : 1u;
                }
                if (col && (x & 1u) == 0 && uq > mmsbp2) row_bad = 1;
                const uint32_t n0 = (x & 1u) * 2;  // this column's samples: n0 (top), n0 + 1
                const uint32_t s0 = col ? (inf >> (4 + n0)) & 1u : 0u;
                const uint32_t s1 = col ? (inf >> (5 + n0)) & 1u : 0u;
                const uint32_t m0 = s0 ? uq - ((inf >> (12 + n0)) & 1u) : 0u;
                const uint32_t m1 = s1 ? uq - ((inf >> (13 + n0)) & 1u) : 0u;
                if (__any(row_bad)) {
                    row_bad = 1;
                    break;
                }
                uint32_t tot;
                const uint32_t off = wave_scan_excl(m0 + m1, tot);
                // every bit this chunk reads, plus the sign bit a zero-length
                // sample reads from the position after the last one
                ms_ring_fill(ms, pos + tot + 1, s.ring, src);
                uint32_t v0, v1;
                const uint32_t b0 = lds_bits<kCoopRingWords - 1>(s.ring, pos + off);
                const uint32_t b1 = lds_bits<kCoopRingWords - 1>(s.ring, pos + off + m0);
                const uint32_t a0 = coop_sample(b0, s0, m0, (inf >> (8 + n0)) & 1u, p, v0);
                const uint32_t a1 = coop_sample(b1, s1, m1, (inf >> (9 + n0)) & 1u, p, v1);
                if (col) {
                    if constexpr (REFINE) {
                        s.blk[y * w + x] = a0;
                        if (second) s.blk[(y + 1) * w + x] = a1;
                        if ((x & 1u) == 0) s.rho[r * ((w + 1) / 2 + 2) + q] = (uint8_t)((inf >> 4) & 0xF);
                    } else {
                        out[y * stride + x] = coop_out<C>(a0, kmax, raw);
                        if (second) out[(y + 1) * stride + x] = coop_out<C>(a1, kmax, raw);
                    }
                }
                ecur[x + 1] = col ? (uint8_t)flog2r(v1) : 0;
                pos += tot;
            }
            __syncthreads();
            if (row_bad) {
                bad = 1;
                break;
            }
        }
        if constexpr (REFINE) {
            if (!bad && passes > 1) bad = coop_refine<WMAX>(s, arena, j, passes, w, h);
            if (!bad)
                for (uint32_t k = lane; k < w * h; k += 64) {
                    const uint32_t yy = k / w, xx = k - yy * w;
                    out[yy * stride + xx] = coop_out<C>(s.blk[k], kmax, raw);
                }
        }
    }
    if (bad || !run_block) {
        for (uint32_t k = lane; k < w * h; k += 64) {
            const uint32_t yy = k / w, xx = k - yy * w;
            out[yy * stride + xx] = (C)0;
        }
    }
    if (lane == 0) status[i] = bad ? kCbBad : kCbOk;
}

}  // namespace sgn
