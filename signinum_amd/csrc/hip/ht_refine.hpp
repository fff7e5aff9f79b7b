// HT SigProp and MagRef passes of one code-block on one lane, over the
// cleanup samples already in the coefficient arena
// (crates/signinum-j2k-native/src/j2c/ht_block_decode.rs :1254-1432).
//
// The including file provides __device__, __forceinline__ and the HIP
// vector types.
#pragma once
#include <stdint.h>

#include "ht_readers4.hpp"

namespace sgn {

// Word source of the refinement readers: the arena in memory.
struct MemSrc {
    const uint8_t *arena;
    __device__ __forceinline__ uint64_t word(uint32_t a) const {
        return *reinterpret_cast<const uint64_t *>(arena + a);
    }
};

// A 4x4 group of a stripe as the passes index it: bit 4 * column + row.
using Group = int32_t[4][4];

// Bits of row r of a 16-bit group mask, one per column.
__device__ __forceinline__ uint32_t group_row(uint32_t s, uint32_t r) {
    return ((s >> r) & 1u) | ((s >> (r + 3)) & 2u) | ((s >> (r + 6)) & 4u) | ((s >> (r + 9)) & 8u);
}

// Column bits (bit c) of up to 8 columns to row-0 positions (bit 4c).
__device__ __forceinline__ uint32_t cols_to_row0(uint32_t m) {
    uint32_t r = 0;
#pragma unroll
    for (uint32_t c = 0; c < 8; ++c) r |= ((m >> c) & 1u) << (4 * c);
    return r;
}

// Samples x .. x+3 of block row y; zeros past the block.
template <typename C>
__device__ __forceinline__ void refine_load(int32_t (&v)[4], const C *out, uint32_t y, uint32_t x,
                                            uint32_t w, uint32_t h, uint32_t stride, bool vec) {
#pragma unroll
    for (int c = 0; c < 4; ++c) v[c] = 0;
    if (y >= h || x >= w) return;
    const C *r = out + y * stride + x;
    if (vec && x + 4 <= w) {
        if constexpr (sizeof(C) == 2) {
            const uint2 q = *reinterpret_cast<const uint2 *>(r);
            v[0] = (int16_t)(q.x & 0xFFFFu);
            v[1] = (int16_t)(q.x >> 16);
            v[2] = (int16_t)(q.y & 0xFFFFu);
            v[3] = (int16_t)(q.y >> 16);
        } else {
            const int4 q = *reinterpret_cast<const int4 *>(r);
            v[0] = q.x;
            v[1] = q.y;
            v[2] = q.z;
            v[3] = q.w;
        }
        return;
    }
#pragma unroll
    for (uint32_t c = 0; c < 4; ++c)
        if (x + c < w) v[c] = (int32_t)r[c];
}

template <typename C>
__device__ __forceinline__ void refine_store(const int32_t (&v)[4], C *out, uint32_t y, uint32_t x,
                                             uint32_t w, uint32_t h, uint32_t stride, bool vec) {
    if (y >= h) return;
    C *r = out + y * stride + x;
    if (vec && x + 4 <= w) {
        if constexpr (sizeof(C) == 2) {
            *reinterpret_cast<uint2 *>(r) =
                make_uint2(((uint32_t)v[0] & 0xFFFFu) | ((uint32_t)v[1] << 16),
                           ((uint32_t)v[2] & 0xFFFFu) | ((uint32_t)v[3] << 16));
        } else {
            *reinterpret_cast<int4 *>(r) = make_int4(v[0], v[1], v[2], v[3]);
        }
        return;
    }
#pragma unroll
    for (uint32_t c = 0; c < 4; ++c)
        if (x + c < w) r[c] = (C)v[c];
}

// Group at column x of the stripe at row y, and (unless the context is
// stripe-causal) row 0 of the stripe below it.
template <typename C>
__device__ __forceinline__ void refine_group(Group &g, int32_t (&below)[4], const C *out,
                                             uint32_t y, uint32_t x, uint32_t w, uint32_t h,
                                             uint32_t stride, bool vec, bool causal) {
#pragma unroll
    for (uint32_t r = 0; r < 4; ++r) refine_load<C>(g[r], out, y + r, x, w, h, stride, vec);
    if (causal) {
#pragma unroll
        for (int c = 0; c < 4; ++c) below[c] = 0;
    } else {
        refine_load<C>(below, out, y + 4, x, w, h, stride, vec);
    }
}

__device__ __forceinline__ uint32_t group_sig(const Group &g) {
    uint32_t s = 0;
#pragma unroll
    for (uint32_t r = 0; r < 4; ++r)
#pragma unroll
        for (uint32_t c = 0; c < 4; ++c) s |= (g[r][c] != 0 ? 1u : 0u) << (4 * c + r);
    return s;
}

__device__ __forceinline__ uint32_t row_sig(const int32_t (&v)[4]) {
    return (v[0] != 0 ? 1u : 0u) | (v[1] != 0 ? 2u : 0u) | (v[2] != 0 ? 4u : 0u) |
           (v[3] != 0 ? 8u : 0u);
}

// The SigProp and MagRef passes of one code-block on the calling lane
// (ht_block_decode.rs:1254-1432), over the cleanup samples in `out`.
//
// After the cleanup a sample is significant exactly when it is nonzero: a
// significant sample's sign-magnitude word is at least 3 << (p - 1), and a
// block holds at most K - Zblk planes, so p >= 31 - K = shift and the sample
// keeps at least one bit.  SigProp gives a sample that becomes significant
// the word 3 << (p - 2); MagRef exchanges bit p - 1 for the refinement bit
// and sets 1 << (p - 2) (the midpoint).  Both act on the magnitude bits,
// which the shift to the sample commutes with, so they are applied to the
// stored samples directly.
//
// The context of a group of 4 columns: its own and the next group's cleanup
// significance, the significance of the row above as the reference's
// prev_row_sig holds it when the group reads it, row 0 of the stripe below
// (cleanup significance, unless stripe-causal) and the last column of the
// group before.  prev_row_sig[g + 1] is written with the cleanup
// significance of group g + 1 of the current stripe while group g is
// decoded (ht_block_decode.rs:1360-1363), so group g >= 1 reads the cleanup
// significance of its own stripe's row 3 as the row above, group 0 the
// previous stripe's final row 3, and the next group's part is the previous
// stripe's final row 3.  `bot` keeps that final row 3, one bit per column.
template <typename C>
__device__ void rows_refine(C *out, uint32_t w, uint32_t h, uint32_t stride, bool vec,
                               const uint8_t *arena, uint32_t ref_off, uint32_t lref,
                               uint32_t passes, bool causal, uint32_t p, uint32_t shift) {
    const MemSrc src{arena};
    // SigProp: forward, LSB-first, 7 bits after 0xFF, zeros past the end
    Fwd4 fsp;
    fwd4_init(fsp, ref_off, lref, 0, src);
    MsBits sp{0, 0, 0};
    // MagRef: backward from the segment's last byte (ht_block_decode.rs:931-999)
    Rev4 rmr;
    rev4_init(rmr, ref_off + lref - 1, lref, src);
    VlcBits mr{0, 0, 1};
    const int32_t vnew = (int32_t)((3u << (p - 2)) >> shift);
    const uint32_t ref0 = ((1u << (p - 1)) | (1u << (p - 2))) >> shift;  // MagRef bit 0
    const uint32_t ref1 = (1u << (p - 2)) >> shift;                      // MagRef bit 1
    uint64_t bot = 0;
    for (uint32_t y = 0; y < h; y += 4) {
        const uint32_t rows = min(4u, h - y);
        const uint32_t pattern = 0x1111u * ((1u << rows) - 1u);
        Group a, b;
        int32_t na[4], nb[4];
        refine_group<C>(a, na, out, y, 0, w, h, stride, vec, causal);
        refine_group<C>(b, nb, out, y, 4, w, h, stride, vec, causal);
        uint32_t prev = 0;
        uint64_t nbot = 0;
        for (uint32_t x = 0; x < w; x += 4) {
            const int s = max(0, (int)x + 4 - (int)w);
            const uint32_t col_pattern = pattern >> (4 * s);
            const uint32_t csa = group_sig(a);
            const uint32_t cs = csa | (group_sig(b) << 16);
            const uint32_t lo = x == 0 ? (uint32_t)bot & 0xFu : group_row(csa, 3);
            const uint32_t hi = x + 4 < 64 ? (uint32_t)(bot >> (x + 4)) & 0xFu : 0u;
            uint32_t u = cols_to_row0(lo | (hi << 4));
            if (!causal) u |= cols_to_row0(row_sig(na) | (row_sig(nb) << 4)) << 3;
            uint32_t mbr = cs | ((cs & 0x77777777u) << 1) | ((cs & 0xEEEEEEEEu) >> 1) | u;
            const uint32_t t = mbr;
            mbr |= (t << 4) | (t >> 4) | (prev >> 12);
            mbr &= col_pattern & ~cs;
            uint32_t nsig = 0;
            bool dirty = false;
            if (mbr) {
                ms_fill(sp, fsp, src);
                uint32_t cwd = (uint32_t)sp.tmp, cnt = 0;
                const uint32_t inv = ~cs & col_pattern;
                uint32_t cand = mbr, done = 0;
                while (cand) {
                    const uint32_t bit = __ffs(cand) - 1, m = 1u << bit;
                    cand &= ~m;
                    done |= m;
                    if (cwd & 1u) {
                        nsig |= m;
                        // the samples below and in the next column this one
                        // makes members
                        const uint32_t r = bit & 3u;
                        const uint32_t nb_ = ((3u << r) & 0xFu) | ((((7u << r) >> 1) & 0xFu) << 4);
                        cand |= (nb_ << (bit & ~3u)) & inv & ~done;
                    }
                    cwd >>= 1;
                    ++cnt;
                }
                if (nsig) {
                    // their signs, in bit order
#pragma unroll
                    for (uint32_t k = 0; k < 16; ++k)
                        if (nsig & (1u << k)) {
                            a[k & 3][k >> 2] = (cwd & 1u) ? -vnew : vnew;
                            cwd >>= 1;
                        }
                    cnt += (uint32_t)__builtin_popcount(nsig);
                    dirty = true;
                }
                sp.tmp >>= cnt;
                sp.bits -= cnt;
            }
            if (passes > 2 && csa) {
                if (mr.bits < 16) vlc_fill(mr, rmr, src);
                uint32_t mw = (uint32_t)mr.tmp;
#pragma unroll
                for (uint32_t k = 0; k < 16; ++k)
                    if (csa & (1u << k)) {
                        const int32_t v = a[k & 3][k >> 2];
                        const uint32_t m = (uint32_t)(v < 0 ? -v : v) ^ ((mw & 1u) ? ref1 : ref0);
                        a[k & 3][k >> 2] = v < 0 ? -(int32_t)m : (int32_t)m;
                        mw >>= 1;
                    }
                const uint32_t used = (uint32_t)__builtin_popcount(csa);
                mr.tmp >>= used;
                mr.bits -= used;
                dirty = true;
            }
            if (dirty) {
#pragma unroll
                for (uint32_t r = 0; r < 4; ++r) refine_store<C>(a[r], out, y + r, x, w, h, stride, vec);
            }
            const uint32_t comb = nsig | cs;
            nbot |= (uint64_t)group_row(comb & 0xFFFFu, 3) << x;
            const uint32_t np = comb | ((comb & 0x7777u) << 1) | ((comb & 0xEEEEu) >> 1);
            prev = (np | u) & 0xF000u;
#pragma unroll
            for (uint32_t r = 0; r < 4; ++r)
#pragma unroll
                for (uint32_t c = 0; c < 4; ++c) a[r][c] = b[r][c];
#pragma unroll
            for (uint32_t c = 0; c < 4; ++c) na[c] = nb[c];
            refine_group<C>(b, nb, out, y, x + 8, w, h, stride, vec, causal);
        }
        bot = nbot;
    }
}

}  // namespace sgn
