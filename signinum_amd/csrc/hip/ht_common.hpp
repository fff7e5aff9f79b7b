// Device-side HTJ2K bit readers shared by the block-decode kernels.
//
// The readers restate the reference decoder's stream semantics
// (crates/signinum-j2k-native/src/j2c/ht_block_decode.rs):
//   MEL      :783-879  forward, MSB-first, 7 bits after 0xFF, last byte | 0x0F
//   forward  :881-929  LSB-first, pad byte after the segment, 7 bits after 0xFF
//   backward :931-999  VLC (starts at Lcup-2 with the Scup nibble removed)
//                      and MagRef (starts unstuffed), drop a bit when the
//                      previous byte > 0x8F and the byte's low 7 bits are ones
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sgn {

struct MelState {
    const uint8_t *d;
    uint32_t pos, remaining;
    uint32_t unstuff, cur, bits_left;
    int32_t k;
    uint32_t num_runs;
    uint64_t runs;
};

__device__ __forceinline__ uint32_t mel_bit(MelState &m) {
    if (m.bits_left == 0) {
        uint32_t b = 0xFF;
        if (m.remaining > 0) {
            b = m.d[m.pos++];
            m.remaining--;
        }
        if (m.remaining == 0) b |= 0x0F;
        m.cur = b;
        m.bits_left = 8 - m.unstuff;
        m.unstuff = (b == 0xFF) ? 1u : 0u;
    }
    m.bits_left--;
    return (m.cur >> m.bits_left) & 1u;
}

__device__ __forceinline__ int mel_run(MelState &m) {
    if (m.num_runs == 0) {
        while (m.num_runs < 8) {
            const int eval = (m.k < 3) ? 0 : (m.k < 6) ? 1 : (m.k < 9) ? 2 : (m.k < 11) ? 3
                                                                    : (m.k == 11) ? 4 : 5;
            const uint32_t first = mel_bit(m);
            uint64_t run;
            if (first) {
                m.k = m.k + 1 > 12 ? 12 : m.k + 1;
                run = (uint64_t)((1u << eval) - 1) << 1;
            } else {
                uint32_t v = 0;
                for (int i = 0; i < eval; ++i) v = (v << 1) | mel_bit(m);
                m.k = m.k > 0 ? m.k - 1 : 0;
                run = ((uint64_t)v << 1) | 1;
            }
            m.runs |= run << (m.num_runs * 7);
            m.num_runs++;
            if (eval == 5 && first == 0 && m.num_runs >= 8) break;
        }
    }
    const int r = (int)(m.runs & 0x7F);
    m.runs >>= 7;
    m.num_runs--;
    return r;
}

struct FwdReader {
    const uint8_t *d;
    uint32_t pos, len;
    uint64_t tmp;
    uint32_t bits, unstuff, pad;
};

__device__ __forceinline__ void fwd_init(FwdReader &r, const uint8_t *d, uint32_t len,
                                         uint32_t pad) {
    r.d = d;
    r.pos = 0;
    r.len = len;
    r.tmp = 0;
    r.bits = 0;
    r.unstuff = 0;
    r.pad = pad;
}

__device__ __forceinline__ uint32_t fwd_fetch(FwdReader &r) {
    while (r.bits <= 32) {
        const uint32_t b = r.pos < r.len ? (uint32_t)r.d[r.pos++] : r.pad;
        r.tmp |= (uint64_t)b << r.bits;
        r.bits += 8 - r.unstuff;
        r.unstuff = (b == 0xFF) ? 1u : 0u;
    }
    return (uint32_t)r.tmp;
}

__device__ __forceinline__ void fwd_advance(FwdReader &r, uint32_t n) {
    r.tmp >>= n;
    r.bits -= n;
}

struct RevReader {
    const uint8_t *d;
    int32_t pos;
    uint32_t remaining;
    uint64_t tmp;
    uint32_t bits, unstuff;
};

__device__ __forceinline__ void rev_init_vlc(RevReader &r, const uint8_t *d, uint32_t lcup,
                                             uint32_t scup) {
    const uint32_t b = d[lcup - 2];
    r.d = d;
    r.tmp = b >> 4;
    r.bits = 4 - (((r.tmp & 7) == 7) ? 1u : 0u);
    r.pos = (int32_t)lcup - 3;
    r.remaining = scup - 2;
    r.unstuff = ((b | 0x0F) > 0x8F) ? 1u : 0u;
}

__device__ __forceinline__ void rev_init_mrp(RevReader &r, const uint8_t *d, uint32_t len) {
    r.d = d;
    r.pos = (int32_t)len - 1;
    r.remaining = len;
    r.tmp = 0;
    r.bits = 0;
    r.unstuff = 1;
}

__device__ __forceinline__ uint32_t rev_fetch(RevReader &r) {
    while (r.bits <= 32) {
        uint32_t b = 0;
        if (r.remaining > 0) {
            b = r.d[r.pos--];
            r.remaining--;
        }
        const uint32_t dbits = 8 - ((r.unstuff && (b & 0x7F) == 0x7F) ? 1u : 0u);
        r.tmp |= (uint64_t)b << r.bits;
        r.bits += dbits;
        r.unstuff = (b > 0x8F) ? 1u : 0u;
    }
    return (uint32_t)r.tmp;
}

__device__ __forceinline__ uint32_t rev_advance(RevReader &r, uint32_t n) {
    r.tmp >>= n;
    r.bits -= n;
    return (uint32_t)r.tmp;
}

// coefficient_to_i32 (ht_block_decode.rs:225-234)
__device__ __forceinline__ int32_t to_i32(uint32_t v, uint32_t kmax) {
    const uint32_t shift = kmax >= 31 ? 0 : 31 - kmax;
    const int32_t m = (int32_t)((v & 0x7FFFFFFFu) >> shift);
    return (v & 0x80000000u) ? -m : m;
}

// Status codes written per code-block.
enum : uint8_t { kCbOk = 0, kCbBad = 1 };

}  // namespace sgn
