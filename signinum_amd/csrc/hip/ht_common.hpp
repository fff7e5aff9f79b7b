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

// ---- chunked byte streams ------------------------------------------------
// The lane kernel reads its three streams through 8-byte aligned loads with
// one chunk prefetched ahead, so the serial bit-unstuffing logic runs on
// registers and a global load is only waited for once every eight bytes.
// Every load is unconditional with its address clamped to the last aligned
// word that holds a byte of the segment: the loaded register is the
// prefetch register itself, so the wait for it falls at its use one chunk
// later, not at the load.  Bytes past the end of a segment are replaced by
// the stream's pad value.  The arena holding the segments is 8-byte aligned
// and padded to a multiple of 16 bytes.

struct FwdBytes {
    const uint64_t *p;     // next aligned chunk to prefetch
    const uint64_t *last;  // last chunk holding a byte of the segment
    uint64_t cur, nxt;
    uint32_t cur_n;        // bytes left in cur
    int32_t left;          // segment bytes not yet delivered
    uint32_t pad;          // value delivered past the end of the segment
};

// The stream starts at byte `off` of `base`.  Addresses are formed from
// `base` by pointer arithmetic only, so the compiler keeps them in the
// global address space (global loads wait on vmcnt alone; a generic pointer
// would make every LDS wait also wait for the prefetch).
__device__ __forceinline__ void fwdb_init(FwdBytes &s, const uint8_t *base, uint32_t off,
                                          uint32_t len, uint32_t pad) {
    const uint64_t *w = reinterpret_cast<const uint64_t *>(base);
    const uint64_t *p = w + (off >> 3);
    const uint32_t sh = off & 7u;
    s.last = w + ((off + (len ? len - 1 : 0)) >> 3);
    s.pad = pad;
    s.left = (int32_t)len;
    s.cur = p[0] >> (8 * sh);
    s.cur_n = 8 - sh;
    s.nxt = *(p + 1 <= s.last ? p + 1 : s.last);
    s.p = p + 2;
}

__device__ __forceinline__ uint32_t fwdb_next(FwdBytes &s) {
    if (s.cur_n == 0) {
        s.cur = s.nxt;
        s.cur_n = 8;
        s.nxt = *(s.p <= s.last ? s.p : s.last);
        ++s.p;
    }
    uint32_t b = (uint32_t)(s.cur & 0xFF);
    s.cur >>= 8;
    s.cur_n--;
    if (s.left <= 0) b = s.pad;
    s.left--;
    return b;
}

// Bytes base[start], base[start-1], ... for `len` bytes, then zeros.
struct RevBytes {
    const uint64_t *p;      // next aligned chunk to prefetch
    const uint64_t *first;  // lowest chunk holding a byte of the segment
    uint64_t cur, nxt;
    uint32_t cur_n;
    int32_t left;
};

// Requires start + 1 >= len.
__device__ __forceinline__ void revb_init(RevBytes &s, const uint8_t *base, uint32_t start,
                                          uint32_t len) {
    const uint64_t *w = reinterpret_cast<const uint64_t *>(base);
    s.left = (int32_t)len;
    if (len == 0) {
        // nothing to deliver: every fetch reads the arena's first word
        s.p = w;
        s.first = w;
        s.cur = 0;
        s.cur_n = 8;
        s.nxt = 0;
        return;
    }
    const uint64_t *p = w + (start >> 3);
    const uint32_t sh = start & 7u;
    s.first = w + ((start + 1 - len) >> 3);
    s.cur = p[0] << (8 * (7 - sh));
    s.cur_n = sh + 1;
    s.nxt = *(p - 1 >= s.first ? p - 1 : s.first);
    s.p = p - 2;
}

__device__ __forceinline__ uint32_t revb_next(RevBytes &s) {
    if (s.cur_n == 0) {
        s.cur = s.nxt;
        s.cur_n = 8;
        s.nxt = *(s.p >= s.first ? s.p : s.first);
        --s.p;
    }
    uint32_t b = (uint32_t)(s.cur >> 56);
    s.cur <<= 8;
    s.cur_n--;
    if (s.left <= 0) b = 0;
    s.left--;
    return b;
}

// MEL decoder over a chunked stream (same symbol logic as mel_bit/mel_run).
struct MelChunked {
    FwdBytes src;
    uint32_t unstuff, cur, bits_left;
    int32_t k;
    uint32_t num_runs;
    uint64_t runs;
};

__device__ __forceinline__ uint32_t melc_bit(MelChunked &m) {
    if (m.bits_left == 0) {
        const bool last = (m.src.left == 1);
        uint32_t b = fwdb_next(m.src);  // 0xFF past the end
        if (last || m.src.left <= 0) b |= 0x0F;
        m.cur = b;
        m.bits_left = 8 - m.unstuff;
        m.unstuff = (b == 0xFF) ? 1u : 0u;
    }
    m.bits_left--;
    return (m.cur >> m.bits_left) & 1u;
}

__device__ __forceinline__ int melc_run(MelChunked &m) {
    if (m.num_runs == 0) {
        while (m.num_runs < 8) {
            const int eval = (m.k < 3) ? 0 : (m.k < 6) ? 1 : (m.k < 9) ? 2 : (m.k < 11) ? 3
                                                                    : (m.k == 11) ? 4 : 5;
            const uint32_t first = melc_bit(m);
            uint64_t run;
            if (first) {
                m.k = m.k + 1 > 12 ? 12 : m.k + 1;
                run = (uint64_t)((1u << eval) - 1) << 1;
            } else {
                uint32_t v = 0;
                for (int i = 0; i < eval; ++i) v = (v << 1) | melc_bit(m);
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

// Forward bit reader (MagSgn) over a chunked stream.
struct FwdChunked {
    FwdBytes src;
    uint64_t tmp;
    uint32_t bits, unstuff;
};

__device__ __forceinline__ uint32_t fwdc_fetch(FwdChunked &r) {
    while (r.bits <= 32) {
        const uint32_t b = fwdb_next(r.src);
        r.tmp |= (uint64_t)b << r.bits;
        r.bits += 8 - r.unstuff;
        r.unstuff = (b == 0xFF) ? 1u : 0u;
    }
    return (uint32_t)r.tmp;
}

__device__ __forceinline__ void fwdc_advance(FwdChunked &r, uint32_t n) {
    r.tmp >>= n;
    r.bits -= n;
}

// Backward VLC bit reader over a chunked stream.
struct RevChunked {
    RevBytes src;
    uint64_t tmp;
    uint32_t bits, unstuff;
};

// The block's cleanup segment is base[src .. src + lcup).
__device__ __forceinline__ void revc_init_vlc(RevChunked &r, const uint8_t *base, uint32_t src,
                                              uint32_t lcup, uint32_t scup) {
    const uint32_t b = base[src + lcup - 2];
    r.tmp = b >> 4;
    r.bits = 4 - (((r.tmp & 7) == 7) ? 1u : 0u);
    r.unstuff = ((b | 0x0F) > 0x8F) ? 1u : 0u;
    revb_init(r.src, base, src + lcup - 3, scup - 2);
}

__device__ __forceinline__ uint32_t revc_fetch(RevChunked &r) {
    while (r.bits <= 32) {
        const uint32_t b = revb_next(r.src);
        const uint32_t dbits = 8 - ((r.unstuff && (b & 0x7F) == 0x7F) ? 1u : 0u);
        r.tmp |= (uint64_t)b << r.bits;
        r.bits += dbits;
        r.unstuff = (b > 0x8F) ? 1u : 0u;
    }
    return (uint32_t)r.tmp;
}

__device__ __forceinline__ uint32_t revc_advance(RevChunked &r, uint32_t n) {
    r.tmp >>= n;
    r.bits -= n;
    return (uint32_t)r.tmp;
}

// Status codes written per code-block.
enum : uint8_t { kCbOk = 0, kCbBad = 1 };

}  // namespace sgn
