// Bit readers of the HT cleanup streams that take four bytes at a time.
//
// Stream semantics restate crates/signinum-j2k-native/src/j2c/
// ht_block_decode.rs: MEL :783-879 (MSB-first, 7 bits after 0xFF, last byte
// ORed with 0x0F, 0xFF past the end), forward MagSgn :881-929 (LSB-first,
// 7 bits after 0xFF, 0xFF pad) and backward VLC :931-999 (a byte whose low
// 7 bits are ones after a byte > 0x8F carries 7 bits, zeros past the end).
// A group of four bytes is appended in one step when no stuffing can occur
// in it, byte by byte otherwise; the delivered bit sequence is the same.
//
// Words come from a source object with `uint64_t word(uint32_t addr) const`
// returning the aligned 8 bytes at byte address addr of the arena (an LDS
// ring in the lane kernel).  The including file provides __device__ and
// __forceinline__.
#pragma once
#include <stdint.h>

namespace sgn {

// ---- forward stream of 4-byte groups (MagSgn, MEL) ------------------------
struct Fwd4 {
    uint64_t cur;     // undelivered bytes, next byte in bits 0..7
    uint32_t cur_n;   // bytes in cur
    uint32_t p;       // arena address of the next 8-byte word
    int32_t words;    // words from p on that hold segment bytes
    int32_t left;     // segment bytes not yet delivered
    uint32_t pad4;    // four pad bytes delivered past the end
};

template <class Src>
__device__ __forceinline__ void fwd4_init(Fwd4 &f, uint32_t off, uint32_t len, uint32_t pad,
                                          const Src &src) {
    const uint32_t a0 = off & ~7u;
    const uint32_t last = (off + (len ? len - 1 : 0)) & ~7u;
    f.left = (int32_t)len;
    f.pad4 = pad * 0x01010101u;
    f.cur = len ? (src.word(a0) >> (8 * (off & 7u))) : 0;
    f.cur_n = 8 - (off & 7u);
    f.p = a0 + 8;
    f.words = len ? (int32_t)((last - a0) >> 3) : 0;
}

// Next four bytes of the stream, first byte in bits 0..7.
template <class Src>
__device__ __forceinline__ uint32_t fwd4_get(Fwd4 &f, const Src &src) {
    uint32_t r;
    if (f.cur_n >= 4) {
        r = (uint32_t)f.cur;
        f.cur >>= 32;
        f.cur_n -= 4;
    } else {
        const uint64_t w = f.words > 0 ? src.word(f.p) : 0;
        f.p += 8;
        f.words--;
        const uint64_t m = f.cur | (w << (8 * f.cur_n));
        r = (uint32_t)m;
        f.cur = w >> (8 * (4 - f.cur_n));
        f.cur_n += 4;
    }
    if (f.left < 4) {
        const uint32_t keep = f.left <= 0 ? 0u : (0xFFFFFFFFu >> (8 * (4 - f.left)));
        r = (r & keep) | (f.pad4 & ~keep);
    }
    f.left -= 4;
    return r;
}

// ---- backward stream of 4-byte groups (VLC) --------------------------------
struct Rev4 {
    uint64_t cur;     // undelivered bytes, next byte in bits 56..63
    uint32_t cur_n;
    uint32_t p;       // arena address of the next (lower) 8-byte word
    int32_t words;    // words from p down that hold segment bytes
    int32_t left;
};

template <class Src>
__device__ __forceinline__ void rev4_init(Rev4 &r, uint32_t start, uint32_t len,
                                          const Src &src) {
    r.left = (int32_t)len;
    if (len == 0) {
        r.cur = 0;
        r.cur_n = 8;
        r.p = 0;
        r.words = 0;
        return;
    }
    const uint32_t a0 = start & ~7u, sh = start & 7u;
    const uint32_t first = (start + 1 - len) & ~7u;
    r.cur = src.word(a0) << (8 * (7 - sh));
    r.cur_n = sh + 1;
    r.p = a0 - 8;
    r.words = (int32_t)((a0 - first) >> 3);
}

// Next four bytes in decreasing address order, first byte in bits 0..7.
template <class Src>
__device__ __forceinline__ uint32_t rev4_get(Rev4 &r, const Src &src) {
    uint64_t m;
    if (r.cur_n >= 4) {
        m = r.cur;
        r.cur <<= 32;
        r.cur_n -= 4;
    } else {
        const uint64_t w = r.words > 0 ? src.word(r.p) : 0;
        r.p -= 8;
        r.words--;
        m = r.cur | (w >> (8 * r.cur_n));
        r.cur = w << (8 * (4 - r.cur_n));
        r.cur_n += 4;
    }
    uint32_t v = __builtin_bswap32((uint32_t)(m >> 32));
    if (r.left < 4) v &= r.left <= 0 ? 0u : (0xFFFFFFFFu >> (8 * (4 - r.left)));
    r.left -= 4;
    return v;
}

// ---- bit readers -----------------------------------------------------------
// MagSgn: LSB-first; a byte after 0xFF carries 7 bits (ht_block_decode.rs:881-929).
struct MsBits {
    uint64_t tmp;
    uint32_t bits, unstuff;
};

__device__ __forceinline__ uint32_t has_ff(uint32_t x) {  // nonzero iff a byte is 0xFF
    const uint32_t y = ~x;
    return (y - 0x01010101u) & ~y & 0x80808080u;
}

// Appends four-byte groups while fewer than 32 bits are held.  Each group
// starts below bit 32, so every appended bit, including the top bit of a
// 7-bit byte (ORed into the next position, as byte-wise appending does),
// stays inside the 64-bit accumulator.
// Appends the next four-byte group (needs b.bits <= 32).
template <class Src>
__device__ __forceinline__ void ms_append(MsBits &b, Fwd4 &f, const Src &src) {
    const uint32_t w = fwd4_get(f, src);
    if (!b.unstuff && !has_ff(w & 0x00FFFFFFu)) {
        b.tmp |= (uint64_t)w << b.bits;
        b.bits += 32;
        b.unstuff = (w >> 24) == 0xFF;
        return;
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const uint32_t c = (w >> (8 * i)) & 0xFF;
        b.tmp |= (uint64_t)c << b.bits;
        b.bits += 8 - b.unstuff;
        b.unstuff = c == 0xFF;
    }
}

template <class Src>
__device__ __forceinline__ void ms_fill(MsBits &b, Fwd4 &f, const Src &src, uint32_t need = 32) {
    while (b.bits < need) ms_append(b, f, src);
}

// VLC: backward; a byte whose low 7 bits are ones after a byte > 0x8F
// carries 7 bits (ht_block_decode.rs:931-999).
struct VlcBits {
    uint64_t tmp;
    uint32_t bits, unstuff;
};

template <class Src>
__device__ __forceinline__ void vlc_fill(VlcBits &b, Rev4 &r, const Src &src) {
    while (b.bits < 32) {
        const uint32_t w = rev4_get(r, src);
        // no byte with its low 7 bits all ones: no stuffing possible in the group
        if (!has_ff(w | 0x80808080u)) {
            b.tmp |= (uint64_t)w << b.bits;
            b.bits += 32;
            b.unstuff = (w >> 24) > 0x8F;
            continue;
        }
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const uint32_t c = (w >> (8 * i)) & 0xFF;
            const uint32_t dbits = 8 - ((b.unstuff && (c & 0x7F) == 0x7F) ? 1u : 0u);
            b.tmp |= (uint64_t)c << b.bits;
            b.bits += dbits;
            b.unstuff = c > 0x8F;
        }
    }
}

// MEL: MSB-first bits, 7 bits after 0xFF, the segment's last byte ORed with
// 0x0F and 0xFF past the end (ht_block_decode.rs:783-879).
// MEL bits in a 64-bit buffer, next bit in bit 63.  A byte after 0xFF
// carries 7 bits (its stuffed MSB is dropped); the segment's last byte is
// ORed with 0x0F and bytes past it read as 0xFF (ht_block_decode.rs:860-929).
struct Mel {
    uint64_t tmp;
    uint32_t bits;     // valid bits in tmp
    int32_t left;      // segment bytes not yet appended
    uint32_t unstuff;  // the last appended byte was 0xFF
    int32_t k;         // run-length state, 0..12
};

__device__ __forceinline__ Mel mel_make(uint32_t len) { return Mel{0, 0, (int32_t)len, 0, 0}; }

// Appends the next four segment bytes.
template <class Src>
__device__ __forceinline__ void mel_fill(Mel &m, Fwd4 &f, const Src &src) {
    const uint32_t w = fwd4_get(f, src);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        uint32_t c = (w >> (8 * i)) & 0xFF;
        if (m.left - i <= 1) c |= 0x0F;
        const uint32_t nb = 8 - m.unstuff;
        m.tmp |= (uint64_t)(c & ((1u << nb) - 1)) << (64 - m.bits - nb);
        m.bits += nb;
        m.unstuff = c == 0xFF;
    }
    m.left -= 4;
}

// Next MEL run: the number of all-insignificant quads before the next
// quad that needs a symbol, times two, plus one when that quad follows
// (ht_block_decode.rs:860-929).  A symbol is one bit (1: a full run of
// 2^E quads, E from the state k) or a zero bit and E bits of run length.
// Decodes one symbol from a buffer holding at least 6 bits.
__device__ __forceinline__ int mel_symbol(Mel &m) {
    // E for k = 0..12, four bits each
    constexpr uint64_t kExp = 0x5433222111000ull;
    const uint32_t e = (uint32_t)(kExp >> (4 * m.k)) & 0xF;
    const bool first = (m.tmp >> 63) != 0;
    const uint32_t v = e ? (uint32_t)((m.tmp << 1) >> (64 - e)) : 0u;
    const uint32_t n = first ? 1u : 1u + e;
    m.tmp <<= n;
    m.bits -= n;
    m.k = first ? (m.k < 12 ? m.k + 1 : 12) : (m.k > 0 ? m.k - 1 : 0);
    return first ? (int)(((1u << e) - 1) << 1) : (int)((v << 1) | 1);
}

template <class Src>
__device__ __forceinline__ int mel_run4(Mel &m, Fwd4 &f, const Src &src) {
    if (m.bits < 6) mel_fill(m, f, src);
    return mel_symbol(m);
}

// One sample of the magnitude/sign phase (ht_block_decode.rs:1010-1032),
// its sign-magnitude word converted to the integer sample in the same step
// (coefficient_to_i32, ht_block_decode.rs:225-234): shift = 31 - K (0 for
// K >= 31).  Insignificant samples give 0.
__device__ __forceinline__ int32_t ms_sample(MsBits &ms, uint32_t inf, uint32_t bit, uint32_t uq,
                                             uint32_t p, uint32_t shift, uint32_t &vn) {
    if ((inf & (1u << (4 + bit))) == 0) {
        vn = 0;
        return 0;
    }
    const uint32_t ms_val = (uint32_t)ms.tmp;
    const uint32_t m_n = uq - ((inf >> (12 + bit)) & 1u);
    ms.tmp >>= m_n;
    ms.bits -= m_n;
    // m_n <= 31: callers reject U_q above Zblk + 2 <= 31
    uint32_t v = ms_val & ((1u << m_n) - 1u);
    v |= ((inf >> (8 + bit)) & 1u) << m_n;
    v |= 1u;
    vn = v;
    const int32_t mag = (int32_t)((((v + 2) << (p - 1)) & 0x7FFFFFFFu) >> shift);
    return (ms_val & 1u) ? -mag : mag;
}

__device__ __forceinline__ uint32_t flog2r(uint32_t v) { return v ? 31u - __builtin_clz(v) : 0u; }

}  // namespace sgn
