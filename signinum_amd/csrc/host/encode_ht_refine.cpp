// HT refinement passes (T.814 7.4-7.5) for the test encoder: a code-block's
// cleanup pass codes the magnitudes above bit 0 (mu >> 1) and one
// significance-propagation (SigProp) and one magnitude-refinement (MagRef)
// pass code bit 0.  SigProp visits, stripe by stripe and 4x4 group by
// group, the samples insignificant after the cleanup that have a
// significant neighbour (significance spreading inside the group in
// visiting order, from the group to the left and from the stripe above),
// one bit each, then the signs of the samples that became significant;
// MagRef gives bit 0 of every sample significant after the cleanup, eight
// columns of a stripe at a time.  The SigProp bits fill the refinement
// segment forwards (LSB first, a 0 above the 7 bits after each 0xFF), the
// MagRef bits backwards from its end (LSB first in reverse byte order; a
// byte whose low seven bits are ones after one above 0x8F carries seven).
// This is the bit layout the decoder reads
// (crates/signinum-j2k-native/src/j2c/ht_block_decode.rs:1254-1432).  A
// sample of magnitude 1 that no SigProp visit reaches is not coded: the
// block is then lossy at bit 0.
#include <cstdint>
#include <vector>

#include "j2k_host.hpp"

namespace sgn {

Status ht_encode_block(const int32_t *coef, uint32_t w, uint32_t h, size_t stride, uint32_t K,
                       std::vector<uint8_t> &out);

namespace {

// Forward SigProp bits: LSB first; after a 0xFF byte the next one has 7.
struct FwdBits {
    std::vector<uint8_t> out;
    uint32_t acc = 0, n = 0, cap = 8;
    void put(uint32_t b) {
        acc |= (b & 1u) << n;
        if (++n == cap) {
            out.push_back(uint8_t(acc));
            cap = uint8_t(acc) == 0xFF ? 7 : 8;
            acc = 0;
            n = 0;
        }
    }
    void flush() {
        if (n) out.push_back(uint8_t(acc));
        n = 0;
        acc = 0;
    }
};

// Backward MagRef bits: LSB first, bytes emitted towards the segment
// start; a byte follows one above 0x8F (in reading order) with only seven
// bits when those seven are all ones.
struct RevBits {
    std::vector<uint8_t> out;  // in reading order (last byte of the segment first)
    uint32_t acc = 0, n = 0, cap = 8;
    bool last_big = true;  // the reader starts as if after a byte above 0x8F
    void put(uint32_t b) {
        acc |= (b & 1u) << n;
        ++n;
        if (n == 7 && last_big && acc == 0x7F) {
            emit();
            return;
        }
        if (n == 8) emit();
    }
    void emit() {
        out.push_back(uint8_t(acc));
        last_big = uint8_t(acc) > 0x8F;
        acc = 0;
        n = 0;
    }
    void flush() {
        if (n) emit();
    }
};

// bit index inside a 4x4 group: 4 * column + row
constexpr uint32_t kSpread[16] = {0x33,    0x76,    0xEC,    0xC8,    0x330,   0x760,
                                  0xEC0,   0xC80,   0x3300,  0x7600,  0xEC00,  0xC800,
                                  0x33000, 0x76000, 0xEC000, 0xC8000};

}  // namespace

// One block with refinement: cleanup of mu >> 1 at K - 1 bit-planes, then
// the SigProp and MagRef passes of bit 0.  cleanup / refine receive the two
// segments; false when no sample is significant in the cleanup (the caller
// then codes the block cleanup-only).
bool ht_encode_block_refined(const int32_t *coef, uint32_t w, uint32_t h, size_t stride,
                             uint32_t K, bool lossless_only, std::vector<uint8_t> &cleanup,
                             std::vector<uint8_t> &refine) {
    if (K < 2) return false;
    const uint32_t gw = (w + 3) / 4, ns = (h + 3) / 4;
    std::vector<int32_t> hi(size_t(w) * h);
    std::vector<uint8_t> mag0(size_t(w) * h), neg(size_t(w) * h);
    // significance after the cleanup, as 16-bit groups per stripe (+1 row)
    std::vector<uint32_t> sig(size_t(ns + 1) * (gw + 2), 0);
    bool any = false;
    for (uint32_t y = 0; y < h; ++y)
        for (uint32_t x = 0; x < w; ++x) {
            const int32_t v = coef[y * stride + x];
            const uint32_t m = uint32_t(v < 0 ? -int64_t(v) : v);
            const size_t i = size_t(y) * w + x;
            hi[i] = v < 0 ? -int32_t(m >> 1) : int32_t(m >> 1);
            mag0[i] = m & 1u;
            neg[i] = v < 0;
            if (m >> 1) {
                sig[size_t(y >> 2) * (gw + 2) + (x >> 2)] |= 1u << (4 * (x & 3) + (y & 3));
                any = true;
            }
        }
    if (!any) return false;
    if (ht_encode_block(hi.data(), w, h, w, K - 1, cleanup)) return false;
    auto pair = [&](uint32_t s, uint32_t g) {
        const size_t b = size_t(s) * (gw + 2) + g;
        return sig[b] | (sig[b + 1] << 16);
    };
    std::vector<uint8_t> reached(size_t(w) * h, 0);
    FwdBits sp;
    std::vector<uint32_t> prev_row(gw + 2, 0);  // previous stripe, with SigProp results
    for (uint32_t s = 0; s < ns; ++s) {
        const uint32_t y = 4 * s;
        uint32_t pattern = 0xFFFF;
        if (h - y < 4) pattern = h - y == 3 ? 0x7777 : (h - y == 2 ? 0x3333 : 0x1111);
        uint32_t prev = 0;
        for (uint32_t g = 0; g < gw; ++g) {
            const uint32_t x = 4 * g;
            uint32_t colp = pattern;
            const int sh = int(x) + 4 - int(w);
            if (sh > 0) colp >>= uint32_t(sh * 4);
            const uint32_t ps = prev_row[g] | (prev_row[g + 1] << 16);
            const uint32_t nsg = pair(s + 1, g);
            const uint32_t u = ((ps & 0x88888888u) >> 3) | ((nsg & 0x11111111u) << 3);
            const uint32_t cs = pair(s, g);
            uint32_t mbr = cs | ((cs & 0x77777777u) << 1) | ((cs & 0xEEEEEEEEu) >> 1) | u;
            const uint32_t t = mbr;
            mbr |= (t << 4) | (t >> 4) | (prev >> 12);
            mbr &= colp & ~cs;
            uint32_t new_sig = 0;
            if (mbr) {
                const uint32_t inv = ~cs & colp;
                uint32_t cand = mbr, done = 0;
                while (cand) {
                    const uint32_t b = uint32_t(__builtin_ctz(cand));
                    cand &= ~(1u << b);
                    done |= 1u << b;
                    const size_t i = size_t(y + (b & 3)) * w + x + (b >> 2);
                    const uint32_t bit = mag0[i];
                    reached[i] = 1;
                    sp.put(bit);
                    if (bit) {
                        new_sig |= 1u << b;
                        cand |= kSpread[b] & inv & ~done;
                    }
                }
                for (uint32_t sb = new_sig; sb; sb &= sb - 1) {
                    const uint32_t b = uint32_t(__builtin_ctz(sb));
                    sp.put(neg[size_t(y + (b & 3)) * w + x + (b >> 2)]);
                }
            }
            const uint32_t comb = new_sig | cs;
            prev_row[g] = comb & 0xFFFF;
            uint32_t np = comb | ((comb & 0x7777) << 1) | ((comb & 0xEEEE) >> 1);
            prev = (np | u) & 0xF000;
        }
    }
    sp.flush();
    if (lossless_only)
        for (size_t i = 0; i < reached.size(); ++i)
            if (mag0[i] && !hi[i] && !reached[i]) return false;
    RevBits mr;
    for (uint32_t s = 0; s < ns; ++s)
        for (uint32_t x0 = 0; x0 < w; x0 += 8) {
            const uint32_t sg = pair(s, x0 / 4);
            for (uint32_t j = 0; j < 8; ++j)
                for (uint32_t r = 0; r < 4; ++r)
                    if ((sg >> (4 * j + r)) & 1u) mr.put(mag0[size_t(4 * s + r) * w + x0 + j]);
        }
    mr.flush();
    refine = sp.out;
    refine.insert(refine.end(), mr.out.rbegin(), mr.out.rend());
    return true;
}

}  // namespace sgn
