// Classic (EBCOT) code-block encoder: the MQ coder of T.800 Annex C (C.2)
// and the three coding passes of Annex D, with selective bypass (raw
// significance and refinement passes after the fourth bit-plane), context
// reset, termination on each pass, vertically causal context and
// segmentation symbols.  Every coding pass is emitted (lossless with the
// 5/3 transform).  The input generator for classic codestreams, the
// counterpart of the reference's bitplane_encode.rs; decoders other than
// this project's (OpenJPEG) read its output.
#include <algorithm>
#include <cstdint>
#include <cstring>
#include <utility>
#include <vector>

#include "encode_classic.hpp"

namespace sgn {
namespace {

// Table C.2 (Qe, next index after MPS / LPS, MPS switch)
struct QeRow {
    uint16_t qe;
    uint8_t nmps, nlps, sw;
};
constexpr QeRow kQe[47] = {
    {0x5601, 1, 1, 1},   {0x3401, 2, 6, 0},   {0x1801, 3, 9, 0},   {0x0AC1, 4, 12, 0},
    {0x0521, 5, 29, 0},  {0x0221, 38, 33, 0}, {0x5601, 7, 6, 1},   {0x5401, 8, 14, 0},
    {0x4801, 9, 14, 0},  {0x3801, 10, 14, 0}, {0x3001, 11, 17, 0}, {0x2401, 12, 18, 0},
    {0x1C01, 13, 20, 0}, {0x1601, 29, 21, 0}, {0x5601, 15, 14, 1}, {0x5401, 16, 14, 0},
    {0x5101, 17, 15, 0}, {0x4801, 18, 16, 0}, {0x3801, 19, 17, 0}, {0x3401, 20, 18, 0},
    {0x3001, 21, 19, 0}, {0x2801, 22, 19, 0}, {0x2401, 23, 20, 0}, {0x2201, 24, 21, 0},
    {0x1C01, 25, 22, 0}, {0x1801, 26, 23, 0}, {0x1601, 27, 24, 0}, {0x1401, 28, 25, 0},
    {0x1201, 29, 26, 0}, {0x1101, 30, 27, 0}, {0x0AC1, 31, 28, 0}, {0x09C1, 32, 29, 0},
    {0x08A1, 33, 30, 0}, {0x0521, 34, 31, 0}, {0x0441, 35, 32, 0}, {0x02A1, 36, 33, 0},
    {0x0221, 37, 34, 0}, {0x0141, 38, 35, 0}, {0x0111, 39, 36, 0}, {0x0085, 40, 37, 0},
    {0x0049, 41, 38, 0}, {0x0025, 42, 39, 0}, {0x0015, 43, 40, 0}, {0x0009, 44, 41, 0},
    {0x0005, 45, 42, 0}, {0x0001, 45, 43, 0}, {0x5601, 46, 46, 0},
};

// MQ encoder (C.2.1-C.2.9).  out[0] is the byte before the codeword (the
// B register's start), dropped at the end.
struct MqEncoder {
    std::vector<uint8_t> out;
    uint32_t a = 0x8000, c = 0, ct = 12;

    MqEncoder() { out.push_back(0); }
    void byteout() {
        uint8_t &b = out.back();
        if (b == 0xFF) {
            out.push_back(uint8_t(c >> 20));
            c &= 0xFFFFF;
            ct = 7;
        } else if (c < 0x8000000) {
            out.push_back(uint8_t(c >> 19));
            c &= 0x7FFFF;
            ct = 8;
        } else {
            ++b;
            if (b == 0xFF) {
                c &= 0x7FFFFFF;
                out.push_back(uint8_t(c >> 20));
                c &= 0xFFFFF;
                ct = 7;
            } else {
                out.push_back(uint8_t(c >> 19));
                c &= 0x7FFFF;
                ct = 8;
            }
        }
    }
    void renorm() {
        do {
            a <<= 1;
            c <<= 1;
            if (--ct == 0) byteout();
        } while (!(a & 0x8000));
    }
    void encode(uint32_t d, uint8_t &cx) {
        uint32_t i = cx & 0x7F, mps = cx >> 7;
        const uint32_t qe = kQe[i].qe;
        a -= qe;
        if (d == mps) {
            if (a & 0x8000) {
                c += qe;
                return;
            }
            if (a < qe)
                a = qe;
            else
                c += qe;
            i = kQe[i].nmps;
        } else {
            if (a < qe)
                c += qe;
            else
                a = qe;
            if (kQe[i].sw) mps ^= 1;
            i = kQe[i].nlps;
        }
        cx = uint8_t(mps << 7 | i);
        renorm();
    }
    // terminates the codeword and returns it (C.2.9)
    std::vector<uint8_t> flush() {
        const uint32_t t = c + a;
        c |= 0xFFFF;
        if (c >= t) c -= 0x8000;
        c <<= ct;
        byteout();
        c <<= ct;
        byteout();
        if (out.back() == 0xFF) out.pop_back();
        std::vector<uint8_t> r(out.begin() + 1, out.end());
        out.assign(1, 0);
        a = 0x8000;
        c = 0;
        ct = 12;
        return r;
    }
};

// Raw bits (D.6): MSB first, a 0 stuffed above the 7 bits after each 0xFF.
struct RawEncoder {
    std::vector<uint8_t> out;
    uint32_t acc = 0, n = 0, cap = 8;

    void put(uint32_t b) {
        acc = (acc << 1) | b;
        if (++n == cap) {
            out.push_back(uint8_t(acc));
            cap = uint8_t(acc) == 0xFF ? 7 : 8;
            acc = 0;
            n = 0;
        }
    }
    std::vector<uint8_t> flush() {
        if (n) out.push_back(uint8_t(acc << (cap - n)));
        // a final 0xFF would make the reader take the next byte's MSB as a
        // stuffed bit: give it one
        if (!out.empty() && out.back() == 0xFF) out.push_back(0);
        std::vector<uint8_t> r;
        r.swap(out);
        acc = 0;
        n = 0;
        cap = 8;
        return r;
    }
};

enum { kCxRl = 17, kCxUni = 18 };

struct BlockCoder {
    uint32_t w, h, pw, band, style;
    std::vector<uint8_t> f;     // flags with a border: 1 sig, 2 neg, 4 visited, 8 refined
    std::vector<uint32_t> mag;  // magnitudes, same layout
    uint8_t cx[19];
    bool raw = false;
    MqEncoder mq;
    RawEncoder rw;

    size_t at(uint32_t x, uint32_t y) const { return size_t(y + 1) * pw + x + 1; }
    void reset_contexts() {
        std::memset(cx, 0, sizeof(cx));
        cx[0] = 4;
        cx[kCxRl] = 3;
        cx[kCxUni] = 46;
    }
    void code(uint32_t d, int label) {
        if (raw)
            rw.put(d);
        else
            mq.encode(d, cx[label]);
    }
    bool hidden_below(uint32_t y) const {
        return (style & 0x08) && y + 1 < h && ((y + 1) >> 2) > (y >> 2);
    }
    void counts(size_t i, uint32_t y, int &hc, int &vc, int &dc) const {
        const bool hide = hidden_below(y);
        auto s = [&](size_t k) { return int(f[k] & 1); };
        hc = s(i - 1) + s(i + 1);
        vc = s(i - pw) + (hide ? 0 : s(i + pw));
        dc = s(i - pw - 1) + s(i - pw + 1) + (hide ? 0 : s(i + pw - 1) + s(i + pw + 1));
    }
    int zc(int hc, int vc, int dc) const {  // Table D.1
        if (band == 3) {
            const int hv = hc + vc;
            if (dc >= 3) return 8;
            if (dc == 2) return hv ? 7 : 6;
            if (dc == 1) return hv >= 2 ? 5 : (hv ? 4 : 3);
            return hv >= 2 ? 2 : hv;
        }
        if (band == 1) std::swap(hc, vc);
        if (hc == 2) return 8;
        if (hc == 1) return vc ? 7 : (dc ? 6 : 5);
        if (vc == 2) return 4;
        if (vc == 1) return 3;
        return dc >= 2 ? 2 : dc;
    }
    void code_sign(size_t i, uint32_t y) {  // Table D.3
        auto ct = [&](uint8_t p, uint8_t q) {
            int s = 0;
            if (p & 1) s += (p & 2) ? -1 : 1;
            if (q & 1) s += (q & 2) ? -1 : 1;
            return s > 0 ? 1 : (s < 0 ? -1 : 0);
        };
        const uint8_t below = hidden_below(y) ? 0 : f[i + pw];
        const int H = ct(f[i - 1], f[i + 1]), V = ct(f[i - pw], below);
        int label, flip;
        if (H == 0) {
            label = V == 0 ? 9 : 10;
            flip = V < 0;
        } else {
            label = 12 + H * V;
            flip = H < 0;
        }
        const uint32_t s = (f[i] & 2) ? 1 : 0;
        code(raw ? s : s ^ uint32_t(flip), label);
    }
    void pass_sigprop(uint32_t bit) {
        for (uint32_t y0 = 0; y0 < h; y0 += 4)
            for (uint32_t x = 0; x < w; ++x)
                for (uint32_t y = y0; y < y0 + 4 && y < h; ++y) {
                    const size_t i = at(x, y);
                    if (f[i] & 1) continue;
                    int hc, vc, dc;
                    counts(i, y, hc, vc, dc);
                    if (!(hc + vc + dc)) continue;
                    const uint32_t b = (mag[i] >> bit) & 1;
                    code(b, zc(hc, vc, dc));
                    f[i] |= 4;
                    if (b) {
                        code_sign(i, y);
                        f[i] |= 1;
                    }
                }
    }
    void pass_refine(uint32_t bit) {
        for (uint32_t y0 = 0; y0 < h; y0 += 4)
            for (uint32_t x = 0; x < w; ++x)
                for (uint32_t y = y0; y < y0 + 4 && y < h; ++y) {
                    const size_t i = at(x, y);
                    if ((f[i] & 5) != 1) continue;
                    int label = 16;
                    if (!(f[i] & 8)) {
                        int hc, vc, dc;
                        counts(i, y, hc, vc, dc);
                        label = hc + vc + dc ? 15 : 14;
                    }
                    code((mag[i] >> bit) & 1, label);
                    f[i] |= 8;
                }
    }
    void cleanup_one(uint32_t x, uint32_t y, uint32_t bit) {
        const size_t i = at(x, y);
        if (f[i] & 5) return;
        int hc, vc, dc;
        counts(i, y, hc, vc, dc);
        const uint32_t b = (mag[i] >> bit) & 1;
        code(b, zc(hc, vc, dc));
        if (b) {
            code_sign(i, y);
            f[i] |= 1;
        }
    }
    void pass_cleanup(uint32_t bit) {
        for (uint32_t y0 = 0; y0 < h; y0 += 4)
            for (uint32_t x = 0; x < w; ++x) {
                uint32_t y = y0;
                if (y0 + 4 <= h) {
                    bool run = true;
                    for (uint32_t k = 0; k < 4 && run; ++k) {
                        const size_t i = at(x, y0 + k);
                        int hc, vc, dc;
                        counts(i, y0 + k, hc, vc, dc);
                        run = !(f[i] & 5) && hc + vc + dc == 0;
                    }
                    if (run) {
                        uint32_t r = 4;
                        for (uint32_t k = 0; k < 4 && r == 4; ++k)
                            if ((mag[at(x, y0 + k)] >> bit) & 1) r = k;
                        code(r < 4, kCxRl);
                        if (r == 4) continue;
                        code(r >> 1, kCxUni);
                        code(r & 1, kCxUni);
                        const size_t i = at(x, y0 + r);
                        code_sign(i, y0 + r);
                        f[i] |= 1;
                        y = y0 + r + 1;
                    }
                }
                for (; y < y0 + 4 && y < h; ++y) cleanup_one(x, y, bit);
            }
    }
};

// Segment of pass p: each pass with termination on each pass, Table D.9 with
// selective bypass, else one for the block.
uint32_t segment_of(uint32_t style, uint32_t p) {
    if (style & 0x04) return p;
    if (style & 0x01) return p < 10 ? 0 : 1 + 2 * ((p - 10) / 3) + ((p - 10) % 3 == 2);
    return 0;
}

}  // namespace

bool classic_encode_block(const int32_t *coef, uint32_t w, uint32_t h, size_t stride, uint32_t K,
                          uint32_t band, uint32_t style, ClassicBlockCode &out) {
    out = ClassicBlockCode();
    uint32_t maxmag = 0;
    for (uint32_t y = 0; y < h; ++y)
        for (uint32_t x = 0; x < w; ++x) {
            const int32_t v = coef[y * stride + x];
            maxmag = std::max<uint32_t>(maxmag, uint32_t(v < 0 ? -int64_t(v) : v));
        }
    if (!maxmag) return true;  // nothing to code: the block is not included
    const uint32_t planes = 32 - __builtin_clz(maxmag);
    if (planes > K) return false;
    out.missing_msbs = K - planes;
    BlockCoder b;
    b.w = w;
    b.h = h;
    b.pw = w + 2;
    b.band = band;
    b.style = style;
    b.f.assign(size_t(b.pw) * (h + 2), 0);
    b.mag.assign(b.f.size(), 0);
    for (uint32_t y = 0; y < h; ++y)
        for (uint32_t x = 0; x < w; ++x) {
            const int32_t v = coef[y * stride + x];
            b.mag[b.at(x, y)] = uint32_t(v < 0 ? -int64_t(v) : v);
            if (v < 0) b.f[b.at(x, y)] |= 2;
        }
    b.reset_contexts();
    const uint32_t passes = 1 + 3 * (planes - 1);
    out.passes = passes;
    uint32_t seg_start = 0;
    for (uint32_t p = 0; p < passes; ++p) {
        const bool raw = (style & 0x01) && p >= 10 && p % 3 != 0;
        b.raw = raw;
        const uint32_t bit = planes - 1 - (p + 2) / 3;
        if (p % 3 == 0) {
            b.pass_cleanup(bit);
            if (style & 0x20)
                for (uint32_t s : {1u, 0u, 1u, 0u}) b.code(s, kCxUni);
            for (uint8_t &v : b.f) v &= uint8_t(~4);
        } else if (p % 3 == 1) {
            b.pass_sigprop(bit);
        } else {
            b.pass_refine(bit);
        }
        if (style & 0x02) b.reset_contexts();
        const bool last = p + 1 == passes;
        if (last || segment_of(style, p + 1) != segment_of(style, p)) {
            ClassicBlockCode::Segment sg;
            sg.bytes = raw ? b.rw.flush() : b.mq.flush();
            sg.passes = p + 1 - seg_start;
            out.segments.push_back(std::move(sg));
            seg_start = p + 1;
        }
    }
    return true;
}

}  // namespace sgn
