// Image-level HTJ2K encoder: forward colour transform, forward DWT,
// quantisation, code-block coding, packetisation and codestream writing.
//
// The reversible path follows the reference encoder
// (crates/signinum-j2k-native/src/j2c/encode.rs:265-509, fdwt.rs,
// forward_mct.rs, quantize.rs:36-118, packet_encode.rs): RCT, 5/3 lifting,
// no quantisation, guard bits 1 (2 with the colour transform), one layer.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "encode_classic.hpp"
#include "j2k_host.hpp"
#include "sgn_api.hpp"

namespace sgn {

bool ht_encode_block_refined(const int32_t *coef, uint32_t w, uint32_t h, size_t stride,
                             uint32_t K, bool lossless_only, std::vector<uint8_t> &cleanup,
                             std::vector<uint8_t> &refine);
Status ht_encode_block(const int32_t *coef, uint32_t w, uint32_t h, size_t stride,
                       uint32_t k_max, std::vector<uint8_t> &out);

namespace {

constexpr double kAlpha = -1.586134342059924;
constexpr double kBeta = -0.052980118572961;
constexpr double kGamma = 0.882911075530934;
constexpr double kDelta = 0.443506852043971;
constexpr double kKappa = 1.230174104914001;

inline size_t ext_l(size_t i) { return i >= 1 ? i - 1 : 1 - i; }
inline size_t ext_r(size_t i, size_t n) { return i + 1 < n ? i + 1 : n - 2 - (i + 1 - n); }

// 1D_SD (F.4.8.1): forward 5/3 lifting in place, absolute origin i0.
void sd53(int32_t *s, size_t n, size_t step, uint32_t i0) {
#define S(i) s[(i) * step]
    if (n == 1) {
        if (i0 & 1) S(0) *= 2;
        return;
    }
    const size_t fe = i0 & 1, fo = 1 - fe;
    for (size_t i = fo; i < n; i += 2) {
        const int32_t a = S(ext_l(i)), b = S(ext_r(i, n));
        S(i) -= static_cast<int32_t>(std::floor((static_cast<double>(a) + b) / 2.0));
    }
    for (size_t i = fe; i < n; i += 2) {
        const int32_t a = S(ext_l(i)), b = S(ext_r(i, n));
        S(i) += static_cast<int32_t>(std::floor((static_cast<double>(a) + b + 2.0) / 4.0));
    }
#undef S
}

// 1D_SD (F.4.8.2): forward 9/7 lifting in double.
void sd97(double *s, size_t n, size_t step, uint32_t i0) {
#define S(i) s[(i) * step]
    if (n == 1) {
        if (i0 & 1) S(0) *= 2.0;
        return;
    }
    const size_t fe = i0 & 1, fo = 1 - fe;
    for (size_t i = fo; i < n; i += 2) S(i) += kAlpha * (S(ext_l(i)) + S(ext_r(i, n)));
    for (size_t i = fe; i < n; i += 2) S(i) += kBeta * (S(ext_l(i)) + S(ext_r(i, n)));
    for (size_t i = fo; i < n; i += 2) S(i) += kGamma * (S(ext_l(i)) + S(ext_r(i, n)));
    for (size_t i = fe; i < n; i += 2) S(i) += kDelta * (S(ext_l(i)) + S(ext_r(i, n)));
    for (size_t i = fe; i < n; i += 2) S(i) *= 1.0 / kKappa;
    for (size_t i = fo; i < n; i += 2) S(i) *= kKappa;
#undef S
}

// 1D_SR in double (the decoder's normalisation) for the band norms.
void sr97(double *s, size_t n, uint32_t i0) {
    if (n == 1) {
        if (i0 & 1) s[0] *= 0.5;
        return;
    }
    const size_t fe = i0 & 1, fo = 1 - fe;
    for (size_t i = fe; i < n; i += 2) s[i] *= kKappa;
    for (size_t i = fo; i < n; i += 2) s[i] *= 1.0 / kKappa;
    for (size_t i = fe; i < n; i += 2) s[i] -= kDelta * (s[ext_l(i)] + s[ext_r(i, n)]);
    for (size_t i = fo; i < n; i += 2) s[i] -= kGamma * (s[ext_l(i)] + s[ext_r(i, n)]);
    for (size_t i = fe; i < n; i += 2) s[i] -= kBeta * (s[ext_l(i)] + s[ext_r(i, n)]);
    for (size_t i = fo; i < n; i += 2) s[i] -= kAlpha * (s[ext_l(i)] + s[ext_r(i, n)]);
}

// L2 norm of the 1D synthesis basis of a low (hi=false) or high band at
// decomposition level `lev` (1 = finest), away from the borders.
double norm97_1d(int lev, bool hi) {
    const size_t n = static_cast<size_t>(64) << lev;
    std::vector<double> sig(n, 0.0);
    // coefficient in the middle of the band at level lev, then lev
    // synthesis steps; the band occupies even (low) or odd (high) slots.
    const size_t nb = n >> lev;
    std::vector<double> band(nb * 2, 0.0);
    // start at level lev: interleaved signal of length 2*nb
    band[(nb & ~static_cast<size_t>(1)) + (hi ? 1 : 0)] = 1.0;
    sr97(band.data(), band.size(), 0);
    std::vector<double> cur = band;
    for (int l = lev - 1; l >= 1; --l) {
        std::vector<double> nx(cur.size() * 2, 0.0);
        for (size_t i = 0; i < cur.size(); ++i) nx[2 * i] = cur[i];
        sr97(nx.data(), nx.size(), 0);
        cur.swap(nx);
    }
    double e = 0;
    for (double v : cur) e += v * v;
    return std::sqrt(e);
}

struct BitWriter {
    std::vector<uint8_t> &out;
    uint32_t acc = 0;
    int n = 0, maxbits = 8;
    explicit BitWriter(std::vector<uint8_t> &o) : out(o) {}
    void bit(uint32_t b) {
        acc = (acc << 1) | (b & 1);
        if (++n == maxbits) {
            out.push_back(static_cast<uint8_t>(acc));
            maxbits = acc == 0xFF ? 7 : 8;
            acc = 0;
            n = 0;
        }
    }
    void bits(uint32_t v, int len) {
        for (int i = len - 1; i >= 0; --i) bit((v >> i) & 1);
    }
    // Pad to a byte; a header ending in 0xFF is followed by a stuffed 0x00.
    void flush() {
        if (n > 0) {
            acc <<= (maxbits - n);
            out.push_back(static_cast<uint8_t>(acc));
            maxbits = acc == 0xFF ? 7 : 8;
            acc = 0;
            n = 0;
        }
        if (maxbits == 7) {
            out.push_back(0);
            maxbits = 8;
        }
    }
};

void tag_encode(TagTree &t, uint32_t x, uint32_t y, uint32_t threshold, BitWriter &bw) {
    uint32_t low = 0;
    for (int l = static_cast<int>(t.levels) - 1; l >= 0; --l) {
        const uint32_t n = t.leaf(x, y, static_cast<uint32_t>(l));
        if (low > t.low[n])
            t.low[n] = low;
        else
            low = t.low[n];
        while (low < threshold) {
            if (low >= t.value[n]) {
                if (!t.known[n]) {
                    bw.bit(1);
                    t.known[n] = 1;
                }
                break;
            }
            bw.bit(0);
            ++low;
        }
        t.low[n] = low;
    }
}

void tag_set_leaves(TagTree &t, const std::vector<uint32_t> &leaves) {
    // leaves in raster order at level 0; parents are minima of children
    for (size_t i = 0; i < leaves.size(); ++i) t.value[i] = leaves[i];
    for (uint32_t l = 1; l < t.levels; ++l)
        for (uint32_t y = 0; y < t.lh[l]; ++y)
            for (uint32_t x = 0; x < t.lw[l]; ++x) {
                uint32_t m = 0xFFFFFFFFu;
                for (uint32_t dy = 0; dy < 2; ++dy)
                    for (uint32_t dx = 0; dx < 2; ++dx) {
                        const uint32_t cx = 2 * x + dx, cy = 2 * y + dy;
                        if (cx < t.lw[l - 1] && cy < t.lh[l - 1])
                            m = std::min(m, t.value[t.off[l - 1] + cy * t.lw[l - 1] + cx]);
                    }
                t.value[t.off[l] + y * t.lw[l] + x] = m;
            }
}

// A classic block's header after its zero bit-plane code (B.10.6-B.10.7):
// the pass count (Table B.4), Lblock raised until every segment length
// fits in Lblock + floor(log2(its passes)) bits, the lengths.
void classic_block_header(BitWriter &bw, const ClassicBlockCode &bc) {
    const uint32_t n = bc.passes;
    if (n == 1) {
        bw.bit(0);
    } else if (n == 2) {
        bw.bits(2, 2);
    } else if (n <= 5) {
        bw.bits(0xC | (n - 3), 4);
    } else if (n <= 36) {
        bw.bits(0xF, 4);
        bw.bits(n - 6, 5);
    } else {
        bw.bits(0x1FF, 9);
        bw.bits(n - 37, 7);
    }
    uint32_t lblock = 3;
    for (const ClassicBlockCode::Segment &sg : bc.segments) {
        const uint32_t lg = 31 - __builtin_clz(sg.passes);
        while (sg.bytes.size() >= (uint64_t(1) << (lblock + lg))) ++lblock;
    }
    for (uint32_t k = 3; k < lblock; ++k) bw.bit(1);
    bw.bit(0);
    for (const ClassicBlockCode::Segment &sg : bc.segments)
        bw.bits(static_cast<uint32_t>(sg.bytes.size()),
                static_cast<int>(lblock + (31 - __builtin_clz(sg.passes))));
}

void put16(std::vector<uint8_t> &o, uint32_t v) {
    o.push_back(static_cast<uint8_t>(v >> 8));
    o.push_back(static_cast<uint8_t>(v));
}
void put32(std::vector<uint8_t> &o, uint32_t v) {
    put16(o, v >> 16);
    put16(o, v & 0xFFFF);
}

}  // namespace

Status encode_image(const void *pixels, uint32_t width, uint32_t height, uint32_t ncomp,
                    uint32_t bit_depth, size_t row_stride, const sgn_j2k_encode_options &opt,
                    std::vector<uint8_t> &cs) {
    if (!pixels || width == 0 || height == 0 || ncomp == 0 || ncomp > 4 || bit_depth == 0 ||
        bit_depth > 16) {
        set_last_error("encode: invalid image description");
        return kErrArg;
    }
    const bool rev = opt.reversible != 0;
    // sub-sampling of component c: nibbles of byte c (0 reads as 1)
    uint32_t sub_x[4], sub_y[4];
    for (uint32_t c = 0; c < 4; ++c) {
        sub_x[c] = (opt.subsampling >> (8 * c)) & 0xF;
        sub_y[c] = (opt.subsampling >> (8 * c + 4)) & 0xF;
        if (!sub_x[c]) sub_x[c] = 1;
        if (!sub_y[c]) sub_y[c] = 1;
    }
    // the colour transform needs the first three components on one grid
    const bool mct = opt.use_mct && ncomp >= 3 && sub_x[0] == sub_x[1] && sub_x[1] == sub_x[2] &&
                     sub_y[0] == sub_y[1] && sub_y[1] == sub_y[2];
    const uint32_t xo = opt.image_x0, yo = opt.image_y0;
    const uint32_t xto = opt.tile_x0, yto = opt.tile_y0;
    const uint64_t xsiz64 = uint64_t(xo) + width, ysiz64 = uint64_t(yo) + height;
    if (xsiz64 >= (1ull << 32) || ysiz64 >= (1ull << 32) || opt.packed_headers > 2) {
        set_last_error("encode: invalid image origin or packed-header mode");
        return kErrArg;
    }
    const uint32_t xsiz = uint32_t(xsiz64), ysiz = uint32_t(ysiz64);
    const uint32_t tw = opt.tile_width ? opt.tile_width : xsiz - xto;
    const uint32_t th = opt.tile_height ? opt.tile_height : ysiz - yto;
    // B-4: the first tile holds a reference grid point of the image
    if (xto > xo || yto > yo || uint64_t(xto) + tw <= xo || uint64_t(yto) + th <= yo) {
        set_last_error("encode: the tile-grid origin must lie in the first tile's image area");
        return kErrArg;
    }
    const uint32_t ntx = ceil_div(xsiz - xto, tw), nty = ceil_div(ysiz - yto, th);
    auto tile_rect = [&](uint32_t ti) {
        const uint32_t x0 = xto + (ti % ntx) * tw, y0 = yto + (ti / ntx) * th;
        Rect r;
        r.x0 = std::max(x0, xo);
        r.y0 = std::max(y0, yo);
        r.x1 = std::min(x0 + tw, xsiz);
        r.y1 = std::min(y0 + th, ysiz);
        return r;
    };
    uint32_t levels = opt.levels;
    {
        // at most log2 of the smallest side of any tile-component
        uint32_t mn = UINT32_MAX;
        for (uint32_t ti = 0; ti < ntx * nty; ++ti) {
            const Rect r = tile_rect(ti);
            for (uint32_t c = 0; c < ncomp; ++c) {
                const uint32_t cw = ceil_div(r.x1, sub_x[c]) - ceil_div(r.x0, sub_x[c]);
                const uint32_t ch = ceil_div(r.y1, sub_y[c]) - ceil_div(r.y0, sub_y[c]);
                mn = std::min(mn, std::min(cw, ch));
            }
        }
        const uint32_t lim = mn ? 31 - __builtin_clz(mn) : 0;
        levels = std::min(levels, lim);
    }
    if (opt.cb_width_log2 < 2 || opt.cb_height_log2 < 2 ||
        opt.cb_width_log2 + opt.cb_height_log2 > 12 || opt.cb_width_log2 > 10 ||
        opt.cb_height_log2 > 10) {
        set_last_error("encode: code-block size out of range");
        return kErrArg;
    }
    if (opt.progression > 4) return kErrArg;
    const uint32_t guard = opt.guard_bits ? opt.guard_bits : (rev ? (mct ? 2 : 1) : 2);

    CompParams cp;
    cp.precision = static_cast<uint8_t>(bit_depth);
    cp.levels = static_cast<uint8_t>(levels);
    cp.xcb = static_cast<uint8_t>(opt.cb_width_log2);
    cp.ycb = static_cast<uint8_t>(opt.cb_height_log2);
    const bool classic = opt.block_coder == 1;
    if (opt.block_coder > 1 || (opt.cb_style & ~0x3Fu) || (!classic && opt.cb_style)) {
        set_last_error("encode: invalid block coder or code-block style");
        return kErrArg;
    }
    // classic: the requested style, less predictable termination (the
    // encoder's flush is the plain C.2.9 one)
    cp.cbstyle = classic ? static_cast<uint8_t>(opt.cb_style & 0x2F) : 0x40;
    cp.transform = rev ? 1 : 0;
    cp.scod = static_cast<uint8_t>((opt.precinct_log2 ? 1 : 0) | (opt.sop ? 2 : 0) |
                                   (opt.eph ? 4 : 0));
    for (uint32_t r = 0; r <= levels; ++r) {
        const uint8_t pp = static_cast<uint8_t>(opt.precinct_log2 ? opt.precinct_log2 : 15);
        cp.ppx[r] = pp;
        cp.ppy[r] = pp;
    }
    cp.guard = static_cast<uint8_t>(guard);
    cp.qstyle = rev ? 0 : 2;
    cp.nsteps = static_cast<uint16_t>(1 + 3 * levels);
    // Step sizes.  Reversible: eps_b = P + gain_b (quantize.rs:44-72).
    for (uint32_t i = 0; i < cp.nsteps; ++i) {
        const uint32_t r = i == 0 ? 0 : 1 + (i - 1) / 3;
        const uint32_t type = i == 0 ? kLL : 1 + (i - 1) % 3;
        const int gain = type == kLL ? 0 : (type == kHH ? 2 : 1);
        if (rev) {
            cp.eps[i] = static_cast<uint16_t>(bit_depth + gain);
            cp.mu[i] = 0;
        } else {
            const int lev = r == 0 ? static_cast<int>(levels) : static_cast<int>(levels - r + 1);
            const bool hx = type == kHL || type == kHH, hy = type == kLH || type == kHH;
            const double norm = lev == 0 ? 1.0 : norm97_1d(lev, hx) * norm97_1d(lev, hy);
            const double base = opt.base_step > 0 ? opt.base_step : 1.0 / 64.0;
            const double d_abs = base * std::ldexp(1.0, static_cast<int>(bit_depth)) / norm;
            const double x = d_abs / std::ldexp(1.0, static_cast<int>(bit_depth) + gain);
            int e = static_cast<int>(std::floor(std::log2(x)));
            double m = x / std::ldexp(1.0, e);
            long mu = std::lround((m - 1.0) * 2048.0);
            if (mu >= 2048) {
                mu = 0;
                e += 1;
            }
            if (-e < 0 || -e > 31) {
                set_last_error("encode: step size out of range");
                return kErrArg;
            }
            cp.eps[i] = static_cast<uint16_t>(-e);
            cp.mu[i] = static_cast<uint16_t>(mu);
        }
    }

    MainHeader mh;
    mh.xsiz = xsiz;
    mh.ysiz = ysiz;
    mh.xo = xo;
    mh.yo = yo;
    mh.xt = tw;
    mh.yt = th;
    mh.xto = xto;
    mh.yto = yto;
    mh.ncomp = static_cast<uint16_t>(ncomp);
    mh.prog = static_cast<uint8_t>(opt.progression);
    mh.layers = 1;
    mh.mct = mct;
    for (uint32_t c = 0; c < ncomp; ++c) {
        mh.comp[c] = cp;
        mh.comp[c].xr = static_cast<uint8_t>(sub_x[c]);
        mh.comp[c].yr = static_cast<uint8_t>(sub_y[c]);
    }

    // ---- main header ----
    cs.clear();
    put16(cs, 0xFF4F);
    put16(cs, 0xFF51);
    put16(cs, 38 + 3 * ncomp);
    put16(cs, classic ? 0 : 0x4000);  // Rsiz: extended capabilities (CAP) for HT
    put32(cs, xsiz);
    put32(cs, ysiz);
    put32(cs, xo);
    put32(cs, yo);
    put32(cs, tw);
    put32(cs, th);
    put32(cs, xto);
    put32(cs, yto);
    put16(cs, ncomp);
    for (uint32_t c = 0; c < ncomp; ++c) {
        cs.push_back(static_cast<uint8_t>(bit_depth - 1));
        cs.push_back(static_cast<uint8_t>(sub_x[c]));
        cs.push_back(static_cast<uint8_t>(sub_y[c]));
    }
    if (!classic) {
        // CAP: Part 15 capabilities, Ccap15 as in codestream_write.rs:166-180
        put16(cs, 0xFF50);
        put16(cs, 8);
        put32(cs, 0x00020000);
        const uint32_t mb = bit_depth - 1;
        const uint32_t bp = mb <= 8 ? 0 : (mb < 28 ? mb - 8 : 13 + (mb >> 2));
        put16(cs, (rev ? 0 : 0x20) | bp);
    }
    put16(cs, 0xFF52);
    put16(cs, 12 + (opt.precinct_log2 ? levels + 1 : 0));
    cs.push_back(cp.scod);
    cs.push_back(static_cast<uint8_t>(opt.progression));
    put16(cs, 1);
    cs.push_back(mct ? 1 : 0);
    cs.push_back(static_cast<uint8_t>(levels));
    cs.push_back(static_cast<uint8_t>(cp.xcb - 2));
    cs.push_back(static_cast<uint8_t>(cp.ycb - 2));
    cs.push_back(cp.cbstyle);
    cs.push_back(rev ? 1 : 0);
    if (opt.precinct_log2)
        for (uint32_t r = 0; r <= levels; ++r)
            cs.push_back(static_cast<uint8_t>((cp.ppy[r] << 4) | cp.ppx[r]));
    put16(cs, 0xFF5C);
    if (rev) {
        put16(cs, 3 + cp.nsteps);
        cs.push_back(static_cast<uint8_t>(guard << 5));
        for (uint32_t i = 0; i < cp.nsteps; ++i) cs.push_back(static_cast<uint8_t>(cp.eps[i] << 3));
    } else {
        put16(cs, 3 + 2 * cp.nsteps);
        cs.push_back(static_cast<uint8_t>((guard << 5) | 2));
        for (uint32_t i = 0; i < cp.nsteps; ++i) put16(cs, (cp.eps[i] << 11) | cp.mu[i]);
    }

    std::vector<uint8_t> blk;
    // tile-parts in codestream order: header bytes (PPT segments) and body
    std::vector<std::vector<uint8_t>> part_hdr, part_body;
    // packet headers per tile-part (packed modes), split at packet boundaries
    std::vector<std::vector<std::vector<uint8_t>>> packed(ntx * nty);
    for (uint32_t ti = 0; ti < ntx * nty; ++ti) {
        const Rect tr = tile_rect(ti);
        const uint32_t w = tr.w(), h = tr.h();
        const size_t plane = static_cast<size_t>(w) * h;
        // ---- sample extraction, DC shift, colour transform ----
        std::vector<int32_t> ip(rev ? plane * ncomp : 0);
        std::vector<double> dp(rev ? 0 : plane * ncomp);
        const int32_t shift = 1 << (bit_depth - 1);
        // component c's samples (B-12: its tile rectangle divided by the
        // sub-sampling), sample (u, v) taken from the image pixel at
        // reference-grid position (u XRsiz, v YRsiz), row pitch the
        // component's width, at the start of the component's plane
        for (uint32_t c = 0; c < ncomp; ++c) {
            const uint32_t xr = sub_x[c], yr = sub_y[c];
            const Rect cr{ceil_div(tr.x0, xr), ceil_div(tr.y0, yr), ceil_div(tr.x1, xr),
                          ceil_div(tr.y1, yr)};
            for (uint32_t v = cr.y0; v < cr.y1; ++v) {
                const uint8_t *row =
                    static_cast<const uint8_t *>(pixels) + size_t(v * yr - yo) * row_stride;
                for (uint32_t u = cr.x0; u < cr.x1; ++u) {
                    const size_t si = static_cast<size_t>(u * xr - xo) * ncomp + c;
                    int32_t s = bit_depth <= 8 ? row[si]
                                               : reinterpret_cast<const uint16_t *>(row)[si];
                    s -= shift;
                    const size_t di = c * plane + size_t(v - cr.y0) * cr.w() + (u - cr.x0);
                    if (rev)
                        ip[di] = s;
                    else
                        dp[di] = s;
                }
            }
        }
        if (mct) {
            for (size_t i = 0; i < plane; ++i) {
                if (rev) {
                    const int32_t r = ip[i], g = ip[plane + i], b = ip[2 * plane + i];
                    ip[i] = static_cast<int32_t>(std::floor((r + 2.0 * g + b) / 4.0));
                    ip[plane + i] = b - g;
                    ip[2 * plane + i] = r - g;
                } else {
                    const double r = dp[i], g = dp[plane + i], b = dp[2 * plane + i];
                    dp[i] = 0.299 * r + 0.587 * g + 0.114 * b;
                    dp[plane + i] = -0.16875 * r - 0.33126 * g + 0.5 * b;
                    dp[2 * plane + i] = 0.5 * r - 0.41869 * g - 0.08131 * b;
                }
            }
        }
        // ---- geometry and forward DWT (vertical, then horizontal) ----
        std::vector<TileCompGeom> tcs(ncomp);
        std::vector<std::vector<int32_t>> qcoef(ncomp);  // per component: bands concatenated
        std::vector<std::vector<size_t>> band_off(ncomp);
        for (uint32_t c = 0; c < ncomp; ++c) {
            Status st = build_tile_comp(mh.comp[c], tr, tcs[c]);
            if (st) return st;
            TileCompGeom &g = tcs[c];
            // work buffer in tile coordinates; each level analyses the
            // current low-low rectangle in place and extracts 3 bands
            std::vector<int32_t> wi;
            std::vector<double> wd;
            if (rev)
                wi.assign(ip.begin() + c * plane, ip.begin() + (c + 1) * plane);
            else
                wd.assign(dp.begin() + c * plane, dp.begin() + (c + 1) * plane);
            size_t total = 0;
            band_off[c].resize(g.bands.size());
            for (size_t b = 0; b < g.bands.size(); ++b) {
                band_off[c][b] = total;
                total += static_cast<size_t>(g.bands[b].rect.w()) * g.bands[b].rect.h();
            }
            qcoef[c].assign(total, 0);
            Rect cur = g.res[g.nres - 1].rect;
            std::vector<int32_t> ci(rev ? plane : 0);
            std::vector<double> cd(rev ? 0 : plane);
            // cur buffer holds the low-low of the current level, w x h of `cur`
            if (rev)
                ci = wi;
            else
                cd = wd;
            for (int r = g.nres - 1; r >= 1; --r) {
                const Rect o = g.res[r].rect;
                const size_t ow = o.w(), oh = o.h();
                if (ow && oh) {
                    if (rev) {
                        for (size_t x = 0; x < ow; ++x) sd53(ci.data() + x, oh, ow, o.y0);
                        for (size_t y = 0; y < oh; ++y) sd53(ci.data() + y * ow, ow, 1, o.x0);
                    } else {
                        for (size_t x = 0; x < ow; ++x) sd97(cd.data() + x, oh, ow, o.y0);
                        for (size_t y = 0; y < oh; ++y) sd97(cd.data() + y * ow, ow, 1, o.x0);
                    }
                }
                // deinterleave: bands of resolution r, low-low to the next level
                const Rect ll = g.res[r - 1].rect;
                std::vector<int32_t> nli(rev ? static_cast<size_t>(ll.w()) * ll.h() : 0);
                std::vector<double> nld(rev ? 0 : static_cast<size_t>(ll.w()) * ll.h());
                const uint32_t lx0 = (o.x0 + 1) / 2, hx0 = o.x0 / 2;
                const uint32_t ly0 = (o.y0 + 1) / 2, hy0 = o.y0 / 2;
                for (size_t y = 0; y < oh; ++y)
                    for (size_t x = 0; x < ow; ++x) {
                        const uint32_t ax = o.x0 + static_cast<uint32_t>(x);
                        const uint32_t ay = o.y0 + static_cast<uint32_t>(y);
                        const int t = ((ax & 1) ? 1 : 0) + ((ay & 1) ? 2 : 0);  // 0 LL 1 HL 2 LH 3 HH
                        const size_t src = y * ow + x;
                        if (t == 0) {
                            const size_t di = static_cast<size_t>(ay / 2 - ly0) * ll.w() + (ax / 2 - lx0);
                            if (rev) nli[di] = ci[src]; else nld[di] = cd[src];
                            continue;
                        }
                        const uint32_t bidx = g.res[r].band[t - 1];
                        const BandGeom &bg = g.bands[bidx];
                        const uint32_t bx = (t & 1) ? ax / 2 - hx0 : ax / 2 - lx0;
                        const uint32_t by = (t & 2) ? ay / 2 - hy0 : ay / 2 - ly0;
                        const size_t di = band_off[c][bidx] + static_cast<size_t>(by) * bg.rect.w() + bx;
                        if (rev) {
                            qcoef[c][di] = ci[src];
                        } else {
                            uint32_t K;
                            float df;
                            band_quant(cp, bg.type, static_cast<uint32_t>(r), K, df);
                            (void)K;
                            const uint32_t idx = 1 + (r - 1) * 3 + (bg.type - 1);
                            const int gain = bg.type == kHH ? 2 : 1;
                            const double dq = std::ldexp(1.0 + cp.mu[idx] / 2048.0,
                                                         static_cast<int>(bit_depth) + gain - cp.eps[idx]);
                            const double v = cd[src];
                            const int32_t q = static_cast<int32_t>(std::floor(std::fabs(v) / dq));
                            qcoef[c][di] = v < 0 ? -q : q;
                        }
                    }
                if (rev) ci.swap(nli); else cd.swap(nld);
                cur = ll;
            }
            // remaining low-low band (resolution 0)
            {
                const BandGeom &bg = g.bands[g.res[0].band[0]];
                const size_t n = static_cast<size_t>(bg.rect.w()) * bg.rect.h();
                for (size_t i = 0; i < n; ++i) {
                    if (rev) {
                        qcoef[c][band_off[c][g.res[0].band[0]] + i] = ci[i];
                    } else {
                        const double dq = std::ldexp(1.0 + cp.mu[0] / 2048.0,
                                                     static_cast<int>(bit_depth) - cp.eps[0]);
                        const int32_t q = static_cast<int32_t>(std::floor(std::fabs(cd[i]) / dq));
                        qcoef[c][band_off[c][g.res[0].band[0]] + i] = cd[i] < 0 ? -q : q;
                    }
                }
            }
        }
        // ---- code-block coding ----
        std::vector<std::vector<std::vector<uint8_t>>> seg(ncomp);
        std::vector<std::vector<ClassicBlockCode>> cls(ncomp);
        // HT blocks with refinement passes: the refinement segment's length
        // (0: cleanup only)
        std::vector<std::vector<uint32_t>> ref_len(ncomp);
        for (uint32_t c = 0; c < ncomp; ++c) {
            TileCompGeom &g = tcs[c];
            seg[c].resize(g.cbs.size());
            ref_len[c].assign(g.cbs.size(), 0);
            if (classic) cls[c].resize(g.cbs.size());
            for (size_t k = 0; k < g.cbs.size(); ++k) {
                const CbGeom &cb = g.cbs[k];
                const BandGeom &bg = g.bands[cb.band];
                uint32_t K;
                float df;
                Status st = band_quant(cp, bg.type, bg.res, K, df);
                if (st) return st;
                const int32_t *base = qcoef[c].data() + band_off[c][cb.band] +
                                      static_cast<size_t>(cb.rect.y0 - bg.rect.y0) * bg.rect.w() +
                                      (cb.rect.x0 - bg.rect.x0);
                if (classic) {
                    ClassicBlockCode &bc = cls[c][k];
                    if (!classic_encode_block(base, cb.rect.w(), cb.rect.h(), bg.rect.w(), K,
                                              bg.type, cp.cbstyle, bc)) {
                        set_last_error("encode: coefficient beyond the band's bit-planes");
                        return kErrArg;
                    }
                    // the body bytes; a block with passes is included (its
                    // first segment, arithmetically coded, is never empty)
                    for (const ClassicBlockCode::Segment &sg : bc.segments)
                        seg[c][k].insert(seg[c][k].end(), sg.bytes.begin(), sg.bytes.end());
                    continue;
                }
                if (opt.ht_refine) {
                    std::vector<uint8_t> cup, ref;
                    if (ht_encode_block_refined(base, cb.rect.w(), cb.rect.h(), bg.rect.w(), K,
                                                opt.ht_refine == 1, cup, ref) &&
                        ref.size() < 2047) {
                        seg[c][k] = std::move(cup);
                        ref_len[c][k] = static_cast<uint32_t>(ref.size());
                        seg[c][k].insert(seg[c][k].end(), ref.begin(), ref.end());
                        continue;
                    }
                }
                st = ht_encode_block(base, cb.rect.w(), cb.rect.h(), bg.rect.w(), K, seg[c][k]);
                if (st) return st;
            }
        }
        // ---- packets ----
        std::vector<PacketId> order;
        Status st = progression(mh.prog, 1, tr, tcs.data(), static_cast<int>(ncomp), order);
        if (st) return st;
        std::vector<uint8_t> body;
        const bool pack = opt.packed_headers != 0;
        uint32_t packet_index = 0;
        for (const PacketId &pk : order) {
            const TileCompGeom &g = tcs[pk.comp];
            const ResGeom &rs = g.res[pk.res];
            if (opt.sop) {
                put16(body, 0xFF91);
                put16(body, 4);
                put16(body, packet_index & 0xFFFF);
            }
            ++packet_index;
            bool any = false;
            for (uint32_t b = 0; b < rs.nbands; ++b) {
                const PrecBand &pb = g.precs[rs.prec_base + pk.prec * rs.nbands + b];
                for (uint32_t k = 0; k < pb.ncbx * pb.ncby; ++k)
                    any |= !seg[pk.comp][pb.cb_base + k].empty();
            }
            // packed modes: the header (and its EPH) goes to its own buffer
            std::vector<uint8_t> phdr;
            std::vector<uint8_t> &hdr = pack ? phdr : body;
            BitWriter bw(hdr);
            bw.bit(any ? 1 : 0);
            if (any) {
                for (uint32_t b = 0; b < rs.nbands; ++b) {
                    const PrecBand &pb = g.precs[rs.prec_base + pk.prec * rs.nbands + b];
                    const uint32_t n = pb.ncbx * pb.ncby;
                    if (!n) continue;
                    const BandGeom &bg = g.bands[g.res[pk.res].band[b]];
                    uint32_t K;
                    float df;
                    band_quant(cp, bg.type, bg.res, K, df);
                    TagTree incl, zbp;
                    incl.init(pb.ncbx, pb.ncby);
                    zbp.init(pb.ncbx, pb.ncby);
                    std::vector<uint32_t> iv(n), zv(n);
                    for (uint32_t k = 0; k < n; ++k) {
                        const bool inc = !seg[pk.comp][pb.cb_base + k].empty();
                        iv[k] = inc ? 0 : 1;
                        zv[k] = classic ? (inc ? cls[pk.comp][pb.cb_base + k].missing_msbs : 0)
                                        : (ref_len[pk.comp][pb.cb_base + k] ? K - 2 : K - 1);
                    }
                    tag_set_leaves(incl, iv);
                    tag_set_leaves(zbp, zv);
                    for (uint32_t k = 0; k < n; ++k) {
                        const std::vector<uint8_t> &s = seg[pk.comp][pb.cb_base + k];
                        const uint32_t xi = k % pb.ncbx, yi = k / pb.ncbx;
                        tag_encode(incl, xi, yi, 1, bw);
                        if (s.empty()) continue;
                        if (classic) {
                            const ClassicBlockCode &bc = cls[pk.comp][pb.cb_base + k];
                            tag_encode(zbp, xi, yi, bc.missing_msbs + 1, bw);  // fully determined
                            classic_block_header(bw, bc);
                            continue;
                        }
                        tag_encode(zbp, xi, yi, K, bw);  // value below K: fully determined
                        const uint32_t lref = ref_len[pk.comp][pb.cb_base + k];
                        if (lref) {
                            // cleanup, SigProp and MagRef: three passes ('11' '00'),
                            // lengths of Lblock and Lblock + 1 bits
                            bw.bits(0xC, 4);
                            const uint32_t lcup = static_cast<uint32_t>(s.size()) - lref;
                            uint32_t lblock = 3;
                            while (lcup >= (1u << lblock) || lref >= (2u << lblock)) {
                                bw.bit(1);
                                ++lblock;
                            }
                            bw.bit(0);
                            bw.bits(lcup, static_cast<int>(lblock));
                            bw.bits(lref, static_cast<int>(lblock + 1));
                            continue;
                        }
                        bw.bit(0);  // one coding pass
                        const uint32_t len = static_cast<uint32_t>(s.size());
                        uint32_t lblock = 3;
                        while (len >= (1u << lblock)) {
                            bw.bit(1);
                            ++lblock;
                        }
                        bw.bit(0);
                        bw.bits(len, static_cast<int>(lblock));
                    }
                }
            }
            bw.flush();
            if (opt.eph) put16(hdr, 0xFF92);
            if (pack) packed[ti].push_back(std::move(phdr));
            if (any)
                for (uint32_t b = 0; b < rs.nbands; ++b) {
                    const PrecBand &pb = g.precs[rs.prec_base + pk.prec * rs.nbands + b];
                    for (uint32_t k = 0; k < pb.ncbx * pb.ncby; ++k) {
                        const std::vector<uint8_t> &s = seg[pk.comp][pb.cb_base + k];
                        body.insert(body.end(), s.begin(), s.end());
                    }
                }
        }
        // ---- tile-part header extras: PPT segments (A.7.5), each holding
        // whole packet headers, Zppt counting up ----
        std::vector<uint8_t> th_extra;
        if (opt.packed_headers == 2) {
            std::vector<uint8_t> segd;
            uint32_t z = 0;
            auto flush_ppt = [&]() {
                if (segd.empty()) return;
                put16(th_extra, 0xFF61);
                put16(th_extra, static_cast<uint32_t>(3 + segd.size()));
                th_extra.push_back(static_cast<uint8_t>(z++));
                th_extra.insert(th_extra.end(), segd.begin(), segd.end());
                segd.clear();
            };
            for (const std::vector<uint8_t> &ph : packed[ti]) {
                if (segd.size() + ph.size() > 65532) flush_ppt();
                segd.insert(segd.end(), ph.begin(), ph.end());
            }
            flush_ppt();
        }
        part_hdr.push_back(std::move(th_extra));
        part_body.push_back(std::move(body));
    }
    // ---- PPM segments (A.7.4) in the main header: per tile-part Nppm (four
    // bytes) and its packet headers, whole entries per segment ----
    if (opt.packed_headers == 1) {
        std::vector<uint8_t> segd;
        uint32_t z = 0;
        auto flush_ppm = [&]() {
            if (segd.empty()) return;
            put16(cs, 0xFF60);
            put16(cs, static_cast<uint32_t>(3 + segd.size()));
            cs.push_back(static_cast<uint8_t>(z++));
            cs.insert(cs.end(), segd.begin(), segd.end());
            segd.clear();
        };
        for (uint32_t ti = 0; ti < ntx * nty; ++ti) {
            std::vector<uint8_t> all;
            for (const std::vector<uint8_t> &ph : packed[ti]) all.insert(all.end(), ph.begin(), ph.end());
            if (all.size() + 4 > 65532) {
                set_last_error("encode: a tile's packet headers exceed one PPM segment");
                return kErrArg;
            }
            if (segd.size() + 4 + all.size() > 65532) flush_ppm();
            put32(segd, static_cast<uint32_t>(all.size()));
            segd.insert(segd.end(), all.begin(), all.end());
        }
        flush_ppm();
    }
    // ---- tile-parts ----
    for (uint32_t ti = 0; ti < ntx * nty; ++ti) {
        const std::vector<uint8_t> &hx = part_hdr[ti], &body = part_body[ti];
        put16(cs, 0xFF90);
        put16(cs, 10);
        put16(cs, ti);
        put32(cs, static_cast<uint32_t>(12 + hx.size() + 2 + body.size()));
        cs.push_back(0);
        cs.push_back(1);
        cs.insert(cs.end(), hx.begin(), hx.end());
        put16(cs, 0xFF93);
        cs.insert(cs.end(), body.begin(), body.end());
    }
    put16(cs, 0xFFD9);
    return kOk;
}

}  // namespace sgn
