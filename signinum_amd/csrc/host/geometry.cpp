// Tile-component geometry, progression orders, quantisation and HT tables.
#include <algorithm>
#include <cmath>
#include <mutex>
#include <tuple>
#include <vector>

#include "ht_vlc_rows.hpp"
#include "j2k_host.hpp"

namespace sgn {

static thread_local std::string g_last_error;

void set_last_error(const std::string &msg) { g_last_error = msg; }
const char *last_error() { return g_last_error.c_str(); }

// B-14 / B-15: resolution and sub-band rectangles; B-16: precinct counts;
// B-17/B-18: code-block size; Figure B.9: code-block grid anchored at 0.
// The precinct origin in sub-band coordinates follows the reference
// (tile.rs:542-570): the resolution-domain anchor floor(x0 / 2^PP) * 2^PP,
// halved together with the exponent for r > 0.
Status build_tile_comp(const CompParams &cp, const Rect &tile, TileCompGeom &g) {
    g = TileCompGeom();
    g.trect = tile;
    g.xr = cp.xr;
    g.yr = cp.yr;
    const Rect trect{ceil_div(tile.x0, cp.xr), ceil_div(tile.y0, cp.yr), ceil_div(tile.x1, cp.xr),
                     ceil_div(tile.y1, cp.yr)};
    g.rect = trect;
    g.nres = static_cast<uint8_t>(cp.levels + 1);
    const uint32_t nl = cp.levels;
    for (uint32_t r = 0; r < g.nres; ++r) {
        ResGeom &rs = g.res[r];
        const uint64_t den = 1ull << (nl - r);
        rs.rect.x0 = static_cast<uint32_t>((trect.x0 + den - 1) / den);
        rs.rect.y0 = static_cast<uint32_t>((trect.y0 + den - 1) / den);
        rs.rect.x1 = static_cast<uint32_t>((trect.x1 + den - 1) / den);
        rs.rect.y1 = static_cast<uint32_t>((trect.y1 + den - 1) / den);
        rs.ppx = cp.ppx[r];
        rs.ppy = cp.ppy[r];
        rs.npx = rs.rect.w() ? ceil_div(rs.rect.x1, 1u << rs.ppx) - (rs.rect.x0 >> rs.ppx) : 0;
        rs.npy = rs.rect.h() ? ceil_div(rs.rect.y1, 1u << rs.ppy) - (rs.rect.y0 >> rs.ppy) : 0;
        uint32_t xcb = cp.xcb, ycb = cp.ycb;
        if (r > 0) {
            if (rs.ppx == 0 || rs.ppy == 0) {
                if (rs.npx && rs.npy) {
                    set_last_error("precinct exponent 0 above resolution 0");
                    return kErrMalformed;
                }
            } else {
                xcb = xcb < rs.ppx - 1u ? xcb : rs.ppx - 1u;
                ycb = ycb < rs.ppy - 1u ? ycb : rs.ppy - 1u;
            }
        } else {
            xcb = xcb < rs.ppx ? xcb : rs.ppx;
            ycb = ycb < rs.ppy ? ycb : rs.ppy;
        }
        rs.cbw = 1u << xcb;
        rs.cbh = 1u << ycb;
        rs.nbands = r == 0 ? 1 : 3;
        const uint32_t nb = r == 0 ? nl : nl - r + 1;
        for (uint32_t bi = 0; bi < rs.nbands; ++bi) {
            BandGeom bg;
            bg.type = static_cast<uint8_t>(r == 0 ? kLL : bi + 1);
            bg.res = static_cast<uint8_t>(r);
            const uint64_t xob = (bg.type == kHL || bg.type == kHH) ? 1 : 0;
            const uint64_t yob = (bg.type == kLH || bg.type == kHH) ? 1 : 0;
            const uint64_t ox = nb ? (1ull << (nb - 1)) * xob : 0;
            const uint64_t oy = nb ? (1ull << (nb - 1)) * yob : 0;
            const uint64_t d = 1ull << nb;
            auto sb = [&](uint64_t v, uint64_t o) {
                return static_cast<uint32_t>(((v > o ? v - o : 0) + d - 1) / d);
            };
            bg.rect.x0 = sb(trect.x0, ox);
            bg.rect.x1 = sb(trect.x1, ox);
            bg.rect.y0 = sb(trect.y0, oy);
            bg.rect.y1 = sb(trect.y1, oy);
            rs.band[bi] = static_cast<uint32_t>(g.bands.size());
            g.bands.push_back(bg);
        }
        rs.prec_base = static_cast<uint32_t>(g.precs.size());
        const uint32_t np = rs.npx * rs.npy;
        if (np == 0) continue;
        uint32_t ppx = rs.ppx, ppy = rs.ppy;
        uint32_t xs = (rs.rect.x0 >> ppx) << ppx, ys = (rs.rect.y0 >> ppy) << ppy;
        if (r > 0) {
            ppx -= 1;
            ppy -= 1;
            xs /= 2;
            ys /= 2;
        }
        for (uint32_t py = 0; py < rs.npy; ++py) {
            for (uint32_t px = 0; px < rs.npx; ++px) {
                Rect pr;
                pr.x0 = px * (1u << ppx) + xs;
                pr.y0 = py * (1u << ppy) + ys;
                pr.x1 = pr.x0 + (1u << ppx);
                pr.y1 = pr.y0 + (1u << ppy);
                for (uint32_t bi = 0; bi < rs.nbands; ++bi) {
                    const Rect sbr = g.bands[rs.band[bi]].rect;
                    PrecBand pb;
                    const uint32_t ax0 = pr.x0 > sbr.x0 ? pr.x0 : sbr.x0;
                    const uint32_t ay0 = pr.y0 > sbr.y0 ? pr.y0 : sbr.y0;
                    const uint32_t ax1 = pr.x1 < sbr.x1 ? pr.x1 : sbr.x1;
                    const uint32_t ay1 = pr.y1 < sbr.y1 ? pr.y1 : sbr.y1;
                    const uint32_t cx0 = (ax0 / rs.cbw) * rs.cbw, cy0 = (ay0 / rs.cbh) * rs.cbh;
                    const uint32_t cx1 = ceil_div(ax1, rs.cbw) * rs.cbw;
                    const uint32_t cy1 = ceil_div(ay1, rs.cbh) * rs.cbh;
                    pb.ncbx = (sbr.w() == 0 || cx1 <= cx0) ? 0 : (cx1 - cx0) / rs.cbw;
                    pb.ncby = (sbr.h() == 0 || cy1 <= cy0) ? 0 : (cy1 - cy0) / rs.cbh;
                    pb.cb_base = static_cast<uint32_t>(g.cbs.size());
                    for (uint32_t j = 0; j < pb.ncby; ++j) {
                        for (uint32_t i = 0; i < pb.ncbx; ++i) {
                            CbGeom cb;
                            const uint32_t bx0 = cx0 + i * rs.cbw, by0 = cy0 + j * rs.cbh;
                            cb.rect.x0 = bx0 > sbr.x0 ? bx0 : sbr.x0;
                            cb.rect.y0 = by0 > sbr.y0 ? by0 : sbr.y0;
                            cb.rect.x1 = bx0 + rs.cbw < sbr.x1 ? bx0 + rs.cbw : sbr.x1;
                            cb.rect.y1 = by0 + rs.cbh < sbr.y1 ? by0 + rs.cbh : sbr.y1;
                            if (cb.rect.x1 < cb.rect.x0) cb.rect.x1 = cb.rect.x0;
                            if (cb.rect.y1 < cb.rect.y0) cb.rect.y1 = cb.rect.y0;
                            cb.band = rs.band[bi];
                            cb.xi = static_cast<uint16_t>(i);
                            cb.yi = static_cast<uint16_t>(j);
                            g.cbs.push_back(cb);
                        }
                    }
                    g.precs.push_back(pb);
                }
            }
        }
    }
    return kOk;
}

// Packet order of one tile (B.12).  Layer- and resolution-major orders
// walk (layer, resolution, component, precinct) directly; the position
// orders sort every precinct of every component and resolution by its
// position on the reference grid (progression.rs:257-345): a precinct row
// or column starts at the tile origin, or at the next multiple of the step
// XRsiz * 2^(PP + NL - r) when the resolution's origin times 2^(NL - r) is
// a multiple of it, and each next one at the next multiple of the step
// (tile.rs:544-645).
Status progression(uint8_t order, uint16_t layers, const Rect &trect, const TileCompGeom *tcs,
                   int ncomp, std::vector<PacketId> &out) {
    (void)trect;
    out.clear();
    uint32_t maxr = 0;
    for (int c = 0; c < ncomp; ++c) maxr = tcs[c].nres > maxr ? tcs[c].nres : maxr;
    auto emit_rc = [&](uint32_t l, uint32_t r, int c) {
        if (r >= tcs[c].nres) return;
        const uint32_t np = tcs[c].res[r].npx * tcs[c].res[r].npy;
        for (uint32_t p = 0; p < np; ++p)
            out.push_back({static_cast<uint16_t>(l), static_cast<uint16_t>(r),
                           static_cast<uint16_t>(c), p});
    };
    if (order == 0) {  // LRCP
        for (uint32_t l = 0; l < layers; ++l)
            for (uint32_t r = 0; r < maxr; ++r)
                for (int c = 0; c < ncomp; ++c) emit_rc(l, r, c);
        return kOk;
    }
    if (order == 1) {  // RLCP
        for (uint32_t r = 0; r < maxr; ++r)
            for (uint32_t l = 0; l < layers; ++l)
                for (int c = 0; c < ncomp; ++c) emit_rc(l, r, c);
        return kOk;
    }
    if (order > 4) return kErrMalformed;
    struct Pos {
        uint64_t y, x;
        uint16_t c, r;
        uint32_t idx;
    };
    std::vector<Pos> v;
    auto first_at = [](uint64_t origin, uint64_t res_origin, uint32_t s, uint64_t step) {
        if (origin % step != 0 && ((res_origin << s) % step) == 0)
            return (origin / step + 1) * step;
        return origin;
    };
    for (int c = 0; c < ncomp; ++c) {
        const TileCompGeom &g = tcs[c];
        for (uint32_t r = 0; r < g.nres; ++r) {
            const ResGeom &rs = g.res[r];
            if (rs.npx * rs.npy == 0) continue;
            const uint32_t s = g.nres - 1 - r;
            const uint64_t stepx = uint64_t(g.xr) << (rs.ppx + s);
            const uint64_t stepy = uint64_t(g.yr) << (rs.ppy + s);
            uint64_t ry = first_at(g.trect.y0, rs.rect.y0, s, stepy);
            for (uint32_t py = 0; py < rs.npy; ++py) {
                uint64_t rx = first_at(g.trect.x0, rs.rect.x0, s, stepx);
                for (uint32_t px = 0; px < rs.npx; ++px) {
                    v.push_back({ry, rx, static_cast<uint16_t>(c), static_cast<uint16_t>(r),
                                 py * rs.npx + px});
                    rx = (rx / stepx + 1) * stepx;
                }
                ry = (ry / stepy + 1) * stepy;
            }
        }
    }
    auto key = [&](const Pos &p) {
        // (major .. minor) per order; the precinct index breaks ties
        if (order == 2) return std::make_tuple(uint64_t(p.r), p.y, p.x, uint64_t(p.c), p.idx);
        if (order == 3) return std::make_tuple(p.y, p.x, uint64_t(p.c), uint64_t(p.r), p.idx);
        return std::make_tuple(uint64_t(p.c), p.y, p.x, uint64_t(p.r), p.idx);
    };
    std::stable_sort(v.begin(), v.end(), [&](const Pos &a, const Pos &b) { return key(a) < key(b); });
    for (const Pos &p : v)
        for (uint32_t l = 0; l < layers; ++l)
            out.push_back({static_cast<uint16_t>(l), p.r, p.c, p.idx});
    return kOk;
}

Status band_quant(const CompParams &cp, uint8_t type, uint32_t res, uint32_t &k_max,
                  float &delta) {
    uint32_t eps = 0, mu = 0;
    if (cp.qstyle == 1) {
        const uint32_t nb = res == 0 ? cp.levels : cp.levels + 1 - res;
        if (cp.eps[0] < cp.levels) return kErrMalformed;
        eps = cp.eps[0] - cp.levels + nb;
        mu = cp.mu[0];
    } else {
        const uint32_t idx = res == 0 ? 0 : 1 + (res - 1) * 3 + (type - 1);
        if (idx >= cp.nsteps) return kErrMalformed;
        eps = cp.eps[idx];
        mu = cp.mu[idx];
    }
    if (cp.guard + eps < 1) return kErrMalformed;
    k_max = cp.guard + eps - 1;
    if (k_max > 32) return kErrMalformed;
    delta = 1.0f;
    if (cp.qstyle != 0) {
        const int gain = type == kLL ? 0 : (type == kHH ? 2 : 1);
        const int e = static_cast<int>(cp.precision) + gain - static_cast<int>(eps);
        const float p2 = e >= 0 ? static_cast<float>(1u << e) : 1.0f / static_cast<float>(1u << -e);
        delta = p2 * (1.0f + static_cast<float>(mu) / 2048.0f);
    }
    return kOk;
}

void TagTree::init(uint32_t w, uint32_t h) {
    levels = 0;
    uint32_t cw = w, ch = h, total = 0;
    if (w == 0 || h == 0) {
        value.clear();
        low.clear();
        known.clear();
        return;
    }
    for (;;) {
        lw[levels] = cw;
        lh[levels] = ch;
        off[levels] = total;
        total += cw * ch;
        levels++;
        if (cw == 1 && ch == 1) break;
        cw = (cw + 1) / 2;
        ch = (ch + 1) / 2;
    }
    value.assign(total, 0xFFFFFFFFu);
    low.assign(total, 0);
    known.assign(total, 0);
}

// ---- HT tables --------------------------------------------------------

static void build_vlc_decode(uint16_t *table, const uint32_t *rows, size_t nrows) {
    for (int idx = 0; idx < 1024; ++idx) {
        const uint32_t c_q = idx >> 7, bits = idx & 0x7F;
        uint16_t e = 0;
        for (size_t i = 0; i < nrows; ++i) {
            const uint32_t w = rows[i];
            const uint32_t len = w & 7, cwd = (w >> 3) & 0x7F, e1 = (w >> 10) & 0xF;
            const uint32_t ek = (w >> 14) & 0xF, rho = (w >> 18) & 0xF, uoff = (w >> 22) & 1;
            if ((w >> 23) != c_q || (bits & ((1u << len) - 1)) != cwd) continue;
            e = static_cast<uint16_t>(len | (uoff << 3) | (rho << 4) | (e1 << 8) | (ek << 12));
            break;
        }
        table[idx] = e;
    }
}

// Encoder choice for (c_q, rho, emb): a row with u_off = (emb != 0) whose
// EMB pattern agrees with emb ((emb & e_k) == e_1); among those the one
// with the most implied MSBs (popcount e_k), then the shortest codeword.
static void build_vlc_encode(uint32_t *table, const uint32_t *rows, size_t nrows) {
    for (uint32_t i = 0; i < 2048; ++i) {
        const uint32_t c_q = i >> 8, rho = (i >> 4) & 0xF, emb = i & 0xF;
        table[i] = 0;
        if ((emb & rho) != emb || (rho == 0 && c_q == 0)) continue;
        int best = -1, best_pop = -1, best_len = 99;
        for (size_t k = 0; k < nrows; ++k) {
            const uint32_t w = rows[k];
            const uint32_t len = w & 7, e1 = (w >> 10) & 0xF, ek = (w >> 14) & 0xF;
            if ((w >> 23) != c_q || ((w >> 18) & 0xF) != rho) continue;
            const uint32_t uoff = (w >> 22) & 1;
            if (uoff != (emb ? 1u : 0u)) continue;
            if (emb && (emb & ek) != e1) continue;
            const int pop = __builtin_popcount(ek);
            if (pop > best_pop || (pop == best_pop && static_cast<int>(len) < best_len)) {
                best = static_cast<int>(k);
                best_pop = pop;
                best_len = static_cast<int>(len);
            }
        }
        if (best < 0) continue;
        const uint32_t w = rows[best];
        table[i] = ((w >> 3) & 0x7F) | ((w & 7) << 8) | (((w >> 14) & 0xF) << 12) | (1u << 16);
    }
}

static int uvlc_prefix(int bits, int &value) {
    if (bits & 1) { value = 1; return 1; }
    if (bits & 2) { value = 2; return 2; }
    if (bits & 4) { value = 3; return 3; }
    value = 5;
    return 3;
}
static int uvlc_suffix_len(int v) { return v < 3 ? 0 : (v == 3 ? 1 : 5); }
static uint16_t uvlc_entry(int pl, int st, int s0, int b0, int b1) {
    return static_cast<uint16_t>(pl | (st << 3) | (s0 << 7) | (b0 << 10) | (b1 << 13));
}

// U-VLC decode tables from the prefix/suffix rules of T.814 (clause 7.3.6),
// including the first-row pair rules.
static void build_uvlc(uint16_t *table, bool initial) {
    const int modes = initial ? 5 : 4;
    for (int mode = 0; mode < modes; ++mode)
        for (int bits = 0; bits < 64; ++bits) {
            uint16_t e = 0;
            int v0, v1;
            if (mode == 1) {
                const int l0 = uvlc_prefix(bits, v0);
                e = uvlc_entry(l0, uvlc_suffix_len(v0), uvlc_suffix_len(v0), v0, 0);
            } else if (mode == 2) {
                const int l1 = uvlc_prefix(bits, v1);
                e = uvlc_entry(l1, uvlc_suffix_len(v1), 0, 0, v1);
            } else if (mode == 3) {
                const int l0 = uvlc_prefix(bits, v0);
                if (initial && v0 > 2) {
                    const int b = (bits >> l0) & 1;
                    e = uvlc_entry(l0 + 1, uvlc_suffix_len(v0), uvlc_suffix_len(v0), v0, b + 1);
                } else {
                    const int l1 = uvlc_prefix(bits >> l0, v1);
                    e = uvlc_entry(l0 + l1, uvlc_suffix_len(v0) + uvlc_suffix_len(v1),
                                   uvlc_suffix_len(v0), v0, v1);
                }
            } else if (mode == 4) {
                const int l0 = uvlc_prefix(bits, v0);
                const int l1 = uvlc_prefix(bits >> l0, v1);
                e = uvlc_entry(l0 + l1, uvlc_suffix_len(v0) + uvlc_suffix_len(v1),
                               uvlc_suffix_len(v0), v0 + 2, v1 + 2);
            }
            table[mode * 64 + bits] = e;
        }
}

const HtTables &ht_tables() {
    static HtTables t;
    static std::once_flag once;
    std::call_once(once, [] {
        build_vlc_decode(t.vlc0, kVlcRows0, sizeof(kVlcRows0) / sizeof(kVlcRows0[0]));
        build_vlc_decode(t.vlc1, kVlcRows1, sizeof(kVlcRows1) / sizeof(kVlcRows1[0]));
        build_uvlc(t.uvlc0, true);
        build_uvlc(t.uvlc1, false);
        build_vlc_encode(t.enc[0], kVlcRows0, sizeof(kVlcRows0) / sizeof(kVlcRows0[0]));
        build_vlc_encode(t.enc[1], kVlcRows1, sizeof(kVlcRows1) / sizeof(kVlcRows1[0]));
    });
    return t;
}

}  // namespace sgn
