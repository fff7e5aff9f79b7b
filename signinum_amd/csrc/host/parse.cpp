// Codestream parsing into device plan records.
//
// Restates the reference's host-side stages for HTJ2K:
//   main header        crates/signinum-j2k-native/src/j2c/codestream.rs:38-272,700-960
//   tile parts         j2c/tile.rs:131-320
//   packet headers     j2c/segment.rs:17-133,304-581 (HT, one layer);
//                      bit unstuffing reader.rs:129-160
//   reduced resolution codestream.rs:137-167 (skipped levels),
//                      crates/signinum-j2k/src/decode.rs:12-33 (target size)
// The result is a flat code-block table, per-level synthesis jobs and one
// fused final job per tile.
#include <algorithm>
#include <cstring>

#include "parse.hpp"

namespace sgn {

namespace {

struct ByteReader {
    const uint8_t *d;
    size_t len, pos;
    bool u8(uint32_t &v) {
        if (pos + 1 > len) return false;
        v = d[pos++];
        return true;
    }
    bool u16(uint32_t &v) {
        if (pos + 2 > len) return false;
        v = (uint32_t(d[pos]) << 8) | d[pos + 1];
        pos += 2;
        return true;
    }
    bool u32(uint32_t &v) {
        if (pos + 4 > len) return false;
        v = (uint32_t(d[pos]) << 24) | (uint32_t(d[pos + 1]) << 16) | (uint32_t(d[pos + 2]) << 8) |
            d[pos + 3];
        pos += 4;
        return true;
    }
    int marker() const {
        if (pos + 2 > len || d[pos] != 0xFF) return -1;
        return d[pos + 1];
    }
    bool skip_segment() {
        uint32_t l;
        if (!u16(l) || l < 2 || pos + (l - 2) > len) return false;
        pos += l - 2;
        return true;
    }
};

Status fail(Status s, const char *msg) {
    set_last_error(msg);
    return s;
}

bool read_coding(ByteReader &b, uint8_t scod, CompParams &c) {
    uint32_t v;
    if (!b.u8(v) || v > 32) return false;
    c.levels = uint8_t(v);
    if (!b.u8(v)) return false;
    c.xcb = uint8_t(v + 2);
    if (!b.u8(v)) return false;
    c.ycb = uint8_t(v + 2);
    // A.6.1: each exponent at most 10, a code-block at most 4096 samples
    // (the block kernels size their per-lane state by that limit)
    if (c.xcb > 10 || c.ycb > 10 || c.xcb + c.ycb > 12) return false;
    if (!b.u8(v)) return false;
    c.cbstyle = uint8_t(v);
    if (!b.u8(v) || v > 1) return false;
    c.transform = uint8_t(v);
    for (int r = 0; r <= c.levels; ++r) {
        if (scod & 1) {
            if (!b.u8(v)) return false;
            c.ppx[r] = uint8_t(v & 0xF);
            c.ppy[r] = uint8_t(v >> 4);
        } else {
            c.ppx[r] = 15;
            c.ppy[r] = 15;
        }
    }
    return true;
}

struct Quant {
    uint8_t style = 0, guard = 0;
    uint16_t n = 0;
    uint16_t eps[3 * kMaxRes + 1] = {}, mu[3 * kMaxRes + 1] = {};
};

bool read_quant(ByteReader &b, size_t remaining, Quant &q) {
    uint32_t v;
    if (!b.u8(v)) return false;
    q.style = uint8_t(v & 0x1F);
    q.guard = uint8_t((v >> 5) & 7);
    if (q.style > 2) return false;
    q.n = 0;
    if (q.style == 0) {
        for (size_t i = 0; i < remaining; ++i) {
            if (!b.u8(v)) return false;
            if (q.n < 3 * kMaxRes + 1) {
                q.eps[q.n] = uint16_t(v >> 3);
                q.mu[q.n++] = 0;
            }
        }
    } else if (q.style == 1) {
        if (!b.u16(v)) return false;
        q.eps[0] = uint16_t(v >> 11);
        q.mu[0] = uint16_t(v & 0x7FF);
        q.n = 1;
        if (remaining > 2) b.pos += remaining - 2;
    } else {
        for (size_t i = 0; i < remaining / 2; ++i) {
            if (!b.u16(v)) return false;
            if (q.n < 3 * kMaxRes + 1) {
                q.eps[q.n] = uint16_t(v >> 11);
                q.mu[q.n++] = uint16_t(v & 0x7FF);
            }
        }
    }
    return true;
}

void apply_quant(CompParams &c, const Quant &q) {
    c.qstyle = q.style;
    c.guard = q.guard;
    c.nsteps = q.n;
    std::memcpy(c.eps, q.eps, sizeof(c.eps));
    std::memcpy(c.mu, q.mu, sizeof(c.mu));
}

// Packet-header bit reader with the B.10.1 stuffing rule.
struct PacketBits {
    const uint8_t *d;
    size_t len;
    size_t bit = 0;
    bool get(uint32_t &out) {
        const size_t byte = bit >> 3;
        if (byte >= len) return false;
        const bool stuff = ((bit & 7) == 7) && d[byte] == 0xFF;
        out = (d[byte] >> (7 - (bit & 7))) & 1;
        ++bit;
        if (stuff) {
            const size_t b2 = bit >> 3;
            if (b2 >= len) return false;
            const uint32_t sb = (d[b2] >> (7 - (bit & 7))) & 1;
            ++bit;
            if (sb) return false;
        }
        return true;
    }
    bool bits(uint32_t n, uint32_t &v) {
        v = 0;
        for (uint32_t i = 0; i < n; ++i) {
            uint32_t b;
            if (!get(b)) return false;
            v = (v << 1) | b;
        }
        return true;
    }
};

// Tag-tree decode against `threshold` (B.10.2); stores the leaf value.
int tag_decode(TagTree &t, uint32_t x, uint32_t y, uint32_t threshold, PacketBits &r,
               uint32_t &value) {
    uint32_t low = 0;
    for (int l = int(t.levels) - 1; l >= 0; --l) {
        const uint32_t n = t.leaf(x, y, uint32_t(l));
        if (low > t.low[n])
            t.low[n] = low;
        else
            low = t.low[n];
        while (low < threshold && low < t.value[n]) {
            uint32_t b;
            if (!r.get(b)) return -1;
            if (b)
                t.value[n] = low;
            else
                ++low;
        }
        t.low[n] = low;
    }
    value = t.value[t.leaf(x, y, 0)];
    return value < threshold ? 1 : 0;
}

struct CbState {
    uint32_t zblk = 0, lblock = 3, passes = 0, lcup = 0, lref = 0;
    int64_t data = -1;  // offset of the block data in the codestream
    bool included = false;
    uint32_t layers_in = 0;            // classic: layers that contributed
    std::vector<ClassicPiece> pieces;  // classic: contributions, offsets set from the body
};

// Table B.4: number of coding passes (segment.rs:481-506).
bool classic_passes(PacketBits &r, uint32_t &n) {
    uint32_t b, v;
    if (!r.get(b)) return false;
    if (!b) return n = 1, true;
    if (!r.get(b)) return false;
    if (!b) return n = 2, true;
    if (!r.bits(2, v)) return false;
    if (v < 3) return n = 3 + v, true;
    if (!r.bits(5, v)) return false;
    if (v < 31) return n = 6 + v, true;
    if (!r.bits(7, v)) return false;
    return n = 37 + v, true;
}

// Codeword segment of coding pass p (segment.rs:223-239, Table D.9).
uint32_t classic_segment(const CbState &s, uint8_t style, uint32_t p) {
    if (style & kCbsTermAll) return p;
    if (style & kCbsBypass) return p < 10 ? 0 : 1 + 2 * ((p - 10) / 3) + ((p - 10) % 3 == 2);
    return s.layers_in;
}

// A classic block's header after inclusion (segment.rs:168-301): passes,
// Lblock increment, one length per codeword segment the new passes touch
// (Lblock + floor(log2(its passes)) bits).
bool classic_block_header(PacketBits &r, CbState &s, uint8_t style) {
    uint32_t added, b;
    if (!classic_passes(r, added)) return false;
    for (;;) {
        if (!r.get(b)) return false;
        if (!b) break;
        ++s.lblock;
    }
    const uint32_t prev = s.passes, cum = prev + added;
    if (cum > 1 + 3 * 31) return false;
    uint32_t last = classic_segment(s, style, prev), count = 0;
    for (uint32_t p = prev; p <= cum; ++p) {
        const uint32_t sg = p < cum ? classic_segment(s, style, p) : last + 1;
        if (p < cum && sg == last) {
            ++count;
            continue;
        }
        if (count) {
            const uint32_t nb = s.lblock + (31 - __builtin_clz(count));
            uint32_t len;
            if (nb > 32 || !r.bits(nb, len)) return false;
            ClassicPiece pc;
            pc.len = len;
            pc.seg = last;
            pc.passes = count;
            s.pieces.push_back(pc);
        }
        last = sg;
        count = 1;
    }
    s.passes = cum;
    ++s.layers_in;
    return true;
}

}  // namespace

Status read_main_header(const uint8_t *data, size_t len, MainHeader &h) {
    ByteReader b{data, len, 0};
    h = MainHeader();
    if (b.marker() != 0x4F) return fail(kErrMalformed, "missing SOC marker");
    b.pos += 2;
    if (b.marker() != 0x51) return fail(kErrMalformed, "missing SIZ marker");
    b.pos += 2;
    uint32_t v, csiz;
    if (!b.u16(v) || !b.u16(v) || !b.u32(h.xsiz) || !b.u32(h.ysiz) || !b.u32(h.xo) ||
        !b.u32(h.yo) || !b.u32(h.xt) || !b.u32(h.yt) || !b.u32(h.xto) || !b.u32(h.yto) ||
        !b.u16(csiz))
        return fail(kErrMalformed, "truncated SIZ");
    if (csiz == 0 || h.xo >= h.xsiz || h.yo >= h.ysiz || h.xt == 0 || h.yt == 0)
        return fail(kErrMalformed, "invalid SIZ");
    if (csiz > kMaxComps) return fail(kErrUnsupported, "too many components");
    h.ncomp = uint16_t(csiz);
    uint8_t prec[kMaxComps], sub_x[kMaxComps], sub_y[kMaxComps];
    for (uint32_t i = 0; i < csiz; ++i) {
        uint32_t s, xr, yr;
        if (!b.u8(s) || !b.u8(xr) || !b.u8(yr)) return fail(kErrMalformed, "truncated SIZ");
        prec[i] = uint8_t((s & 0x7F) + 1);
        if (xr == 0 || yr == 0 || prec[i] > 32) return fail(kErrMalformed, "invalid component");
        sub_x[i] = uint8_t(xr);
        sub_y[i] = uint8_t(yr);
        // the signed flag (s & 0x80) changes nothing: the reference decodes
        // signed components exactly as unsigned ones (codestream.rs:647-649)
    }
    // the first tile holds a reference grid point of the image (B-4,
    // codestream.rs:582-596)
    if (uint64_t(h.xto) + h.xt <= h.xo || uint64_t(h.yto) + h.yt <= h.yo || h.xto > h.xo ||
        h.yto > h.yo)
        return fail(kErrMalformed, "invalid tile-grid origin");
    std::vector<std::pair<uint8_t, std::pair<size_t, size_t>>> ppm;
    bool have_cod = false, have_qcd = false;
    CompParams cod;
    uint8_t cod_scod = 0;
    Quant qcd;
    bool coc_set[kMaxComps] = {}, qcc_set[kMaxComps] = {};
    CompParams coc[kMaxComps];
    Quant qcc[kMaxComps];
    for (;;) {
        const int m = b.marker();
        if (m < 0) return fail(kErrMalformed, "bad marker in main header");
        if (m == 0x90) break;
        b.pos += 2;
        if (m == 0x52) {
            uint32_t l, scod, prog, nl, mct;
            if (!b.u16(l) || !b.u8(scod) || !b.u8(prog) || !b.u16(nl) || !b.u8(mct))
                return fail(kErrMalformed, "truncated COD");
            if (prog > 4 || nl == 0 || nl > 32) return fail(kErrMalformed, "invalid COD");
            h.prog = uint8_t(prog);
            h.layers = uint16_t(nl);
            h.mct = mct == 1;
            cod = CompParams();
            if (!read_coding(b, uint8_t(scod), cod)) return fail(kErrMalformed, "invalid COD");
            cod_scod = uint8_t(scod);
            have_cod = true;
        } else if (m == 0x53) {
            uint32_t l, ci, s;
            if (!b.u16(l) || !(csiz < 257 ? b.u8(ci) : b.u16(ci)) || !b.u8(s) || ci >= csiz)
                return fail(kErrMalformed, "invalid COC");
            coc[ci] = CompParams();
            if (!read_coding(b, uint8_t(s), coc[ci])) return fail(kErrMalformed, "invalid COC");
            coc[ci].scod = uint8_t(s);
            coc_set[ci] = true;
        } else if (m == 0x5C) {
            uint32_t l;
            if (!b.u16(l) || l < 3 || !read_quant(b, l - 3, qcd))
                return fail(kErrMalformed, "invalid QCD");
            have_qcd = true;
        } else if (m == 0x5D) {
            uint32_t l, ci;
            if (!b.u16(l) || !(csiz < 257 ? b.u8(ci) : b.u16(ci)) || ci >= csiz)
                return fail(kErrMalformed, "invalid QCC");
            const size_t isz = csiz < 257 ? 1 : 2;
            if (l < 3 + isz || !read_quant(b, l - 3 - isz, qcc[ci]))
                return fail(kErrMalformed, "invalid QCC");
            qcc_set[ci] = true;
        } else if (m == 0x60) {
            // PPM (codestream.rs:761-780): Zppm, then (16-bit length,
            // bytes) pairs; empty pairs are dropped
            uint32_t l, z, nb;
            if (!b.u16(l) || l < 3 || b.pos + (l - 2) > len) return fail(kErrMalformed, "invalid PPM");
            const size_t end = b.pos + (l - 2);
            if (!b.u8(z)) return fail(kErrMalformed, "invalid PPM");
            while (b.pos < end) {
                if (!b.u16(nb) || b.pos + nb > end) return fail(kErrMalformed, "invalid PPM");
                if (nb) ppm.push_back({uint8_t(z), {b.pos, size_t(nb)}});
                b.pos += nb;
            }
        } else if (m == 0x50 || m == 0x59 || m == 0x5E || m == 0x55 || m == 0x64 || m == 0x63) {
            if (!b.skip_segment()) return fail(kErrMalformed, "truncated marker segment");
        } else if (m >= 0x30 && m <= 0x3F) {
        } else {
            return fail(kErrUnsupported, "unsupported main-header marker");
        }
    }
    if (!have_cod || !have_qcd) return fail(kErrMalformed, "missing COD or QCD");
    h.first_sot = b.pos;
    std::stable_sort(ppm.begin(), ppm.end(),
                     [](const auto &a, const auto &c) { return a.first < c.first; });
    for (const auto &e : ppm) h.ppm.push_back(e.second);
    for (uint32_t i = 0; i < csiz; ++i) {
        CompParams &c = h.comp[i];
        if (coc_set[i]) {
            c = coc[i];
            c.scod = uint8_t(coc[i].scod | cod_scod);
        } else {
            c = cod;
            c.scod = cod_scod;
        }
        apply_quant(c, qcc_set[i] ? qcc[i] : qcd);
        c.precision = prec[i];
        c.xr = sub_x[i];
        c.yr = sub_y[i];
        if (c.nsteps == 0) return fail(kErrMalformed, "missing step sizes");
        if (c.qstyle != 1 && c.levels > 0 && 1 + (c.levels - 1) * 3 + 2 >= c.nsteps)
            return fail(kErrMalformed, "insufficient step sizes");
    }
    return kOk;
}

uint32_t skipped_levels(const MainHeader &h, uint32_t denominator) {
    if (denominator <= 1) return 0;
    const Sampling smp = sampling_of(h);
    const uint32_t iw = ceil_div(h.xsiz - h.xo, smp.sx), ih = ceil_div(h.ysiz - h.yo, smp.sy);
    const uint32_t tw = ceil_div(iw, denominator);
    const uint32_t th = ceil_div(ih, denominator);
    // image_width / target (codestream.rs:149-156)
    const uint32_t wr = iw / tw, hr = ih / th;
    const uint32_t wl = wr ? 31 - __builtin_clz(wr) : 0, hl = hr ? 31 - __builtin_clz(hr) : 0;
    uint32_t s = std::min(wl, hl);
    uint32_t min_nr = 255;
    for (int c = 0; c < h.ncomp; ++c) min_nr = std::min<uint32_t>(min_nr, h.comp[c].levels + 1);
    return std::min(s, min_nr - 1);
}

Status parse_tiles(const uint8_t *data, size_t len, const MainHeader &h,
                   std::vector<TileBody> &tiles) {
    const uint32_t ntx = h.tiles_x(), nty = h.tiles_y();
    // SOT's Isot is 16 bits (A.4.2): a grid of more tiles cannot be coded
    if (uint64_t(ntx) * nty > 65535u) return fail(kErrMalformed, "more than 65535 tiles");
    tiles.assign(size_t(ntx) * nty, TileBody());
    for (uint32_t i = 0; i < ntx * nty; ++i) {
        TileBody &t = tiles[i];
        const uint32_t x0 = h.xto + (i % ntx) * h.xt, y0 = h.yto + (i / ntx) * h.yt;
        t.rect = {std::max(x0, h.xo), std::max(y0, h.yo), std::min(x0 + h.xt, h.xsiz),
                  std::min(y0 + h.yt, h.ysiz)};
        for (int c = 0; c < h.ncomp; ++c) t.comp[c] = h.comp[c];
        t.prog = h.prog;
        t.mct = h.mct;
        t.layers = h.layers;
    }
    ByteReader b{data, len, h.first_sot};
    size_t tp_index = 0;  // tile-parts seen, in codestream order
    while (b.marker() == 0x90) {
        b.pos += 2;
        std::vector<std::pair<uint8_t, std::pair<size_t, size_t>>> ppt;
        uint32_t lsot, isot, psot, tp, tn;
        if (!b.u16(lsot) || !b.u16(isot) || !b.u32(psot) || !b.u8(tp) || !b.u8(tn))
            return fail(kErrMalformed, "truncated SOT");
        if (isot >= tiles.size()) return fail(kErrMalformed, "tile index out of range");
        const size_t start = b.pos;
        size_t dlen;
        if (psot == 0)
            dlen = len - b.pos;
        else if (psot < 12)
            return fail(kErrMalformed, "invalid Psot");
        else
            dlen = psot - 12;
        TileBody &t = tiles[isot];
        bool sod = false;
        for (;;) {
            const int m = b.marker();
            if (m < 0) return kOk;  // non-strict: stop at a broken tile-part
            b.pos += 2;
            if (m == 0x93) {
                sod = true;
                break;
            }
            if (m == 0x52 || m == 0x53) {
                uint32_t l, ci = 0, scod, prog = 0, nl = 0, mct = 0;
                if (!b.u16(l)) return fail(kErrMalformed, "truncated COD/COC");
                if (m == 0x52) {
                    if (!b.u8(scod) || !b.u8(prog) || !b.u16(nl) || !b.u8(mct) || prog > 4)
                        return fail(kErrMalformed, "invalid tile COD");
                } else if (!(h.ncomp < 257 ? b.u8(ci) : b.u16(ci)) || !b.u8(scod) ||
                           ci >= h.ncomp) {
                    return fail(kErrMalformed, "invalid tile COC");
                }
                CompParams cc;
                if (!read_coding(b, uint8_t(scod), cc)) return fail(kErrMalformed, "invalid COD");
                if (m == 0x52) {
                    t.mct = mct == 1;
                    t.layers = uint16_t(nl);
                    t.prog = uint8_t(prog);
                }
                for (int c = 0; c < h.ncomp; ++c) {
                    if (m == 0x53 && uint32_t(c) != ci) continue;
                    CompParams &tc = t.comp[c];
                    tc.scod = uint8_t(tc.scod | scod);
                    tc.levels = cc.levels;
                    tc.xcb = cc.xcb;
                    tc.ycb = cc.ycb;
                    tc.cbstyle = cc.cbstyle;
                    tc.transform = cc.transform;
                    std::memcpy(tc.ppx, cc.ppx, sizeof(cc.ppx));
                    std::memcpy(tc.ppy, cc.ppy, sizeof(cc.ppy));
                }
            } else if (m == 0x5C || m == 0x5D) {
                uint32_t l, ci = 0;
                if (!b.u16(l)) return fail(kErrMalformed, "truncated QCD/QCC");
                size_t isz = 0;
                if (m == 0x5D) {
                    if (!(h.ncomp < 257 ? b.u8(ci) : b.u16(ci)) || ci >= h.ncomp)
                        return fail(kErrMalformed, "invalid tile QCC");
                    isz = h.ncomp < 257 ? 1 : 2;
                }
                Quant q;
                if (l < 3 + isz || !read_quant(b, l - 3 - isz, q))
                    return fail(kErrMalformed, "invalid tile QCD");
                for (int c = 0; c < h.ncomp; ++c)
                    if (m == 0x5C || uint32_t(c) == ci) apply_quant(t.comp[c], q);
            } else if (m == 0xD9) {
                break;
            } else if (m == 0x61) {
                // PPT (tile.rs:251-258,697-706); not beside PPM
                uint32_t l, z;
                if (!h.ppm.empty()) return fail(kErrMalformed, "PPM and PPT in one codestream");
                if (!b.u16(l) || l < 3 || b.pos + (l - 2) > len || !b.u8(z))
                    return fail(kErrMalformed, "invalid PPT");
                ppt.push_back({uint8_t(z), {b.pos, size_t(l - 3)}});
                b.pos += l - 3;
            } else if (m == 0x58 || m == 0x64) {
                if (!b.skip_segment()) return fail(kErrMalformed, "truncated marker segment");
            } else if (m >= 0x30 && m <= 0x3F) {
            } else {
                return fail(kErrUnsupported, "unsupported tile-part marker");
            }
        }
        if (!sod) return kOk;
        const size_t used = b.pos - start;
        if (used > dlen) return kOk;
        const size_t remain = dlen - used;
        if (b.pos + remain > len) return fail(kErrMalformed, "tile-part exceeds codestream");
        TilePart part;
        part.off = b.pos;
        part.len = remain;
        std::stable_sort(ppt.begin(), ppt.end(),
                         [](const auto &a, const auto &c) { return a.first < c.first; });
        for (const auto &e : ppt) part.hdr.push_back(e.second);
        if (tp_index < h.ppm.size()) part.hdr.push_back(h.ppm[tp_index]);
        ++tp_index;
        t.parts.push_back(std::move(part));
        b.pos += remain;
    }
    return kOk;
}

// Packet headers of one tile (HT, single layer).  Fills the code-block
// states in TileCompGeom::cbs order per component; offsets are relative to
// the codestream start.  Tile-parts are read one after another with the
// progression continuing (segment.rs:36-133): while the header stream is not
// at its end, the next packet; an SOP marker in the body is skipped; the
// header comes from the body, or with packed headers from the tile-part's
// header pieces (the next piece once the current one is exhausted,
// tile.rs:55-67); the header is aligned, followed by EPH when signalled,
// and the included blocks' bytes come from the body.  Non-strict: a
// malformed packet ends its tile-part; the next tile-part continues
// (segment.rs:36-50).
Status parse_packets(const uint8_t *cs, const TileBody &t, int ncomp, const TileCompGeom *tcs,
                     std::vector<std::vector<CbSeg>> &segs) {
    std::vector<PacketId> order;
    Status st = progression(t.prog, t.layers, t.rect, tcs, ncomp, order);
    if (st) return st;
    segs.assign(static_cast<size_t>(ncomp), {});
    std::vector<std::vector<CbState>> state(static_cast<size_t>(ncomp));
    std::vector<std::vector<TagTree>> incl(static_cast<size_t>(ncomp)), zbp(static_cast<size_t>(ncomp));
    for (int c = 0; c < ncomp; ++c) {
        state[c].assign(tcs[c].cbs.size(), CbState());
        incl[c].resize(tcs[c].precs.size());
        zbp[c].resize(tcs[c].precs.size());
        for (size_t p = 0; p < tcs[c].precs.size(); ++p) {
            incl[c][p].init(tcs[c].precs[p].ncbx, tcs[c].precs[p].ncby);
            zbp[c][p].init(tcs[c].precs[p].ncbx, tcs[c].precs[p].ncby);
        }
    }
    size_t next = 0;  // next packet of the progression
    for (const TilePart &tp : t.parts) {
        const uint8_t *body = cs + tp.off;
        const size_t blen = tp.len;
        const bool packed = !tp.hdr.empty();
        size_t hi = 0;
        PacketBits r = packed ? PacketBits{cs + tp.hdr[0].first, tp.hdr[0].second, 0}
                              : PacketBits{body, blen, 0};
        size_t bpos = 0;  // body byte position (packed headers)
        for (;;) {
            if (packed && (r.bit >> 3) >= r.len && hi + 1 < tp.hdr.size()) {
                ++hi;
                r = PacketBits{cs + tp.hdr[hi].first, tp.hdr[hi].second, 0};
            }
            if ((r.bit >> 3) >= r.len) break;
            if (next >= order.size()) break;
            const PacketId &pk = order[next++];
            const TileCompGeom &g = tcs[pk.comp];
            const ResGeom &rs = g.res[pk.res];
            const CompParams &cp = t.comp[pk.comp];
            bool ok = true;
            {
                const size_t at = packed ? bpos : (r.bit >> 3);
                if ((cp.scod & 2) && at + 1 < blen && body[at] == 0xFF && body[at + 1] == 0x91) {
                    if (at + 6 > blen) break;
                    if (packed)
                        bpos += 6;
                    else
                        r.bit += 48;
                }
            }
            uint32_t nonzero;
            if (!r.get(nonzero)) break;
            if (nonzero) {
                for (uint32_t bi = 0; bi < rs.nbands && ok; ++bi) {
                    const uint32_t pi = rs.prec_base + pk.prec * rs.nbands + bi;
                    const PrecBand &pb = g.precs[pi];
                    for (uint32_t k = 0; k < pb.ncbx * pb.ncby && ok; ++k) {
                        CbState &s = state[pk.comp][pb.cb_base + k];
                        const CbGeom &cg = g.cbs[pb.cb_base + k];
                        const bool first = !s.included;
                        uint32_t inc, v;
                        if (s.included) {
                            if (!r.get(inc)) { ok = false; break; }
                        } else {
                            const int rc = tag_decode(incl[pk.comp][pi], cg.xi, cg.yi, pk.layer + 1u, r, v);
                            if (rc < 0) { ok = false; break; }
                            inc = v <= pk.layer;
                        }
                        if (!inc) continue;
                        if (first) {
                            uint32_t thr = 1;
                            for (;;) {
                                const int rc = tag_decode(zbp[pk.comp][pi], cg.xi, cg.yi, thr, r, v);
                                if (rc < 0) { ok = false; break; }
                                if (rc) break;
                                if (++thr > 255) { ok = false; break; }
                            }
                            if (!ok) break;
                            s.zblk = v;
                        }
                        s.included = true;
                        if (!(cp.cbstyle & kCbsHt)) {
                            if (!classic_block_header(r, s, cp.cbstyle)) { ok = false; break; }
                            continue;
                        }
                        if (!first || s.passes != 0) { ok = false; break; }
                        // HT pass count codeword (segment.rs:508-530)
                        uint32_t np = 1, b1;
                        if (!r.get(b1)) { ok = false; break; }
                        if (b1) {
                            np = 2;
                            if (!r.get(b1)) { ok = false; break; }
                            if (b1) {
                                uint32_t e;
                                if (!r.bits(2, e)) { ok = false; break; }
                                np = 3 + e;
                                if (e == 3) {
                                    if (!r.bits(5, e)) { ok = false; break; }
                                    np = 6 + e;
                                    if (e == 31) {
                                        if (!r.bits(7, e)) { ok = false; break; }
                                        np = 37 + e;
                                    }
                                }
                            }
                        }
                        if (np > 94) { ok = false; break; }
                        // placeholder passes and lengths (segment.rs:542-581)
                        const uint32_t groups = (np - 1) / 3;
                        s.zblk += groups;
                        const uint32_t ph = groups * 3, actual = np - ph;
                        for (;;) {
                            uint32_t bb;
                            if (!r.get(bb)) { ok = false; break; }
                            if (!bb) break;
                            ++s.lblock;
                        }
                        if (!ok) break;
                        const uint32_t nb = s.lblock + (31 - __builtin_clz(ph + 1));
                        uint32_t lc = 0, lr = 0;
                        if (nb > 32 || !r.bits(nb, lc) || lc < 2 || lc >= 65535) { ok = false; break; }
                        if (actual > 1) {
                            const uint32_t nb2 = s.lblock + (actual > 2 ? 1 : 0);
                            if (nb2 > 32 || !r.bits(nb2, lr) || lr >= 2047) { ok = false; break; }
                        }
                        s.passes = actual;
                        s.lcup = lc;
                        s.lref = lr;
                    }
                }
            }
            if (!ok) break;
            if (r.bit & 7) r.bit += 8 - (r.bit & 7);
            if (cp.scod & 4) {
                const size_t bb = r.bit >> 3;
                if (bb + 2 > r.len || r.d[bb] != 0xFF || r.d[bb + 1] != 0x92) break;
                r.bit += 16;
            }
            if (!packed) bpos = r.bit >> 3;
            if (nonzero) {
                for (uint32_t bi = 0; bi < rs.nbands && ok; ++bi) {
                    const PrecBand &pb = g.precs[rs.prec_base + pk.prec * rs.nbands + bi];
                    for (uint32_t k = 0; k < pb.ncbx * pb.ncby; ++k) {
                        CbState &s = state[pk.comp][pb.cb_base + k];
                        if (!(cp.cbstyle & kCbsHt)) {
                            // this packet's contributions: the ones without an offset yet
                            for (ClassicPiece &pc : s.pieces) {
                                if (pc.offset >= 0) continue;
                                if (bpos + pc.len > blen) { ok = false; break; }
                                pc.offset = int64_t(tp.off + bpos);
                                bpos += pc.len;
                            }
                            if (!ok) break;
                            continue;
                        }
                        if (s.passes == 0 || s.data >= 0) continue;
                        const size_t need = size_t(s.lcup) + s.lref;
                        if (bpos + need > blen) { ok = false; break; }
                        s.data = int64_t(tp.off + bpos);
                        bpos += need;
                    }
                }
            }
            if (!ok) break;
            if (!packed) r.bit = bpos * 8;
        }
    }
    for (int c = 0; c < ncomp; ++c) {
        segs[c].resize(state[c].size());
        for (size_t k = 0; k < state[c].size(); ++k) {
            const CbState &s = state[c][k];
            CbSeg &o = segs[c][k];
            o.passes = s.data >= 0 ? s.passes : 0;
            if (!(t.comp[c].cbstyle & kCbsHt)) {
                // contributions whose packet body was never read have no bytes
                o.passes = s.passes;
                o.pieces = s.pieces;
            }
            o.zblk = s.zblk;
            o.lcup = s.lcup;
            o.lref = s.lref;
            o.offset = s.data >= 0 ? s.data : -1;
        }
    }
    return kOk;
}

}  // namespace sgn

namespace sgn {

// Packet tables of one tile (HT, one layer, headers in the bodies) for the
// device header parser: the same walk parse_packets makes, laid out flat.
Status packet_walk(const TileBody &t, int ncomp, const TileCompGeom *tcs, LocalPlan &lp,
                   std::vector<std::vector<uint32_t>> &cb_index) {
    std::vector<PacketId> order;
    Status st = progression(t.prog, t.layers, t.rect, tcs, ncomp, order);
    if (st) return st;
    PktTile T{};
    T.part0 = static_cast<uint32_t>(lp.pk_parts.size());
    T.nparts = static_cast<uint32_t>(t.parts.size());
    for (const TilePart &tp : t.parts)
        lp.pk_parts.push_back({static_cast<uint32_t>(tp.off), static_cast<uint32_t>(tp.len)});
    auto tree_nodes = [](uint32_t w, uint32_t h) {
        uint32_t n = 0;
        if (!w || !h) return n;
        for (;;) {
            n += w * h;
            if (w == 1 && h == 1) return n;
            w = (w + 1) / 2;
            h = (h + 1) / 2;
        }
    };
    // per (component, precinct-band): its first PktCb and its two trees
    T.cb0 = static_cast<uint32_t>(lp.pk_cbs.size());
    T.node0 = lp.pk_nodes;
    uint32_t nodes = 0;
    cb_index.assign(static_cast<size_t>(ncomp), {});
    std::vector<std::vector<PktBand>> pbs(static_cast<size_t>(ncomp));
    for (int c = 0; c < ncomp; ++c) {
        const TileCompGeom &g = tcs[c];
        cb_index[c].assign(g.cbs.size(), kPktNoJob);
        pbs[c].resize(g.precs.size());
        for (size_t p = 0; p < g.precs.size(); ++p) {
            const PrecBand &pb = g.precs[p];
            PktBand &b = pbs[c][p];
            b.cb0 = static_cast<uint32_t>(lp.pk_cbs.size());
            b.ncbx = static_cast<uint16_t>(pb.ncbx);
            b.ncby = static_cast<uint16_t>(pb.ncby);
            for (uint32_t k = 0; k < pb.ncbx * pb.ncby; ++k) {
                cb_index[c][pb.cb_base + k] = static_cast<uint32_t>(lp.pk_cbs.size());
                lp.pk_cbs.push_back({kPktNoJob, kPktNoJob, 0});
            }
            const uint32_t tn = tree_nodes(pb.ncbx, pb.ncby);
            b.incl = T.node0 + nodes;
            b.zbp = b.incl + tn;
            nodes += 2 * tn;
        }
    }
    T.pkt0 = static_cast<uint32_t>(lp.pk_pkts.size());
    for (const PacketId &pk : order) {
        const TileCompGeom &g = tcs[pk.comp];
        const ResGeom &rs = g.res[pk.res];
        const CompParams &cp = t.comp[pk.comp];
        PktPacket P{};
        P.band0 = static_cast<uint32_t>(lp.pk_bands.size());
        P.nbands = rs.nbands;
        P.sop = (cp.scod & 2) ? 1 : 0;
        P.eph = (cp.scod & 4) ? 1 : 0;
        for (uint32_t bi = 0; bi < rs.nbands; ++bi)
            lp.pk_bands.push_back(pbs[pk.comp][rs.prec_base + pk.prec * rs.nbands + bi]);
        lp.pk_pkts.push_back(P);
    }
    T.npkts = static_cast<uint32_t>(lp.pk_pkts.size()) - T.pkt0;
    T.nnodes = nodes;
    T.ncbs = static_cast<uint32_t>(lp.pk_cbs.size()) - T.cb0;
    lp.pk_nodes += nodes;
    lp.pk_tiles.push_back(T);
    return kOk;
}

}  // namespace sgn
