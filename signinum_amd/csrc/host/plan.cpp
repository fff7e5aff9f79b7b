// Plan building: one codestream -> code-block table, synthesis-level jobs
// and fused final jobs.  Mirrors the data flow of the reference's
// decode_tile (crates/signinum-j2k-native/src/j2c/decode.rs:796-870):
// sub-band decode of the retained resolutions, per-level synthesis
// (idwt.rs:54-193), store with sign shift and inverse MCT (decode.rs:
// 1720-1902, mct.rs:11-156), then output conversion (lib.rs:1772-1835,
// 2202-2295).
#include <algorithm>

#include "parse.hpp"
#include "sgn_api.hpp"

namespace sgn {

static int fmt_channels(int fmt) {
    switch (fmt) {
    case SGN_GRAY8:
    case SGN_GRAY16:
        return 1;
    case SGN_RGB8:
    case SGN_RGB16:
        return 3;
    case SGN_RGBA8:
        return 4;
    default:
        return 0;
    }
}

Sampling sampling_of(const MainHeader &h) {
    Sampling s;
    bool same = true;
    for (int c = 0; c < h.ncomp; ++c) {
        s.any |= h.comp[c].xr != 1 || h.comp[c].yr != 1;
        same &= h.comp[c].xr == h.comp[0].xr && h.comp[c].yr == h.comp[0].yr;
    }
    s.mixed = s.any && !same;
    // the output grid: the components' own when they share one
    // (codestream.rs:663-683)
    s.sx = same ? h.comp[0].xr : 1;
    s.sy = same ? h.comp[0].yr : 1;
    return s;
}

Status output_geometry(const MainHeader &h, int fmt, uint32_t denominator, OutputGeom &g) {
    const int ch = fmt_channels(fmt);
    if (!ch) {
        set_last_error("unknown pixel format");
        return kErrArg;
    }
    if (denominator != 1 && denominator != 2 && denominator != 4 && denominator != 8) {
        set_last_error("downscale denominator must be 1, 2, 4 or 8");
        return kErrArg;
    }
    if ((ch == 1 && h.ncomp != 1) || (ch >= 3 && h.ncomp != 3)) {
        set_last_error("colour space cannot be mapped to the requested pixel format");
        return kErrUnsupported;
    }
    const uint32_t sk = skipped_levels(h, denominator);
    const Sampling smp = sampling_of(h);
    // a reduced resolution of sub-sampled components is stored with
    // full-resolution component offsets (decode.rs:1858-1898)
    if (sk && smp.any) {
        set_last_error("reduced resolution of sub-sampled components");
        return kErrUnsupported;
    }
    // The reference's store places reduced-resolution tiles by subtracting
    // the full-resolution image offset from reduced coordinates
    // (decode.rs:1762-1790): with a non-zero image origin only full
    // resolution has a well-defined output.
    if (sk && (h.xo || h.yo)) {
        set_last_error("reduced resolution of an image with a non-zero origin");
        return kErrUnsupported;
    }
    // image_width / image_height (codestream.rs:544-553)
    g.width = ceil_div(h.xsiz - h.xo, smp.sx << sk);
    g.height = ceil_div(h.ysiz - h.yo, smp.sy << sk);
    g.ncomp = h.ncomp;
    g.channels = uint32_t(ch);
    g.bytes_per_sample = (fmt == SGN_GRAY16 || fmt == SGN_RGB16) ? 2 : 1;
    g.row_bytes = size_t(g.width) * ch * g.bytes_per_sample;
    return kOk;
}

Status region_output(const MainHeader &h, const OutputGeom &og, uint32_t denominator,
                     const Region &roi, Rect &out) {
    if (!roi.on) {
        out = {0, 0, og.width, og.height};
        return kOk;
    }
    const Sampling smp = sampling_of(h);
    const uint64_t fw = ceil_div(h.xsiz - h.xo, smp.sx), fh = ceil_div(h.ysiz - h.yo, smp.sy);
    if (uint64_t(roi.x) + roi.w > fw || uint64_t(roi.y) + roi.h > fh) {
        set_last_error("region outside the image");
        return kErrRegion;
    }
    const uint32_t d = denominator ? denominator : 1;
    out.x0 = roi.x / d;
    out.y0 = roi.y / d;
    out.x1 = ceil_div(roi.x + roi.w, d);
    out.y1 = ceil_div(roi.y + roi.h, d);
    if (out.x1 > og.width || out.y1 > og.height) {
        set_last_error("scaled region outside the decoded image");
        return kErrRegion;
    }
    return kOk;
}

namespace {

// Synthesis inputs of output samples [a, b) of one level, as sample
// indices of its low- or high-pass band: index k sits at 2k (low) or 2k + 1
// (high), and the 9/7 synthesis reaches two coefficients to either side of
// a sample (5/3: one); a margin of four covers both with room to spare.
constexpr uint32_t kRegionMargin = 4;

void band_need(uint32_t a, uint32_t b, uint32_t lo, uint32_t hi, uint32_t &na, uint32_t &nb) {
    const uint32_t x0 = a / 2, x1 = (b + 1) / 2 + kRegionMargin;
    na = std::max(lo, x0 > kRegionMargin ? x0 - kRegionMargin : 0u);
    nb = std::min(hi, x1);
    if (nb < na) nb = na;
}

Rect need_of(const Rect &n, const Rect &band) {
    Rect r;
    band_need(n.x0, n.x1, band.x0, band.x1, r.x0, r.x1);
    band_need(n.y0, n.y1, band.y0, band.y1, r.y0, r.y1);
    return r;
}

bool meets(const Rect &a, const Rect &b) {
    return a.x0 < b.x1 && b.x0 < a.x1 && a.y0 < b.y1 && b.y0 < a.y1;
}

}  // namespace

// A classic code-block's job (decode.rs:1189-1343; bitplane.rs:90-186):
// its contributions joined in codestream order (in place when they are
// contiguous there), and with selective bypass or termination on each pass
// the codeword segments, split where the segment index changes; a segment
// is raw when it starts after pass 9 on a significance-propagation or
// refinement pass.
static Status classic_job(const uint8_t *cs, const CompParams &cp, const BandGeom &bg, const CbGeom &cg,
                   const CbSeg &sg, uint32_t K, uint32_t band_off, LocalPlan &lp) {
    ClassicJob j{};
    j.stride = coef_pitch(bg.rect.w());
    j.dst = band_off + (cg.rect.y0 - bg.rect.y0) * j.stride + (cg.rect.x0 - bg.rect.x0);
    j.w = static_cast<uint16_t>(cg.rect.w());
    j.h = static_cast<uint16_t>(cg.rect.h());
    if (j.w == 0 || j.h == 0) return kOk;
    // the decoder's per-block row store: (w + 3) / 2 dwords per stripe, at
    // most 768 (a 4x1024 block)
    if (((uint32_t(j.w) + 3) >> 1) * ((uint32_t(j.h) + 3) >> 2) > 768 || j.w > 1024) {
        set_last_error("classic code-block shape too large");
        return kErrUnsupported;
    }
    if (K > 32) {
        set_last_error("too many bit-planes");
        return kErrMalformed;
    }
    const uint32_t planes = K > sg.zblk ? K - sg.zblk : 0;  // non-strict: saturating
    if (planes > 31) {
        set_last_error("32 magnitude bit-planes in a classic code-block");
        return kErrUnsupported;
    }
    j.planes = static_cast<uint8_t>(planes);
    j.passes = static_cast<uint8_t>(sg.passes);
    j.band = static_cast<uint8_t>(bg.type & 3u);
    j.style = static_cast<uint8_t>(cp.cbstyle & 0x3F);
    // data: in place when the contributions follow each other
    int64_t first = -1, end = -1;
    bool contiguous = true;
    uint64_t total = 0;
    for (const ClassicPiece &pc : sg.pieces) {
        if (pc.offset < 0 || pc.len == 0) continue;
        if (first < 0) first = pc.offset;
        else if (pc.offset != end) contiguous = false;
        end = pc.offset + pc.len;
        total += pc.len;
    }
    if (total >= (1ull << 31)) {
        set_last_error("classic code-block data too long");
        return kErrMalformed;
    }
    j.len = static_cast<uint32_t>(total);
    if (contiguous || total == 0) {
        j.src = first < 0 ? 0u : static_cast<uint32_t>(first);
    } else {
        j.src = static_cast<uint32_t>(lp.joined.size());
        j.style |= kClassicJoined;
        for (const ClassicPiece &pc : sg.pieces)
            if (pc.offset >= 0) lp.joined.insert(lp.joined.end(), cs + pc.offset, cs + pc.offset + pc.len);
    }
    if (j.style & (kCbsBypass | kCbsTermAll)) {
        j.seg = static_cast<uint32_t>(lp.cseg.size());
        uint32_t at = 0, passes = 0;
        for (size_t q = 0; q < sg.pieces.size(); ++q) {
            const ClassicPiece &pc = sg.pieces[q];
            if (q == 0 || pc.seg != sg.pieces[q - 1].seg) {
                ClassicSeg s{};
                s.off = at;
                s.start = static_cast<uint8_t>(passes);
                s.raw = (j.style & kCbsBypass) && passes > 9 && passes % 3 != 0;
                lp.cseg.push_back(s);
            }
            ClassicSeg &s = lp.cseg.back();
            if (pc.offset >= 0) {
                at += pc.len;
                s.len += pc.len;
            }
            passes += pc.passes;
            s.end = static_cast<uint8_t>(passes);
        }
        j.nseg = static_cast<uint32_t>(lp.cseg.size()) - j.seg;
    }
    j.cost = j.planes && j.passes ? j.len + uint32_t(j.w) * j.h * j.passes / 8 : 0;
    lp.classic.push_back(j);
    return kOk;
}

Status build_local_plan(const uint8_t *cs, size_t len, int fmt, uint32_t denominator,
                        uint64_t dst, uint64_t pitch, uint64_t capacity, LocalPlan &lp,
                        const Region &roi, bool device_packets) {
    lp = LocalPlan();
    MainHeader h;
    Status st = read_main_header(cs, len, h);
    if (st) return st;
    OutputGeom og;
    st = output_geometry(h, fmt, denominator, og);
    if (st) return st;
    // the output: the whole decoded image or the region's covering rectangle
    Rect reg;
    st = region_output(h, og, denominator, roi, reg);
    if (st) return st;
    const size_t row_bytes = size_t(reg.w()) * og.channels * og.bytes_per_sample;
    lp.width = reg.w();
    lp.height = reg.h();
    lp.row_bytes = row_bytes;
    lp.out_bytes = uint64_t(row_bytes) * reg.h();
    if (pitch < row_bytes || (reg.h() && capacity < pitch * (reg.h() - 1) + row_bytes)) {
        set_last_error("output surface too small for the decoded image");
        return kErrBuffer;
    }
    if (roi.on && (reg.w() == 0 || reg.h() == 0)) return kOk;  // an empty region
    const Sampling smp = sampling_of(h);
    // the image origin on the output grid; with components sharing a
    // sub-sampling it must fall on that grid (else the reference's store
    // replicates an edge sample, decode.rs:1858-1898)
    const uint32_t xoff = ceil_div(h.xo, smp.sx), yoff = ceil_div(h.yo, smp.sy);
    // an image origin off a component's sampling grid leaves the first
    // columns / rows without a sample of it (the reference writes zeros
    // there, decode.rs:1858-1898): not decoded here
    for (int c = 0; c < h.ncomp; ++c)
        if (h.xo % h.comp[c].xr || h.yo % h.comp[c].yr) {
            set_last_error("image origin off a component's sampling grid");
            return kErrUnsupported;
        }
    // the region on the (component) grid at the decoded resolution (a
    // non-zero image origin only at full resolution)
    const Rect regA{reg.x0 + xoff, reg.y0 + yoff, reg.x1 + xoff, reg.y1 + yoff};
    // region decodes skip code-blocks and synthesis blocks, except with
    // sub-sampled components (roi.rs:24-30: decoded whole, then cropped)
    const bool prune = roi.on && !smp.any;
    const uint32_t sk = skipped_levels(h, denominator);
    std::vector<TileBody> tiles;
    st = parse_tiles(cs, len, h, tiles);
    if (st) return st;
    const int nc = h.ncomp;
    const uint32_t bpp = og.channels * og.bytes_per_sample;
    bool same8 = true;
    for (int c = 0; c < nc; ++c) same8 &= h.comp[c].precision == 8;

    std::vector<TileCompGeom> tcs(static_cast<size_t>(nc));
    std::vector<std::vector<CbSeg>> segs;
    // per component: the samples each retained resolution must deliver
    // (region decodes), absolute coordinates of that resolution
    std::vector<std::vector<Rect>> need(static_cast<size_t>(nc));
    for (const TileBody &t : tiles) {
        for (int c = 0; c < nc; ++c) {
            // several quality layers: classic code-blocks only (an HT block
            // included twice is malformed in the packet parser)
            if (t.layers > 1 && (t.comp[c].cbstyle & kCbsHt)) {
                set_last_error("multi-layer HTJ2K packet assembly");
                return kErrUnsupported;
            }
            st = build_tile_comp(t.comp[c], t.rect, tcs[c]);
            if (st) return st;
            if (tcs[c].nres <= sk) {
                set_last_error("resolution levels below the requested scale");
                return kErrMalformed;
            }
        }
        if (prune) {
            // a tile outside the region is not decoded at all
            const Rect F = tcs[0].res[tcs[0].nres - sk - 1].rect;
            if (!meets(F, regA)) {
                for (int c = 0; c < nc; ++c) lp.pruned_cbs += tcs[c].cbs.size();
                continue;
            }
            for (int c = 0; c < nc; ++c) {
                const TileCompGeom &g = tcs[c];
                const uint32_t R = g.nres - sk;
                need[c].assign(R, Rect());
                const Rect fr = g.res[R - 1].rect;
                need[c][R - 1] = {std::max(fr.x0, regA.x0), std::max(fr.y0, regA.y0),
                                  std::min(fr.x1, regA.x1), std::min(fr.y1, regA.y1)};
                for (uint32_t r = R - 1; r >= 1; --r)
                    need[c][r - 1] = need_of(need[c][r], g.res[r - 1].rect);
            }
        }
        // packet headers on the device: HT code-blocks, one layer, headers
        // in the tile-part bodies (no PPM / PPT)
        bool dev = device_packets && t.layers == 1;
        for (int c = 0; c < nc && dev; ++c) dev = (t.comp[c].cbstyle & kCbsHt) != 0;
        for (const TilePart &tp : t.parts) dev = dev && tp.hdr.empty();
        std::vector<std::vector<uint32_t>> pk_index;
        if (dev) {
            st = packet_walk(t, nc, tcs.data(), lp, pk_index);
            segs.assign(static_cast<size_t>(nc), {});
            for (int c = 0; c < nc; ++c) segs[c].resize(tcs[c].cbs.size());
        } else {
            st = parse_packets(cs, t, nc, tcs.data(), segs);
        }
        if (st) return st;
        const bool mct = t.mct && nc >= 3;
        if (mct && smp.mixed) {
            set_last_error("colour transform over components on different sampling grids");
            return kErrUnsupported;
        }
        if (mct && (t.comp[0].transform != t.comp[1].transform ||
                    t.comp[1].transform != t.comp[2].transform)) {
            set_last_error("inverse MCT over components with different transforms");
            return kErrMalformed;
        }
        // A retained level one sample wide (or high) at an odd origin halves
        // that sample in f32 (idwt.rs:726-731), which later levels carry as a
        // half: every reversible component of such a tile is then lifted in
        // f32 with the 5/3 steps, as the reference computes it throughout.
        bool lone53 = false;
        for (int c = 0; c < nc; ++c) {
            if (t.comp[c].transform != 1) continue;
            const uint32_t R = tcs[c].nres - sk;
            for (uint32_t r = 1; r < R; ++r) {
                const Rect &rr = tcs[c].res[r].rect;
                lone53 |= (rr.w() == 1 && (rr.x0 & 1)) || (rr.h() == 1 && (rr.y0 & 1));
            }
        }
        FinalJob fj{};
        Rect out_rect;
        for (int c = 0; c < nc; ++c) {
            const TileCompGeom &g = tcs[c];
            const CompParams &cp = t.comp[c];
            const uint32_t R = g.nres - sk;
            const bool rev = cp.transform == 1 && !lone53;
            const bool f53 = cp.transform == 1 && lone53;
            const uint64_t comp_start = lp.coef_elems;
            // coefficient layout: bands of the retained resolutions
            std::vector<uint32_t> band_off(g.bands.size(), 0);
            std::vector<float> band_delta(g.bands.size(), 1.0f);
            std::vector<uint32_t> band_k(g.bands.size(), 0);
            for (size_t b = 0; b < g.bands.size(); ++b) {
                const BandGeom &bg = g.bands[b];
                if (bg.res >= R) continue;
                band_off[b] = static_cast<uint32_t>(lp.coef_elems);
                lp.coef_elems += uint64_t(coef_pitch(bg.rect.w())) * bg.rect.h();
                uint32_t K;
                float d;
                st = band_quant(cp, bg.type, bg.res, K, d);
                if (st) {
                    set_last_error("invalid quantisation parameters");
                    return st;
                }
                if (cp.transform == 1 && d != 1.0f) {
                    set_last_error("quantised reversible transform");
                    return kErrUnsupported;
                }
                band_delta[b] = d;
                band_k[b] = K;
                lp.max_k = std::max(lp.max_k, K);
            }
            // code-blocks (decode.rs:1117-1187,1554-1718)
            for (size_t k = 0; k < g.cbs.size(); ++k) {
                const CbGeom &cg = g.cbs[k];
                const BandGeom &bg = g.bands[cg.band];
                if (bg.res >= R) continue;
                const CbSeg &sg = segs[c][k];
                const uint32_t K = band_k[cg.band];
                if (!(cp.cbstyle & kCbsHt)) {
                    if (prune) {
                        const Rect bn = bg.res == 0 ? need[c][0] : need_of(need[c][bg.res], bg.rect);
                        if (!meets(cg.rect, bn)) {
                            ++lp.pruned_cbs;
                            continue;
                        }
                    }
                    st = classic_job(cs, cp, bg, cg, sg, K, band_off[cg.band], lp);
                    if (st) return st;
                    continue;
                }
                CbJob j{};
                j.stride = coef_pitch(bg.rect.w());
                j.dst = band_off[cg.band] + (cg.rect.y0 - bg.rect.y0) * j.stride +
                        (cg.rect.x0 - bg.rect.x0);
                j.w = static_cast<uint16_t>(cg.rect.w());
                j.h = static_cast<uint16_t>(cg.rect.h());
                j.kmax = static_cast<uint8_t>(K);
                j.flags = static_cast<uint8_t>(((cp.cbstyle & 0x08) ? kCbStripeCausal : 0) |
                                               ((bg.type & 3u) << kCbBandShift));
                if (j.w == 0 || j.h == 0) continue;
                const uint32_t actual = K > sg.zblk ? K - sg.zblk : 0;
                if (K > 31 && sg.passes > 0) {
                    set_last_error("too many bit-planes");
                    return kErrMalformed;
                }
                if (dev) {
                    // a zero block until the device reads its packet header
                    j.flags |= kCbZero;
                    lp.pk_cbs[pk_index[c][k]].kmax = K;
                } else if (K == 0 || sg.passes == 0 || actual == 0 || sg.offset < 0) {
                    j.flags |= kCbZero;
                } else {
                    j.src = static_cast<uint32_t>(sg.offset);
                    j.lcup = static_cast<uint16_t>(sg.lcup);
                    j.lref = static_cast<uint16_t>(sg.lref);
                    j.zblk = static_cast<uint8_t>(std::min<uint32_t>(sg.zblk, 255));
                    j.passes = static_cast<uint8_t>(sg.passes);
                }
                if (prune) {
                    // a block none of the region's samples depends on is not
                    // decoded: the synthesis of the samples the region needs
                    // never reads its part of the band
                    const Rect bn = bg.res == 0 ? need[c][0] : need_of(need[c][bg.res], bg.rect);
                    if (!meets(cg.rect, bn)) {
                        ++lp.pruned_cbs;
                        continue;
                    }
                }
                j.cost = dev ? (uint32_t(j.w) * j.h) / 4
                       : (j.flags & kCbZero) ? 0 : sg.lcup + sg.lref + (uint32_t(j.w) * j.h) / 4;
                if (dev)
                    lp.pk_cbs[pk_index[c][k]].job = lp.pk_cbs[pk_index[c][k]].lane_pos =
                        static_cast<uint32_t>(lp.cbs.size());
                lp.cbs.push_back(j);
                lp.cb_band.push_back(band_off[cg.band]);
            }
            // synthesis levels (idwt.rs:54-193)
            uint32_t prev_out = band_off[g.res[0].band[0]];
            uint32_t prev_w = coef_pitch(g.bands[g.res[0].band[0]].rect.w());
            float prev_delta = band_delta[g.res[0].band[0]];
            bool prev_float = false;
            // with components on different grids the last level goes to a
            // plane of its own, read by the component's store job
            uint32_t plane = 0;
            auto to_plane = [&](LevelJob &lj, const Rect &rr, uint32_t depth) {
                lj.nbx = static_cast<uint16_t>(ceil_div(rr.w(), 64));
                lj.nby = static_cast<uint16_t>(ceil_div(rr.h(), 64));
                lj.bx0 = lj.by0 = 0;
                lj.out = static_cast<uint32_t>(lp.coef_elems);
                plane = lj.out;
                lp.coef_elems += uint64_t(coef_pitch(rr.w())) * rr.h();
                lp.levels.push_back(lj);
                lp.level_depth.push_back(static_cast<uint8_t>(depth));
            };
            if (R == 1) {
                LevelJob lj{};
                lj.band[0] = prev_out;
                lj.bw[0] = prev_w;
                lj.delta[0] = prev_delta;
                const Rect rr = g.res[0].rect;
                lj.x0 = rr.x0;
                lj.y0 = rr.y0;
                lj.x1 = rr.x1;
                lj.y1 = rr.y1;
                lj.reversible = rev;
                lj.lift53 = f53 ? 1 : 0;
                lj.levels = 0;
                lj.narrow = 1;  // slot 0 is the code-block band itself
                if (smp.mixed) {
                    to_plane(lj, rr, 1);
                } else {
                    fj.level[c] = static_cast<uint32_t>(lp.finals_lv.size());
                    lp.finals_lv.push_back(lj);
                }
            }
            for (uint32_t r = 1; r < R; ++r) {
                const ResGeom &rs = g.res[r];
                LevelJob lj{};
                lj.lift53 = f53 ? 1 : 0;
                lj.band[0] = prev_out;
                lj.bw[0] = prev_w;
                lj.delta[0] = prev_delta;
                lj.float_in = prev_float ? 1 : 0;
                // slots filled by code-blocks: the detail bands, and the
                // low-pass band at the first level
                lj.narrow = static_cast<uint8_t>(r == 1 ? 0xF : 0xE);
                for (int b = 0; b < 3; ++b) {
                    const uint32_t bi = rs.band[b];
                    lj.band[b + 1] = band_off[bi];
                    lj.bw[b + 1] = coef_pitch(g.bands[bi].rect.w());
                    lj.delta[b + 1] = band_delta[bi];
                }
                lj.x0 = rs.rect.x0;
                lj.y0 = rs.rect.y0;
                lj.x1 = rs.rect.x1;
                lj.y1 = rs.rect.y1;
                lj.reversible = rev;
                lj.levels = 1;
                lj.nbx = static_cast<uint16_t>(ceil_div(rs.rect.w(), 64));
                lj.nby = static_cast<uint16_t>(ceil_div(rs.rect.h(), 64));
                if (prune) {
                    // only the 64x64 blocks holding samples the region needs
                    const Rect &n = need[c][r];
                    const uint32_t bx0 = (n.x0 - rs.rect.x0) / 64, by0 = (n.y0 - rs.rect.y0) / 64;
                    const uint32_t bx1 = ceil_div(n.x1 - rs.rect.x0, 64);
                    const uint32_t by1 = ceil_div(n.y1 - rs.rect.y0, 64);
                    lj.bx0 = static_cast<uint16_t>(bx0);
                    lj.by0 = static_cast<uint16_t>(by0);
                    lj.nbx = static_cast<uint16_t>(bx1 > bx0 ? bx1 - bx0 : 0);
                    lj.nby = static_cast<uint16_t>(by1 > by0 ? by1 - by0 : 0);
                }
                if (r + 1 < R) {
                    lj.out = static_cast<uint32_t>(lp.coef_elems);
                    lp.coef_elems += uint64_t(coef_pitch(rs.rect.w())) * rs.rect.h();
                    lp.levels.push_back(lj);
                    lp.level_depth.push_back(static_cast<uint8_t>(r));
                    prev_out = lj.out;
                    prev_w = coef_pitch(rs.rect.w());
                    prev_delta = 1.0f;
                    prev_float = !rev;
                } else if (smp.mixed) {
                    to_plane(lj, rs.rect, r);
                } else {
                    fj.level[c] = static_cast<uint32_t>(lp.finals_lv.size());
                    lp.finals_lv.push_back(lj);
                }
            }
            // the IDWT kernels address a component's slots with 32-bit byte
            // offsets from its lowest slot
            if ((lp.coef_elems - comp_start) * 4 > (uint64_t(1) << 31)) {
                set_last_error("tile-component too large (coefficients over 2 GiB)");
                return kErrUnsupported;
            }
            const Rect fr = g.res[R - 1].rect;
            if (smp.mixed) {
                // output pixels whose grid position falls on this tile's
                // samples: [xr cx0, xr cx1) of the reference grid, inside
                // the image and the region
                StoreJob sj{};
                sj.plane = plane;
                sj.pitch = coef_pitch(fr.w());
                sj.cx0 = fr.x0;
                sj.cy0 = fr.y0;
                const uint64_t gx0 = std::max<uint64_t>(uint64_t(g.xr) * fr.x0, h.xo);
                const uint64_t gy0 = std::max<uint64_t>(uint64_t(g.yr) * fr.y0, h.yo);
                const uint64_t gx1 = std::min<uint64_t>(uint64_t(g.xr) * fr.x1, h.xsiz);
                const uint64_t gy1 = std::min<uint64_t>(uint64_t(g.yr) * fr.y1, h.ysiz);
                sj.ox0 = std::max<uint32_t>(uint32_t(gx0 - h.xo), reg.x0);
                sj.oy0 = std::max<uint32_t>(uint32_t(gy0 - h.yo), reg.y0);
                sj.ox1 = gx1 > h.xo ? std::min<uint32_t>(uint32_t(gx1 - h.xo), reg.x1) : 0;
                sj.oy1 = gy1 > h.yo ? std::min<uint32_t>(uint32_t(gy1 - h.yo), reg.y1) : 0;
                sj.xo = h.xo;
                sj.yo = h.yo;
                sj.dst = dst - uint64_t(reg.y0) * pitch - uint64_t(reg.x0) * bpp;
                sj.opitch = pitch;
                sj.nbx = static_cast<uint16_t>(sj.ox1 > sj.ox0 ? ceil_div(sj.ox1 - sj.ox0, 64) : 0);
                sj.nby = static_cast<uint16_t>(sj.oy1 > sj.oy0 ? ceil_div(sj.oy1 - sj.oy0, 64) : 0);
                sj.xr = g.xr;
                sj.yr = g.yr;
                sj.chan = static_cast<uint8_t>(c);
                sj.nchan = static_cast<uint8_t>(og.channels);
                sj.bps = static_cast<uint8_t>(og.bytes_per_sample);
                sj.precision = cp.precision;
                sj.prec0 = t.comp[0].precision;
                sj.flags = static_cast<uint8_t>((rev ? 1 : 0) | (same8 ? 2 : 0) |
                                                (c == 0 && og.channels == 4 ? 4 : 0));
                if (sj.nbx && sj.nby) lp.stores.push_back(sj);
                continue;
            }
            if (c == 0)
                out_rect = fr;
            else if (fr.x0 != out_rect.x0 || fr.y0 != out_rect.y0 || fr.x1 != out_rect.x1 ||
                     fr.y1 != out_rect.y1) {
                set_last_error("components decode to different rectangles");
                return kErrUnsupported;
            }
            fj.precision[c < 4 ? c : 3] = cp.precision;
        }
        if (smp.mixed) continue;  // the store jobs write this tile
        // image coordinates: the (component) grid minus the image origin
        // (full resolution whenever the origin is not zero)
        out_rect.x0 -= xoff;
        out_rect.x1 -= xoff;
        out_rect.y0 -= yoff;
        out_rect.y1 -= yoff;
        if (out_rect.x1 > og.width || out_rect.y1 > og.height) {
            set_last_error("tile outside the output image");
            return kErrMalformed;
        }
        // the tile's pixel (0, 0) in the output, which starts at the
        // region's corner (possibly outside it: the kernel writes only the
        // crop rectangle)
        fj.dst = dst + uint64_t(int64_t(out_rect.y0) - int64_t(reg.y0)) * pitch +
                 uint64_t(int64_t(out_rect.x0) - int64_t(reg.x0)) * bpp;
        fj.pitch = pitch;
        fj.out_x = out_rect.x0;
        fj.out_y = out_rect.y0;
        fj.width = out_rect.w();
        fj.height = out_rect.h();
        fj.cx0 = std::max(out_rect.x0, reg.x0) - out_rect.x0;
        fj.cy0 = std::max(out_rect.y0, reg.y0) - out_rect.y0;
        fj.cx1 = std::min(out_rect.x1, reg.x1) - out_rect.x0;
        fj.cy1 = std::min(out_rect.y1, reg.y1) - out_rect.y0;
        fj.bx0 = static_cast<uint16_t>(fj.cx0 / 64);
        fj.by0 = static_cast<uint16_t>(fj.cy0 / 64);
        fj.nbx = static_cast<uint16_t>(ceil_div(fj.cx1, 64) - fj.bx0);
        fj.nby = static_cast<uint16_t>(ceil_div(fj.cy1, 64) - fj.by0);
        fj.ncomp = static_cast<uint8_t>(nc);
        fj.fmt = static_cast<uint8_t>(fmt);
        fj.mct = mct ? 1 : 0;
        fj.reversible = t.comp[0].transform == 1 && !lone53;
        fj.rct_f32 = mct && lone53 && t.comp[0].transform == 1 ? 1 : 0;
        fj.same8 = same8 ? 1 : 0;
        if (fj.cx1 > fj.cx0 && fj.cy1 > fj.cy0) lp.finals.push_back(fj);
    }
    return kOk;
}

}  // namespace sgn
