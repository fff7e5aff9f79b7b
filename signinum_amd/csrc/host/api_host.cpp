// C ABI entry points that need no device: inspection and scaled geometry.
#include <algorithm>
#include <exception>

#include "parse.hpp"
#include "sgn_api.hpp"

extern "C" {

int sgn_j2k_inspect(const uint8_t *bytes, size_t len, sgn_j2k_info *info) {
    try {
        if (!bytes || !info) return SGN_E_ARG;
        sgn::Container c;
        sgn::Status st = sgn::locate_codestream(bytes, len, c);
        if (st) return st;
        sgn::MainHeader h;
        st = sgn::read_main_header(c.cs, c.cs_len, h);
        if (st) return st;
        *info = sgn_j2k_info{};
        const sgn::Sampling smp = sgn::sampling_of(h);
        info->width = sgn::ceil_div(h.xsiz - h.xo, smp.sx);
        info->height = sgn::ceil_div(h.ysiz - h.yo, smp.sy);
        info->num_components = h.ncomp;
        info->bit_depth = h.comp[0].precision;
        uint32_t nl = 255;
        for (int c = 0; c < h.ncomp; ++c) nl = h.comp[c].levels < nl ? h.comp[c].levels : nl;
        info->num_levels = nl;
        info->reversible = h.comp[0].transform == 1;
        info->mct = h.mct;
        info->high_throughput = (h.comp[0].cbstyle & 0x40) != 0;
        info->tiles_x = h.tiles_x();
        info->tiles_y = h.tiles_y();
        info->cb_width = 1u << h.comp[0].xcb;
        info->cb_height = 1u << h.comp[0].ycb;
        info->progression = h.prog;
        // crates/signinum-j2k/src/parse/mod.rs:41-49,68-74: the JP2 colour
        // specification, else inferred from components, MCT and transform
        int cs = c.colorspace;
        if (cs < 0)
            cs = h.ncomp == 1 ? sgn::kCsSGray
                 : h.ncomp == 3 ? (!h.mct ? sgn::kCsRgb
                                   : h.comp[0].transform == 1 ? sgn::kCsRct : sgn::kCsIct)
                                : sgn::kCsIccTagged;
        info->colorspace = uint32_t(cs);
        info->container = c.jp2 ? 1 : 0;
        return SGN_OK;
    } catch (const std::exception &e) {
        sgn::set_last_error(e.what());
        return SGN_E_ARG;
    }
}

int sgn_j2k_scaled_dims(const uint8_t *bytes, size_t len, uint32_t denominator, uint32_t *width,
                        uint32_t *height) {
    if (!bytes) return SGN_E_ARG;
    sgn::Container c;
    sgn::Status st = sgn::locate_codestream(bytes, len, c);
    if (st) return st;
    sgn::MainHeader h;
    st = sgn::read_main_header(c.cs, c.cs_len, h);
    if (st) return st;
    if (denominator != 1 && denominator != 2 && denominator != 4 && denominator != 8) {
        sgn::set_last_error("downscale denominator must be 1, 2, 4 or 8");
        return SGN_E_ARG;
    }
    const uint32_t sk = sgn::skipped_levels(h, denominator);
    const sgn::Sampling smp = sgn::sampling_of(h);
    if (sk && (h.xo || h.yo)) {
        sgn::set_last_error("reduced resolution of an image with a non-zero origin");
        return SGN_E_UNSUPPORTED;
    }
    if (sk && smp.any) {
        sgn::set_last_error("reduced resolution of sub-sampled components");
        return SGN_E_UNSUPPORTED;
    }
    if (width) *width = sgn::ceil_div(h.xsiz - h.xo, smp.sx << sk);
    if (height) *height = sgn::ceil_div(h.ysiz - h.yo, smp.sy << sk);
    return SGN_OK;
}

int sgn_j2k_region_dims(const uint8_t *bytes, size_t len, uint32_t denominator, uint32_t x,
                        uint32_t y, uint32_t w, uint32_t h, uint32_t *out_x, uint32_t *out_y,
                        uint32_t *out_w, uint32_t *out_h) {
    if (!bytes) return SGN_E_ARG;
    sgn::Container c;
    sgn::Status st = sgn::locate_codestream(bytes, len, c);
    if (st) return st;
    sgn::MainHeader hd;
    st = sgn::read_main_header(c.cs, c.cs_len, hd);
    if (st) return st;
    // the format only sets the byte layout: any one the components map to
    sgn::OutputGeom og;
    st = sgn::output_geometry(hd, hd.ncomp == 1 ? SGN_GRAY8 : SGN_RGB8, denominator, og);
    if (st == sgn::kErrUnsupported && hd.ncomp != 1 && hd.ncomp != 3) {
        sgn::MainHeader g = hd;
        g.ncomp = 1;
        st = sgn::output_geometry(g, SGN_GRAY8, denominator, og);
    }
    if (st) return st;
    sgn::Region roi;
    roi.on = true;
    roi.x = x;
    roi.y = y;
    roi.w = w;
    roi.h = h;
    sgn::Rect r;
    st = sgn::region_output(hd, og, denominator, roi, r);
    if (st) return st;
    if (out_x) *out_x = r.x0;
    if (out_y) *out_y = r.y0;
    if (out_w) *out_w = r.w();
    if (out_h) *out_h = r.h();
    return SGN_OK;
}

}  // extern "C"

extern "C" int sgn_j2k_code_blocks(const uint8_t *bytes, size_t len, sgn_ht_code_block_job *jobs,
                                   size_t capacity, size_t *count) {
    try {
        if (!bytes || !count) return SGN_E_ARG;
        sgn::Container c;
        sgn::Status st = sgn::locate_codestream(bytes, len, c);
        if (st) return st;
        // data offsets stay relative to the caller's input
        const uint64_t base = c.cs_offset;
        bytes = c.cs;
        len = c.cs_len;
        sgn::MainHeader h;
        st = sgn::read_main_header(bytes, len, h);
        if (st) return st;
        std::vector<sgn::TileBody> tiles;
        st = sgn::parse_tiles(bytes, len, h, tiles);
        if (st) return st;
        std::vector<sgn_ht_code_block_job> out;
        std::vector<sgn::TileCompGeom> tcs(h.ncomp);
        std::vector<std::vector<sgn::CbSeg>> segs;
        uint64_t packed = 0;
        for (const sgn::TileBody &t : tiles) {
            for (int c = 0; c < h.ncomp; ++c) {
                if (!(t.comp[c].cbstyle & sgn::kCbsHt)) {
                    sgn::set_last_error("classic (EBCOT) code-blocks have no HT block jobs");
                    return SGN_E_UNSUPPORTED;
                }
                st = sgn::build_tile_comp(t.comp[c], t.rect, tcs[c]);
                if (st) return st;
            }
            st = sgn::parse_packets(bytes, t, h.ncomp, tcs.data(), segs);
            if (st) return st;
            for (int c = 0; c < h.ncomp; ++c) {
                const sgn::TileCompGeom &g = tcs[c];
                for (size_t k = 0; k < g.cbs.size(); ++k) {
                    const sgn::CbGeom &cg = g.cbs[k];
                    const sgn::BandGeom &bg = g.bands[cg.band];
                    const sgn::CbSeg &sg = segs[c][k];
                    if (cg.rect.w() == 0 || cg.rect.h() == 0) continue;
                    uint32_t K;
                    float delta;
                    st = sgn::band_quant(t.comp[c], bg.type, bg.res, K, delta);
                    if (st) return st;
                    sgn_ht_code_block_job j{};
                    j.data_offset = sg.offset >= 0 ? base + uint64_t(sg.offset) : 0;
                    j.cleanup_length = sg.offset >= 0 ? sg.lcup : 0;
                    j.refinement_length = sg.offset >= 0 ? sg.lref : 0;
                    j.width = cg.rect.w();
                    j.height = cg.rect.h();
                    j.output_offset = static_cast<uint32_t>(packed);
                    j.output_stride = j.width;
                    j.missing_bit_planes = static_cast<uint8_t>(std::min<uint32_t>(sg.zblk, 255));
                    j.number_of_coding_passes = static_cast<uint8_t>(sg.offset >= 0 ? sg.passes : 0);
                    j.num_bitplanes = static_cast<uint8_t>(K);
                    j.stripe_causal = (t.comp[c].cbstyle & 0x08) ? 1 : 0;
                    j.dequantization_step = delta;
                    packed += uint64_t(j.width) * j.height;
                    out.push_back(j);
                }
            }
        }
        *count = out.size();
        if (!jobs) return SGN_OK;
        if (capacity < out.size()) {
            sgn::set_last_error("code-block job buffer too small");
            return SGN_E_BUFFER;
        }
        std::copy(out.begin(), out.end(), jobs);
        return SGN_OK;
    } catch (const std::exception &e) {
        sgn::set_last_error(e.what());
        return SGN_E_ARG;
    }
}
