// Host-side codestream model shared by the decode-plan builder and the
// test-input encoder of signinum-j2k-hip.
//
// Geometry follows ITU-T T.800 Annex B as the reference restates it:
// resolution and sub-band rectangles (crates/signinum-j2k-native/src/j2c/
// tile.rs:385-489), precinct partition (tile.rs:512-645) and code-block
// partition (build.rs:113-258).  Everything here is plain data in flat
// vectors so it can be copied into device tables without pointer chasing.
#pragma once

#include <cstddef>
#include <cstdint>
#include <string>
#include <utility>
#include <vector>

namespace sgn {

enum Status : int {
    kOk = 0,
    kErrArg = -1,
    kErrUnsupported = -2,
    kErrMalformed = -3,
    kErrBlock = -4,
    kErrBuffer = -5,
    kErrNoMem = -6,
    kErrDevice = -7,
    kErrRegion = -8,
};

void set_last_error(const std::string &msg);
const char *last_error();

inline uint32_t ceil_div(uint32_t a, uint32_t b) {
    return static_cast<uint32_t>((static_cast<uint64_t>(a) + b - 1) / b);
}

struct Rect {
    uint32_t x0 = 0, y0 = 0, x1 = 0, y1 = 0;
    uint32_t w() const { return x1 > x0 ? x1 - x0 : 0; }
    uint32_t h() const { return y1 > y0 ? y1 - y0 : 0; }
};

constexpr int kMaxRes = 33;
constexpr int kMaxComps = 16;

// Per-component coding parameters after COD/COC/QCD/QCC merging.
struct CompParams {
    uint8_t precision = 8;
    uint8_t scod = 0;  // bit0 user precincts, bit1 SOP allowed, bit2 EPH
    uint8_t levels = 0;
    uint8_t xcb = 6, ycb = 6;  // code-block size exponents
    uint8_t cbstyle = 0x40;
    uint8_t transform = 1;  // 1 = reversible 5/3, 0 = irreversible 9/7
    uint8_t ppx[kMaxRes] = {}, ppy[kMaxRes] = {};
    uint8_t qstyle = 0;  // 0 none, 1 scalar derived, 2 scalar expounded
    uint8_t guard = 1;
    uint16_t nsteps = 0;
    uint16_t eps[3 * kMaxRes + 1] = {};
    uint16_t mu[3 * kMaxRes + 1] = {};
    uint8_t xr = 1, yr = 1;  // sub-sampling on the reference grid (XRsiz, YRsiz)
};

struct MainHeader {
    uint32_t xsiz = 0, ysiz = 0, xo = 0, yo = 0, xt = 0, yt = 0, xto = 0, yto = 0;
    uint16_t ncomp = 0;
    uint8_t prog = 0;
    uint16_t layers = 1;
    bool mct = false;
    CompParams comp[kMaxComps];
    size_t first_sot = 0;
    // PPM packets (codestream.rs:761-780, 164-177): (offset, length) in the
    // codestream, the packet headers of tile-part 0, 1, ... in Zppm order
    std::vector<std::pair<size_t, size_t>> ppm;
    uint32_t tiles_x() const { return ceil_div(xsiz - xto, xt); }
    uint32_t tiles_y() const { return ceil_div(ysiz - yto, yt); }
};

enum BandType : uint8_t { kLL = 0, kHL = 1, kLH = 2, kHH = 3 };

struct BandGeom {
    uint8_t type = kLL;
    uint8_t res = 0;
    Rect rect;
};

struct PrecBand {
    uint32_t ncbx = 0, ncby = 0;
    uint32_t cb_base = 0;  // index of the first code-block in TileCompGeom::cbs
};

struct CbGeom {
    Rect rect;  // in sub-band coordinates (absolute, i.e. including band origin)
    uint32_t band = 0;
    uint16_t xi = 0, yi = 0;
};

struct ResGeom {
    Rect rect;
    uint8_t ppx = 15, ppy = 15;
    uint32_t npx = 0, npy = 0;
    uint32_t cbw = 64, cbh = 64;
    uint8_t nbands = 0;
    uint32_t band[3] = {};
    uint32_t prec_base = 0;  // PrecBand index of (precinct p, band b) = prec_base + p * nbands + b
};

struct TileCompGeom {
    Rect rect;         // the tile-component, in component coordinates (B-12)
    Rect trect;        // the tile on the reference grid
    uint8_t xr = 1, yr = 1;
    uint8_t nres = 0;
    ResGeom res[kMaxRes];
    std::vector<BandGeom> bands;
    std::vector<PrecBand> precs;
    std::vector<CbGeom> cbs;
};

// Builds the resolution / sub-band / precinct / code-block partition of one
// tile-component; tile_rect is the tile on the reference grid, divided by
// the component's sub-sampling (B-12).
Status build_tile_comp(const CompParams &cp, const Rect &tile_rect, TileCompGeom &g);

struct PacketId {
    uint16_t layer, res, comp;
    uint32_t prec;
};

// Packet order of one tile for the five progression orders (B.12).
Status progression(uint8_t order, uint16_t layers, const Rect &tile_rect, const TileCompGeom *tcs,
                   int ncomp, std::vector<PacketId> &out);

// Sub-band quantisation: exponent / mantissa (codestream.rs:225-272),
// bit-plane count K = G + eps - 1 and dequantisation step
// (decode.rs:1130-1166).
Status band_quant(const CompParams &cp, uint8_t type, uint32_t res, uint32_t &k_max,
                  float &delta);

// Tag tree (B.10.2) shared by the packet-header reader and writer.
struct TagTree {
    uint32_t levels = 0;
    uint32_t lw[40] = {}, lh[40] = {}, off[40] = {};
    std::vector<uint32_t> value, low;
    std::vector<uint8_t> known;
    void init(uint32_t w, uint32_t h);
    uint32_t leaf(uint32_t x, uint32_t y, uint32_t level) const {
        return off[level] + (y >> level) * lw[level] + (x >> level);
    }
};

// HT look-up tables.  Decode tables use the entry layout of the reference
// decoder (ht_block_decode.rs:1085-1201); the encoder table maps
// (c_q, rho, emb) to a codeword.
struct HtTables {
    uint16_t vlc0[1024], vlc1[1024];
    uint16_t uvlc0[320], uvlc1[256];
    // encoder: [table][c_q << 8 | rho << 4 | emb] = cwd | len << 8 | e_k << 12, 0 if none
    uint32_t enc[2][2048];
};
const HtTables &ht_tables();

}  // namespace sgn
