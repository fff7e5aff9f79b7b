// Codestream parsing and plan building (host side).
#pragma once
#include <utility>
#include <vector>

#include "j2k_host.hpp"
#include "plan_types.h"

namespace sgn {

// One tile-part: the body and, when packet headers are packed, the header
// pieces (PPT segments of the tile-part in Zppt order, then its PPM packet;
// tile.rs:291-309), all as (offset, length) in the codestream.
struct TilePart {
    size_t off = 0, len = 0;
    std::vector<std::pair<size_t, size_t>> hdr;
};

struct TileBody {
    Rect rect;
    CompParams comp[kMaxComps];
    uint8_t prog = 0;
    bool mct = false;
    uint16_t layers = 1;
    std::vector<TilePart> parts;
};

// One packet's contribution to a classic code-block: `len` bytes at
// codestream offset `offset` of codeword segment `seg`, carrying `passes`
// coding passes.
struct ClassicPiece {
    int64_t offset = -1;
    uint32_t len = 0, seg = 0, passes = 0;
};

struct CbSeg {
    int64_t offset = -1;  // byte offset of cleanup||refinement in the codestream
    uint32_t lcup = 0, lref = 0, zblk = 0, passes = 0;
    std::vector<ClassicPiece> pieces;  // classic code-blocks: contributions in order
};

// signinum-core Colorspace (crates/signinum-core/src/types.rs:9-20), as
// sgn_j2k_info.colorspace reports it.
enum ColorspaceCode : int {
    kCsGrayscale = 0,
    kCsYCbCr = 1,
    kCsRgb = 2,
    kCsCmyk = 3,
    kCsYcck = 4,
    kCsSRgb = 5,
    kCsSGray = 6,
    kCsIccTagged = 7,
    kCsRct = 8,
    kCsIct = 9,
};

// Where the codestream of an input lies: the input itself (raw codestream)
// or the payload of a JP2 file's jp2c box, with what the JP2 header says.
struct Container {
    const uint8_t *cs = nullptr;
    size_t cs_len = 0, cs_offset = 0;
    bool jp2 = false;
    int colorspace = -1;     // from colr (-1: none; inferred from the codestream)
    uint32_t enumcs = 0;     // EnumCS of the first colr box (0: none or ICC)
    bool saw_colr = false;
    bool palette = false;    // pclr present
    bool cdef = false, cdef_identity = false;
};
Status fail_status(Status s, const char *msg);
bool looks_like_jp2(const uint8_t *data, size_t len);
Status locate_codestream(const uint8_t *data, size_t len, Container &c);
// kOk when the device decode produces the reference's samples for this
// container (no palette, channel reordering or colour conversion after the
// codestream decode), else kErrUnsupported.
Status device_decodable(const Container &c);

// Frames of DICOM encapsulated Pixel Data: each a list of (offset, length)
// byte ranges of the input, concatenated in order.
struct DicomFrame {
    std::vector<std::pair<size_t, size_t>> pieces;
    size_t bytes = 0;
};
Status dicom_frames(const uint8_t *data, size_t len, size_t limit, std::vector<DicomFrame> &frames);

Status read_main_header(const uint8_t *data, size_t len, MainHeader &h);
uint32_t skipped_levels(const MainHeader &h, uint32_t denominator);
Status parse_tiles(const uint8_t *data, size_t len, const MainHeader &h,
                   std::vector<TileBody> &tiles);
Status parse_packets(const uint8_t *cs, const TileBody &t, int ncomp, const TileCompGeom *tcs,
                     std::vector<std::vector<CbSeg>> &segs);

// Output geometry of a decode at 1/denominator.
struct OutputGeom {
    uint32_t width = 0, height = 0, ncomp = 0, channels = 0, bytes_per_sample = 1;
    size_t row_bytes = 0;
};
Status output_geometry(const MainHeader &h, int fmt, uint32_t denominator, OutputGeom &g);

// Component sampling: any sub-sampled, mixed (not all alike), and the
// output grid's factors (the common sub-sampling, else 1).
struct Sampling {
    bool any = false, mixed = false;
    uint32_t sx = 1, sy = 1;
};
Sampling sampling_of(const MainHeader &h);

// Plan records of one codestream with offsets local to it; merged into a
// batch plan by rebasing (arena, coefficient and job offsets).
struct LocalPlan {
    std::vector<CbJob> cbs;          // every code-block of the retained resolutions
    std::vector<LevelJob> levels;    // intermediate synthesis levels
    std::vector<uint8_t> level_depth;// r of each intermediate level job (1-based)
    std::vector<LevelJob> finals_lv; // last level per tile and component
    std::vector<FinalJob> finals;    // one per codestream tile
    std::vector<uint32_t> cb_band;   // element offset of each code-block's band
    uint64_t coef_elems = 0;         // int32 elements of the coefficient arena
    uint32_t max_k = 0;              // largest magnitude bit-plane count of any band
    uint32_t width = 0, height = 0;
    size_t row_bytes = 0;
    uint64_t out_bytes = 0;
    uint64_t pruned_cbs = 0;         // code-blocks a region decode does not need
    // classic (EBCOT) code-blocks: jobs with `src` in the codestream, or, for
    // blocks whose contributions are not contiguous there (several layers),
    // in `joined` (flag kClassicJoined); segment table of the blocks with
    // selective bypass or termination on each pass
    std::vector<ClassicJob> classic;
    std::vector<ClassicSeg> cseg;
    std::vector<uint8_t> joined;
    // components on different sampling grids: per tile-component output
    // jobs in place of the fused final stage
    std::vector<StoreJob> stores;
    // packet headers left to the device (build_local_plan with
    // device_packets): tables with codestream offsets and local job indices
    std::vector<PktTile> pk_tiles;
    std::vector<PktPart> pk_parts;
    std::vector<PktPacket> pk_pkts;
    std::vector<PktBand> pk_bands;
    std::vector<PktCb> pk_cbs;
    uint32_t pk_nodes = 0;
};

// Region of interest in full-resolution image coordinates (Rect of
// TileRegionScaledDecodeJob, crates/signinum-j2k/src/batch.rs:34-46).
struct Region {
    bool on = false;
    uint32_t x = 0, y = 0, w = 0, h = 0;
};

// Output rectangle of a region decode at 1/denominator: the region scaled
// to cover it (Rect::scaled_covering, crates/signinum-core/src/types.rs:
// 69-83), checked against the full and the decoded image
// (crates/signinum-j2k/src/decode.rs:36-71,377-389).
Status region_output(const MainHeader &h, const OutputGeom &og, uint32_t denominator,
                     const Region &roi, Rect &out);

// Parses one codestream into plan records; the output of tile t starts at
// dst + y0 * pitch + x0 * bytes_per_pixel.  With a region, only the region
// is decoded and written (from dst): tiles outside it are dropped and the
// code-blocks and synthesis blocks its pixels do not depend on are pruned.
// device_packets: tiles whose packet headers the device can parse (HT, one
// layer, headers in the bodies) get packet tables instead of parsed
// code-block lengths; their jobs start as zero blocks.
Status build_local_plan(const uint8_t *cs, size_t len, int fmt, uint32_t denominator,
                        uint64_t dst, uint64_t pitch, uint64_t capacity, LocalPlan &lp,
                        const Region &roi = Region(), bool device_packets = false);

// The packet tables of one tile for the device parser; cb_index[c][k]: the
// PktCb entry of code-block k of component c.
Status packet_walk(const TileBody &t, int ncomp, const TileCompGeom *tcs, LocalPlan &lp,
                   std::vector<std::vector<uint32_t>> &cb_index);

}  // namespace sgn
