// Device-visible plan records shared by the host plan builder and the HIP
// kernels.  Plain structs of fixed size, uploaded once per plan.
#pragma once
#include <stdint.h>

namespace sgn {

// One HT code-block (the device form of HtCodeBlockDecodeJob,
// crates/signinum-j2k-native/src/lib.rs:149-174).
struct CbJob {
    uint32_t src;     // byte offset of cleanup||refinement data in the arena
    uint32_t dst;     // index of the block's first sample in the coefficient arena, in
                      // elements of the plan's code-block sample type (int32, or int16
                      // in a narrow plan)
    uint32_t stride;  // band row width in elements
    uint16_t lcup, lref;
    uint16_t w, h;
    uint8_t zblk, passes, kmax, flags;  // flags: kCbStripeCausal | kCbZero
    uint32_t cost;    // predicted cost (ordering key)
    uint32_t tile;    // index of the owning tile job in the plan
};
static_assert(sizeof(CbJob) == 32, "CbJob layout");

// flags: bit 0 stripe-causal, bit 1 decodes to zero, bits 2-3 sub-band
// (0 LL, 1 HL, 2 LH, 3 HH; a feature of the decode-cost model)
enum : uint8_t { kCbStripeCausal = 1, kCbZero = 2, kCbBandShift = 2 };

// One classic (EBCOT) code-block (the device form of J2kCodeBlockDecodeJob,
// crates/signinum-j2k-native/src/lib.rs:268-296): MQ / raw codeword
// segments decoded into int32 coefficients.
struct ClassicJob {
    uint32_t src;     // byte offset of the block's joined segments in the arena
    uint32_t dst;     // int32 index of the block's first sample in the coefficient arena
    uint32_t stride;  // band row pitch in elements
    uint32_t len;     // bytes of joined segments
    uint16_t w, h;
    uint8_t planes;   // bit-planes: the band's, less the block's missing ones
    uint8_t passes;   // coding passes present
    uint8_t band;     // 0 LL, 1 HL, 2 LH, 3 HH
    uint8_t style;    // code-block style bits (Table A.19) | kClassicJoined
    uint32_t seg;     // first entry of the block's segments in the segment table
    uint32_t nseg;    // 0: one arithmetic segment over all of the data
    uint32_t tile;    // owning tile job
    uint32_t cost;    // ordering key
};
static_assert(sizeof(ClassicJob) == 40, "ClassicJob layout");

// A codeword segment: bytes [off, off + len) of the block's data carry
// coding passes [start, end); raw: bypass (no arithmetic coding).
struct ClassicSeg {
    uint32_t off, len;
    uint8_t start, end, raw, pad;
};
static_assert(sizeof(ClassicSeg) == 12, "ClassicSeg layout");

enum : uint8_t {
    kCbsBypass = 0x01, kCbsReset = 0x02, kCbsTermAll = 0x04, kCbsCausal = 0x08,
    kCbsSegSym = 0x20, kCbsHt = 0x40,
    kClassicJoined = 0x80  // src indexes the plan's joined-data run, not the codestream
};

// Input features of the code-block decode-cost model (5 values, each
// roughly in [0, 1]): coded bytes, coding passes, missing MSBs (Zblk),
// sub-band and sample count.
struct CbFeatures {
    float f[5];
};
constexpr inline CbFeatures cb_features(const CbJob &j) {
    return CbFeatures{{float(uint32_t(j.lcup) + j.lref) * (1.0f / 4096.0f),
                       float(j.passes) * (1.0f / 3.0f), float(j.zblk) * (1.0f / 16.0f),
                       float((j.flags >> kCbBandShift) & 3u) * (1.0f / 3.0f),
                       float(uint32_t(j.w) * j.h) * (1.0f / 4096.0f)}};
}

// Packet headers parsed on the device (HT, one quality layer, headers in
// the tile-part bodies).  The host lays out what the headers refer to:
// per tile its tile-parts and packets in progression order; per packet the
// precinct-bands it carries; per precinct-band its code-block grid (the
// two tag trees) and, per code-block in raster order, the plan job it
// fills.  Offsets into the next table are absolute within the plan.
struct PktTile {
    uint32_t part0, nparts;  // PktPart range
    uint32_t pkt0, npkts;    // PktPacket range, progression order
    uint32_t node0, nnodes;  // tag-tree nodes (value, low) pairs in the node scratch
    uint32_t cb0, ncbs;      // PktCb range (every code-block of the tile's packets)
};
struct PktPart {
    uint32_t off, len;  // tile-part body in the arena
};
struct PktPacket {
    uint32_t band0;  // PktBand range
    uint16_t nbands;
    uint8_t sop, eph;  // SOP markers may precede the packet; EPH ends its header
};
struct PktBand {
    uint32_t cb0;           // PktCb of the band's first code-block
    uint16_t ncbx, ncby;    // code-block grid of the precinct-band
    uint32_t incl, zbp;     // first node of the inclusion / zero-bit-plane tag tree
};
constexpr uint32_t kPktNoJob = 0xFFFFFFFFu;
constexpr uint32_t kPktWave = 0x80000000u;  // job index in the wave list
struct PktCb {
    uint32_t job;       // per-tile lane-list index, kPktWave | wave-list index, or kPktNoJob
    uint32_t lane_pos;  // its index in the batched (ordered) lane list
    uint32_t kmax;      // magnitude bit-planes of its band
};

// Row pitch (elements) of every band and intermediate level in the
// coefficient arena: a multiple of 16, so that rows of int16 samples start
// 32-byte aligned (the HT kernels store aligned 32-byte row chunks) and rows
// of int32 samples 64-byte aligned.
constexpr inline uint32_t coef_pitch(uint32_t w) { return (w + 15u) & ~15u; }

// One 2D synthesis level of one tile-component (IDWT step,
// J2kDirectIdwtStep in crates/signinum-j2k-native/src/direct_plan.rs).
struct LevelJob {
    uint32_t band[4];   // element offsets: LL, HL, LH, HH (int16 units for `narrow` slots)
    uint32_t bw[4];     // band row pitches (coef_pitch of the band width)
    float delta[4];     // dequantisation steps (irreversible), 1 for float inputs
    uint32_t out;       // element offset of the output (intermediate levels)
    uint32_t x0, y0, x1, y1;  // level rectangle (absolute coordinates)
    uint32_t block_base;      // first work-group of this job
    uint16_t nbx, nby;        // 64x64 output blocks
    uint8_t reversible;
    uint8_t float_in;         // bit b set: band b holds float samples, not integers
    uint8_t levels;           // 1: one synthesis level, 0: copy LL only
    uint8_t narrow;           // bit b set: band b holds int16 code-block samples
    uint16_t bx0, by0;        // first 64x64 block of the nbx x nby computed (region decodes)
    // with reversible == 0: lift with the 5/3 steps in f32 (a reversible
    // tile-component with a level one sample wide or high at an odd origin,
    // whose halved sample the integer 5/3 path cannot carry)
    uint8_t lift53;
    uint8_t pad[3];
};

// Final stage of one tile: last synthesis level of every component fused
// with inverse MCT, sign shift, rounding and packing into the output.
struct FinalJob {
    uint32_t level[4];        // LevelJob index per component
    uint64_t dst;             // device address of the tile's first output pixel
    uint64_t pitch;           // bytes per output row
    uint32_t width, height;   // output rectangle
    uint32_t block_base;
    uint16_t nbx, nby;
    uint8_t ncomp, fmt, mct, reversible;
    uint8_t precision[4];
    uint8_t same8;            // all components 8-bit
    uint8_t rct_f32;          // with reversible == 0 and mct: the reversible transform in f32
    uint8_t pad[2];
    uint32_t out_x, out_y;    // position of this J2K tile inside the output image
    uint16_t bx0, by0;        // first 64x64 block of the nbx x nby decoded
    // Pixels written, relative to the tile's rectangle: [cx0, cx1) x [cy0,
    // cy1) (the whole rectangle unless the decode is of a region); dst is
    // the address the tile's pixel (0, 0) would have, which lies outside the
    // output when the region starts inside the tile.
    uint32_t cx0, cy0, cx1, cy1;
};

// Output of one tile-component of an image whose components are sampled
// on different grids: every output pixel (image coordinates X, Y) of
// [ox0, ox1) x [oy0, oy1) takes the component sample at ((X + xo) / xr,
// (Y + yo) / yr) from the component's plane (the last synthesis level's
// output), shifted, converted and written as channel `chan`
// (decode.rs:1858-1898: each sample covers xr x yr pixels of the grid).
struct StoreJob {
    uint32_t plane;               // element offset of the plane (int32 5/3, f32 9/7)
    uint32_t pitch;               // its row pitch in elements
    uint32_t cx0, cy0;            // component coordinates of the plane's (0, 0)
    uint32_t ox0, oy0, ox1, oy1;  // output pixels written
    uint32_t xo, yo;              // image origin on the reference grid
    uint64_t dst;                 // device address of output pixel (0, 0)
    uint64_t opitch;              // output row bytes
    uint32_t block_base;          // first work-group (64x64 output blocks)
    uint16_t nbx, nby;
    uint8_t xr, yr, chan, nchan;  // sub-sampling, channel written, channels per pixel
    uint8_t bps, precision, prec0, flags;  // flags: 1 reversible, 2 all 8-bit, 4 alpha
};

}  // namespace sgn
