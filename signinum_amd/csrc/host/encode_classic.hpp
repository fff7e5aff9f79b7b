// Classic (EBCOT) code-block encoder (T.800 Annex C and D).
#pragma once
#include <cstddef>
#include <cstdint>
#include <vector>

namespace sgn {

// One code-block, every coding pass: its missing MSBs (the zero bit-plane
// tag-tree value) and the terminated codeword segments in order, each with
// the passes it carries.  An all-zero block has no passes.
struct ClassicBlockCode {
    struct Segment {
        std::vector<uint8_t> bytes;
        uint32_t passes = 0;
    };
    uint32_t missing_msbs = 0, passes = 0;
    std::vector<Segment> segments;
};

// coef: w x h signed integer coefficients (row stride `stride`) of a band
// with K magnitude bit-planes; band 0 LL, 1 HL, 2 LH, 3 HH; style the
// code-block style bits (0x01 bypass, 0x02 reset, 0x04 termination on each
// pass, 0x08 vertically causal, 0x20 segmentation symbols).  False when a
// magnitude needs more than K bit-planes.
bool classic_encode_block(const int32_t *coef, uint32_t w, uint32_t h, size_t stride, uint32_t K,
                          uint32_t band, uint32_t style, ClassicBlockCode &out);

}  // namespace sgn
