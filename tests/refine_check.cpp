// CPU differential check of the lane refinement passes
// (signinum_amd/csrc/hip/ht_refine.hpp, compiled here for the host) against
// the oracle's HT block decode.  Input file (written by test_readers.py):
// the codestream, then per code-block with refinement passes its offset,
// lengths, size, missing MSBs, pass count, K and stripe-causal flag.  For
// each block the oracle's cleanup-only samples are put into a plane (int16
// and int32, aligned and unaligned rows, inside a guard band), the passes
// run on them, and the result must equal the oracle's full decode, with
// the guard band untouched; each block runs again with the stripe-causal
// context forced on (both sides decode the same bits under that rule).
// Prints "ok <blocks>" and exits 0 when everything agrees.
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>

#define __device__
#define __forceinline__ inline
struct uint2 {
    uint32_t x, y;
};
struct int4 {
    int32_t x, y, z, w;
};
static inline uint2 make_uint2(uint32_t a, uint32_t b) { return {a, b}; }
static inline int4 make_int4(int32_t a, int32_t b, int32_t c, int32_t d) { return {a, b, c, d}; }
static inline int __ffs(uint32_t x) { return __builtin_ffs((int)x); }
using std::max;
using std::min;
#include <algorithm>
#include "../signinum_amd/csrc/hip/ht_refine.hpp"

extern "C" {
#include "../oracle/ht_block.c"
#include "../oracle/ht_tables.c"
}

struct Block {
    uint64_t off;
    uint32_t lcup, lref, w, h, zblk, passes, kmax, causal;
};

template <typename C>
static int check(const std::vector<uint8_t> &arena, const Block &b, uint32_t off, bool vec,
                 bool causal) {
    const uint32_t w = b.w, h = b.h, stride = w;
    std::vector<uint32_t> cl(w * h, 0), full(w * h, 0);
    const uint8_t *d = arena.data() + b.off;
    if (oracle_ht_decode_block(d, b.lcup, b.lref, w, h, b.zblk, b.passes, causal, 0, cl.data(),
                               stride) ||
        oracle_ht_decode_block(d, b.lcup, b.lref, w, h, b.zblk, b.passes, causal, 2, full.data(),
                               stride)) {
        std::printf("oracle failed on a block\n");
        return 1;
    }
    const uint32_t pitch = ((w + 7) & ~7u) + 16;
    const C guard = (C)0x5A5A;
    std::vector<C> plane(pitch * (h + 1), guard);
    for (uint32_t y = 0; y < h; ++y)
        for (uint32_t x = 0; x < w; ++x)
            plane[off + y * pitch + x] = (C)oracle_coefficient_to_i32(cl[y * stride + x], b.kmax);
    uint32_t passes = b.passes;
    if (b.lref == 0 || b.zblk >= 29) passes = 1;
    if (passes > 1)
        sgn::rows_refine<C>(plane.data() + off, w, h, pitch, vec, arena.data(),
                            (uint32_t)b.off + b.lcup, b.lref, passes, causal, 30u - b.zblk,
                            b.kmax >= 31 ? 0u : 31u - b.kmax);
    for (uint32_t y = 0; y <= h; ++y)
        for (uint32_t x = 0; x < pitch; ++x) {
            const bool in = y < h && x >= off && x < off + w;
            const int32_t want =
                in ? oracle_coefficient_to_i32(full[y * stride + x - off], b.kmax) : (int32_t)guard;
            if ((int32_t)plane[y * pitch + x] != want) {
                std::printf("mismatch %ux%u block at (%u,%u): %d, want %d (passes %u, causal %d)\n",
                            w, h, x, y, (int)plane[y * pitch + x], want, b.passes, (int)causal);
                return 1;
            }
        }
    return 0;
}

int main(int argc, char **argv) {
    if (argc != 2) return 2;
    FILE *f = std::fopen(argv[1], "rb");
    if (!f) return 2;
    uint64_t cs_len = 0, nblocks = 0;
    if (std::fread(&cs_len, 8, 1, f) != 1 || std::fread(&nblocks, 8, 1, f) != 1) return 2;
    std::vector<uint8_t> arena(cs_len + 64, 0);
    if (std::fread(arena.data(), 1, cs_len, f) != cs_len) return 2;
    std::vector<Block> blocks(nblocks);
    for (Block &b : blocks)
        if (std::fread(&b, sizeof(Block), 1, f) != 1) return 2;
    std::fclose(f);
    for (const Block &b : blocks) {
        for (int forced = 0; forced < 2; ++forced) {
            const bool causal = forced ? true : b.causal != 0;
            // int16 planes hold blocks of at most 15 bit-planes
            const bool narrow = b.kmax <= 15;
            if ((narrow && check<int16_t>(arena, b, 8, true, causal)) ||
                check<int32_t>(arena, b, 8, true, causal) ||
                (narrow && check<int16_t>(arena, b, 3, false, causal)) ||
                check<int32_t>(arena, b, 5, false, causal))
                return 1;
        }
    }
    std::printf("ok %zu\n", blocks.size());
    return 0;
}
