// CPU differential check of the device packet-header parser
// (signinum_amd/csrc/hip/packets.hpp, compiled here for the host) against
// the host parser: for each codestream given, plans built with host-parsed
// packet headers and with packet tables (build_local_plan with
// device_packets) at full and half resolution, a region, and with bytes of
// the packet headers and bodies overwritten; the device kernel runs on the
// packet tables tile by tile, and every code-block job must equal the host
// plan's (offset, lengths, Zblk, passes, flags).  Prints "ok <jobs>" and
// exits 0 when everything agrees.
#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>

#define __device__
#define __host__
#define __global__
#define __forceinline__ inline
#define __launch_bounds__(...)
#define __restrict__
struct Dim3 {
    uint32_t x = 0, y = 0, z = 0;
};
static Dim3 threadIdx, blockIdx;
static void __syncthreads() {}
using std::max;
using std::min;
#include "../signinum_amd/csrc/hip/packets.hpp"
#include "../signinum_amd/csrc/host/parse.hpp"
#include "../signinum_amd/csrc/host/j2k_host.hpp"
using namespace sgn;
int main(int argc, char **argv) {
    size_t total = 0, bad = 0, tiles = 0;
    for (int a = 1; a < argc; ++a) {
        FILE *f = fopen(argv[a], "rb");
        std::vector<uint8_t> cs(1 << 26);
        size_t len = fread(cs.data(), 1, cs.size(), f);
        fclose(f);
        cs.resize(len + 64);
        for (int mode = 0; mode < 6; ++mode) {
            std::vector<uint8_t> keep = cs;
            if (mode >= 3) {
                // bytes of packet headers and bodies overwritten (the lengths stay)
                const size_t at = len * (mode == 3 ? 30 : mode == 4 ? 55 : 80) / 100;
                for (size_t q = 0; q < 12 && at + q < len; ++q) cs[at + q] = (uint8_t)(0x5A * (q + 1) + mode);
            }
            const uint32_t den = mode == 1 ? 2u : 1u;
            Region roi;
            if (mode == 2) { roi.on = true; roi.x = 20; roi.y = 10; roi.w = 50; roi.h = 30; }
            const size_t L = len;
            LocalPlan ref, dev;
            Status s1 = kErrArg, s2 = kErrArg;
            for (int fmt = 0; fmt < 5 && s1; ++fmt) s1 = build_local_plan(cs.data(), L, fmt, den, 0, 1 << 16, 1ull << 40, ref, roi);
            for (int fmt = 0; fmt < 5 && s2; ++fmt) s2 = build_local_plan(cs.data(), L, fmt, den, 0, 1 << 16, 1ull << 40, dev, roi, true);
            if (s1 || s2) { if (s1 != s2) { printf("%s mode %d: plan %d %d\n", argv[a], mode, s1, s2); ++bad; } continue; }
            if (ref.cbs.size() != dev.cbs.size()) { printf("%s: job count %zu vs %zu\n", argv[a], ref.cbs.size(), dev.cbs.size()); ++bad; continue; }
            std::vector<uint32_t> nodes(2 * dev.pk_nodes + 2), lens(dev.pk_cbs.size() + 1);
            std::vector<uint8_t> st(dev.pk_tiles.size() + 1);
            std::vector<CbJob> jobs = dev.cbs;
            for (size_t t = 0; t < dev.pk_tiles.size(); ++t) {
                // one work-group per tile; lane 0 (which walks the headers
                // after the barrier) runs last
                blockIdx.x = (uint32_t)t;
                for (int l = 63; l > 0; --l) {
                    threadIdx.x = (uint32_t)l;
                    sgn_packet_headers(dev.pk_tiles.data(), (uint32_t)dev.pk_tiles.size(), dev.pk_parts.data(), dev.pk_pkts.data(), dev.pk_bands.data(), dev.pk_cbs.data(), nodes.data(), lens.data(), cs.data(), jobs.data(), nullptr, nullptr, st.data());
                }
                threadIdx.x = 0;
                sgn_packet_headers(dev.pk_tiles.data(), (uint32_t)dev.pk_tiles.size(), dev.pk_parts.data(), dev.pk_pkts.data(), dev.pk_bands.data(), dev.pk_cbs.data(), nodes.data(), lens.data(), cs.data(), jobs.data(), nullptr, nullptr, st.data());
                ++tiles;
            }
            for (size_t k = 0; k < jobs.size(); ++k) {
                const CbJob &x = jobs[k], &y = ref.cbs[k];
                const bool zx = x.flags & kCbZero, zy = y.flags & kCbZero;
                bool same = zx == zy && x.flags == y.flags;
                if (same && !zx) same = x.src == y.src && x.lcup == y.lcup && x.lref == y.lref && x.zblk == y.zblk && x.passes == y.passes;
                if (!same) { if (bad < 10) printf("%s mode %d job %zu: src %u/%u lcup %u/%u lref %u/%u zblk %u/%u passes %u/%u flags %x/%x\n", argv[a], mode, k, x.src, y.src, x.lcup, y.lcup, x.lref, y.lref, x.zblk, y.zblk, x.passes, y.passes, x.flags, y.flags); ++bad; }
                ++total;
            }
            cs = keep;
        }
    }
    if (bad) {
        printf("%zu jobs over %zu tile parses, %zu mismatches\n", total, tiles, bad);
        return 1;
    }
    printf("ok %zu\n", total);
    return 0;
}
