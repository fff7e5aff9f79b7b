// Mutation check of the host planner (build under AddressSanitizer): each
// codestream or JP2 file given is damaged many times over (bytes flipped, a run set to
// 0xFF, marker bytes planted, truncation, bytes dropped), and every damaged
// stream goes through build_local_plan at full and reduced resolution, with
// and without a region and with the device packet tables.  The planner must
// either refuse the stream or return a plan whose every job stays inside
// the buffers the kernels will index: code-block data inside the
// codestream (or the joined classic data), code-block samples inside the
// coefficient arena, segment tables and packet tables inside their own
// ranges, tile outputs inside the output capacity.  The device kernels
// trust these offsets, so a plan that breaks one of them would turn a
// malformed file into an out-of-bounds access on the GPU.  Prints the
// status of each undamaged stream ("plan <file> <status>"), then
// "ok <plans> <refused>", and exits 0 when no plan breaks an invariant.
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../signinum_amd/csrc/host/parse.hpp"
#include "../signinum_amd/csrc/host/j2k_host.hpp"
using namespace sgn;

static uint64_t rng_state = 0x9E3779B97F4A7C15ull;
static uint32_t rnd() {
    rng_state = rng_state * 6364136223846793005ull + 1442695040888963407ull;
    return uint32_t(rng_state >> 33);
}

static std::vector<uint8_t> mutate(const std::vector<uint8_t> &src) {
    std::vector<uint8_t> cs = src;
    const size_t n = cs.size();
    switch (rnd() % 6) {
    case 0:  // a few bytes flipped anywhere
        for (int k = 0, m = 1 + rnd() % 4; k < m; ++k) cs[rnd() % n] ^= uint8_t(1u << (rnd() % 8));
        break;
    case 1: {  // a run of 0xFF
        const size_t at = rnd() % n, len = 1 + rnd() % 16;
        for (size_t q = at; q < at + len && q < n; ++q) cs[q] = 0xFF;
        break;
    }
    case 2: {  // a marker planted (SOT, SOP, EPH, EOC, COD, QCD, PPT)
        static const uint8_t mk[] = {0x90, 0x91, 0x92, 0xD9, 0x52, 0x5C, 0x61};
        const size_t at = rnd() % (n - 1);
        cs[at] = 0xFF;
        cs[at + 1] = mk[rnd() % sizeof(mk)];
        break;
    }
    case 3:  // truncated
        cs.resize(1 + rnd() % n);
        break;
    case 4: {  // bytes dropped from the middle
        const size_t at = rnd() % n, len = 1 + rnd() % 32;
        cs.erase(cs.begin() + at, cs.begin() + (at + len < n ? at + len : n));
        break;
    }
    default:  // a header byte (the first 200) set to a random value
        cs[rnd() % (n < 200 ? n : 200)] = uint8_t(rnd());
        break;
    }
    return cs;
}

static int check(const LocalPlan &lp, size_t len, uint64_t capacity, const char *what) {
    int bad = 0;
    auto fail = [&](const char *m, size_t i) {
        if (bad++ < 5) printf("%s: %s (entry %zu)\n", what, m, i);
    };
    // in a narrow plan dst counts int16 elements; the arena holds
    // coef_elems int32 elements, so 2 * coef_elems bounds either unit
    const uint64_t elems = 2 * lp.coef_elems;
    for (size_t i = 0; i < lp.cbs.size(); ++i) {
        const CbJob &j = lp.cbs[i];
        if (!(j.flags & kCbZero) && uint64_t(j.src) + j.lcup + j.lref > len)
            fail("code-block bytes past the codestream", i);
        if (j.w == 0 || j.h == 0 || uint32_t(j.w) * j.h > 4096 || j.stride < j.w) {
            char m[96];
            snprintf(m, sizeof m, "code-block shape %ux%u stride %u", j.w, j.h, j.stride);
            fail(m, i);
        }
        else if (uint64_t(j.dst) + uint64_t(j.h - 1) * j.stride + j.w > elems)
            fail("code-block samples past the coefficient arena", i);
    }
    for (size_t i = 0; i < lp.classic.size(); ++i) {
        const ClassicJob &j = lp.classic[i];
        const uint64_t data = (j.style & kClassicJoined) ? lp.joined.size() : len;
        if (uint64_t(j.src) + j.len > data) fail("classic bytes past their data", i);
        if (j.w == 0 || j.h == 0 || uint32_t(j.w) * j.h > 4096 || j.stride < j.w) {
            char m[96];
            snprintf(m, sizeof m, "classic shape %ux%u stride %u", j.w, j.h, j.stride);
            fail(m, i);
        }
        else if (uint64_t(j.dst) + uint64_t(j.h - 1) * j.stride + j.w > lp.coef_elems)
            fail("classic samples past the coefficient arena", i);
        if (uint64_t(j.seg) + j.nseg > lp.cseg.size()) fail("classic segment range", i);
        for (uint32_t s = 0; s < j.nseg && uint64_t(j.seg) + s < lp.cseg.size(); ++s) {
            const ClassicSeg &g = lp.cseg[j.seg + s];
            if (uint64_t(g.off) + g.len > j.len) fail("classic segment past the block data", i);
        }
    }
    for (size_t i = 0; i < lp.pk_parts.size(); ++i)
        if (uint64_t(lp.pk_parts[i].off) + lp.pk_parts[i].len > len) fail("tile-part body", i);
    for (size_t i = 0; i < lp.pk_tiles.size(); ++i) {
        const PktTile &t = lp.pk_tiles[i];
        if (uint64_t(t.part0) + t.nparts > lp.pk_parts.size() ||
            uint64_t(t.pkt0) + t.npkts > lp.pk_pkts.size() ||
            uint64_t(t.node0) + t.nnodes > lp.pk_nodes || uint64_t(t.cb0) + t.ncbs > lp.pk_cbs.size())
            fail("packet table ranges", i);
    }
    for (size_t i = 0; i < lp.pk_pkts.size(); ++i)
        if (uint64_t(lp.pk_pkts[i].band0) + lp.pk_pkts[i].nbands > lp.pk_bands.size())
            fail("packet band range", i);
    for (size_t i = 0; i < lp.pk_bands.size(); ++i) {
        const PktBand &b = lp.pk_bands[i];
        const uint64_t ncb = uint64_t(b.ncbx) * b.ncby;
        if (b.cb0 + ncb > lp.pk_cbs.size()) fail("packet band code-blocks", i);
        // a tag tree over ncbx x ncby leaves has fewer than 2 * ncb + levels nodes
        if (uint64_t(b.incl) + ncb > lp.pk_nodes || uint64_t(b.zbp) + ncb > lp.pk_nodes)
            fail("tag-tree nodes", i);
    }
    for (size_t i = 0; i < lp.pk_cbs.size(); ++i) {
        const PktCb &c = lp.pk_cbs[i];
        if (c.job != kPktNoJob && !(c.job & kPktWave) && c.lane_pos >= lp.cbs.size() &&
            c.job >= lp.cbs.size())
            fail("packet code-block job", i);
    }
    if (lp.out_bytes > capacity) fail("output past the capacity", 0);
    return bad;
}

// What the decode entry points do: find the codestream (raw, or a JP2
// file's jp2c box), then plan it.  `cs_len`: the codestream's length, the
// bound of the plan's data offsets.
static Status plan(const std::vector<uint8_t> &in, int fmt, uint32_t den, uint64_t capacity,
                   LocalPlan &lp, const Region &roi, bool dev, size_t &cs_len) {
    Container c;
    const Status st = locate_codestream(in.data(), in.size(), c);
    if (st != kOk) return st;
    if (c.cs < in.data() || c.cs + c.cs_len > in.data() + in.size()) {
        printf("codestream outside the input\n");
        return kErrArg;
    }
    cs_len = c.cs_len;
    return build_local_plan(c.cs, c.cs_len, fmt, den, 0, 1 << 16, capacity, lp, roi, dev);
}

int main(int argc, char **argv) {
    // SGN_FUZZ_ROUNDS / SGN_FUZZ_SEED: a longer or different run
    const char *er = getenv("SGN_FUZZ_ROUNDS"), *es = getenv("SGN_FUZZ_SEED");
    const int rounds = er ? atoi(er) : 400;
    if (es) rng_state ^= strtoull(es, nullptr, 10) * 0xD1B54A32D192ED03ull;
    size_t plans = 0, refused = 0, bad = 0;
    const uint64_t capacity = 1ull << 40;
    for (int a = 1; a < argc; ++a) {
        FILE *f = fopen(argv[a], "rb");
        if (!f) return 2;
        std::vector<uint8_t> src;
        uint8_t buf[65536];
        size_t got;
        while ((got = fread(buf, 1, sizeof buf, f)) > 0) src.insert(src.end(), buf, buf + got);
        fclose(f);
        if (src.size() < 8) return 2;
        {
            // the stream as given, first, in the first output format it takes
            std::vector<uint8_t> cs(src);
            LocalPlan lp;
            Status st = kErrArg;
            size_t cs_len = 0;
            for (int fmt = 0; fmt < 5 && st != kOk; ++fmt) {
                lp = LocalPlan();
                st = plan(cs, fmt, 1, capacity, lp, Region(), false, cs_len);
            }
            printf("plan %s %d\n", argv[a], int(st));
            if (st == kOk) bad += check(lp, cs_len, capacity, argv[a]);
        }
        for (int r = 0; r < rounds; ++r) {
            const std::vector<uint8_t> m = mutate(src);
            // an exact-size heap copy: a read past the end is a sanitizer error
            std::vector<uint8_t> cs(m);
            const int fmt = int(rnd() % 5);
            const uint32_t den = 1u << (rnd() % 3);
            Region roi;
            if (rnd() % 3 == 0) {
                roi.on = true;
                roi.x = rnd() % 64;
                roi.y = rnd() % 64;
                roi.w = 1 + rnd() % 96;
                roi.h = 1 + rnd() % 96;
            }
            const bool dev = rnd() % 2;
            LocalPlan lp;
            size_t cs_len = 0;
            const Status st = plan(cs, fmt, den, capacity, lp, roi, dev, cs_len);
            if (st != kOk) {
                ++refused;
                continue;
            }
            ++plans;
            char what[512];
            snprintf(what, sizeof what, "%s round %d fmt %d den %u roi %d dev %d", argv[a], r, fmt,
                     den, int(roi.on), int(dev));
            bad += check(lp, cs_len, capacity, what);
        }
    }
    if (bad) {
        printf("bad %zu\n", bad);
        return 1;
    }
    printf("ok %zu %zu\n", plans, refused);
    return 0;
}
