// CPU differential check of the lane kernel's 4-byte-group HT bit readers
// (signinum_amd/csrc/hip/ht_readers4.hpp, compiled here for the host)
// against the oracle's byte-wise MEL, MagSgn and VLC readers, on random
// streams rich in stuffing patterns, at every alignment, read far past the
// end of each segment.  Prints "ok" and exits 0 when every value agrees.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <vector>

#define __device__
#define __forceinline__ inline
#include "../signinum_amd/csrc/hip/ht_readers4.hpp"

extern "C" {
#include "../oracle/ht_block.c"
#include "../oracle/ht_tables.c"
}

struct HostSrc {
    const uint8_t *arena;
    uint64_t word(uint32_t a) const {
        uint64_t w;
        std::memcpy(&w, arena + a, 8);
        return w;
    }
};

static int fail(const char *what, int trial, int step, uint32_t a, uint32_t b) {
    std::printf("%s mismatch: trial %d step %d: %08x vs %08x\n", what, trial, step, a, b);
    return 1;
}

int main() {
    std::mt19937_64 rng(20261202);
    const uint32_t kArena = 4096;
    std::vector<uint8_t> arena(kArena + 64, 0);
    auto fill = [&]() {
        for (uint32_t i = 0; i < kArena; ++i) {
            const uint32_t r = rng() % 100;
            arena[i] = r < 15 ? 0xFF : r < 30 ? 0x7F : r < 40 ? 0xFE : uint8_t(rng());
        }
    };
    const HostSrc src{arena.data()};
    for (int trial = 0; trial < 4000; ++trial) {
        fill();
        const uint32_t off = 64 + rng() % 1024;
        const uint32_t len = rng() % 200;
        // ---- MagSgn (forward, pad 0xFF) ----
        {
            fwd_t ref;
            fwd_init(&ref, arena.data() + off, len, 0xFF);
            sgn::Fwd4 f;
            sgn::fwd4_init(f, off, len, 0xFF, src);
            sgn::MsBits b{0, 0, 0};
            for (int step = 0; step < 120; ++step) {
                const uint32_t m = rng() % 32;
                const uint32_t a = fwd_fetch(&ref);
                sgn::ms_fill(b, f, src);
                const uint32_t c = (uint32_t)b.tmp;
                if (a != c) return fail("magsgn", trial, step, a, c);
                fwd_advance(&ref, m);
                b.tmp >>= m;
                b.bits -= m;
            }
        }
        // ---- VLC (backward, from Lcup - 2 with the Scup nibble) ----
        {
            const uint32_t lcup = 2 + rng() % 300;
            const uint32_t scup = 2 + rng() % (lcup - 1);
            rev_t ref;
            rev_init_vlc(&ref, arena.data() + off, lcup, scup);
            const uint32_t b0 = arena[off + lcup - 2];
            sgn::VlcBits b;
            b.tmp = b0 >> 4;
            b.bits = 4 - (((b.tmp & 7) == 7) ? 1u : 0u);
            b.unstuff = ((b0 | 0x0F) > 0x8F) ? 1u : 0u;
            sgn::Rev4 r;
            sgn::rev4_init(r, off + lcup - 3, scup - 2, src);
            for (int step = 0; step < 120; ++step) {
                const uint32_t m = rng() % 28;
                const uint32_t a = rev_fetch(&ref);
                sgn::vlc_fill(b, r, src);
                const uint32_t c = (uint32_t)b.tmp;
                if (a != c) return fail("vlc", trial, step, a, c);
                rev_advance(&ref, m);
                b.tmp >>= m;
                b.bits -= m;
            }
        }
        // ---- MEL (forward runs) ----
        {
            const uint32_t mlen = 1 + rng() % 40;
            mel_t ref;
            std::memset(&ref, 0, sizeof ref);
            ref.data = arena.data();
            ref.pos = off;
            ref.remaining = mlen;
            sgn::Fwd4 f;
            sgn::fwd4_init(f, off, mlen, 0xFF, src);
            sgn::Mel m = sgn::mel_make((uint32_t)mlen);
            for (int step = 0; step < 150; ++step) {
                const int a = mel_get_run(&ref);
                const int c = sgn::mel_run4(m, f, src);
                if (a != c) return fail("mel", trial, step, (uint32_t)a, (uint32_t)c);
            }
        }
    }
    std::printf("ok\n");
    return 0;
}
