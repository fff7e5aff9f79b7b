// HT cleanup decode, one code-block per lane, with the compressed streams
// staged through LDS.
//
// Each wave decodes 64 code-blocks (at most 64x64; the host sorts blocks by
// predicted cost so the lanes of a wave carry similar work).  Of a block
// with SigProp / MagRef passes this kernel decodes the cleanup; sgn_ht_refine
// runs the passes after it over the samples the cleanup wrote.
// The decode walks every block in segments of kSegPairs quad pairs of one
// quad row.  Between segments the wave stops at a uniform point and refills,
// for all 64 lanes at once, three per-lane LDS rings (MagSgn, VLC, MEL) with
// LDS-DMA loads of whole 16-byte units: every HBM line of a stream is read
// once, coalesced, and the serial bit-unstuffing chains inside a segment read
// LDS only.  The only global memory operations inside a segment are the
// coefficient stores (four samples of a row per 16-byte store) and, when a
// stream outruns its ring, a synchronous fallback load.
//
// MEL, VLC, UVLC and the magnitude/sign phase run fused per quad pair as in
// crates/signinum-j2k-native/src/j2c/ht_block_decode.rs (decode_cleanup_
// symbols :1056-1204, decode_magnitude_sign_phase :1435-1560); the bit
// readers take four bytes at a time when no stuffing can occur in them.
#include <hip/hip_runtime.h>

#include "../host/plan_types.h"
#include "ht_common.hpp"
#include "ht_readers4.hpp"
#include "ht_refine.hpp"

namespace sgn {

constexpr int kRowsMaxQ = 34;    // quads per row of a 64-wide block, plus 2
// Bytes of a block row that a lane gathers in registers and stores at once
// (a multiple of the 32-byte HBM write granule).  The coefficient pitch is
// a multiple of this many int16 samples.
constexpr uint32_t kRowChunkBytes = 32;

// Synchronous 8-byte load for a stream that ran past its staged range.  The
// wait sits inside the rarely taken branch, so the common path never waits
// on vector memory (which would include the in-flight coefficient stores).
__device__ __forceinline__ uint64_t load_now(const uint64_t *p) {
    uint64_t v;
    asm volatile("global_load_dwordx2 %0, %1, off\n\ts_waitcnt vmcnt(0)"
                 : "=v"(v)
                 : "v"(p)
                 : "memory");
    return v;
}

// Units [lo, hi) (16-byte units of the arena) held by a lane's ring of
// SLOTS units; unit u lives in slot u % SLOTS.
struct Staged {
    uint32_t lo, hi;
};

// Word at arena byte address a (8-aligned) from the lane's ring, or from
// memory when it is not staged.
template <int SLOTS>
__device__ __forceinline__ uint64_t ring_word(const uint8_t *ring, Staged s, uint32_t a,
                                              const uint8_t *arena) {
    const uint32_t u = a >> 4;
    if (u >= s.lo && u < s.hi)
        return *reinterpret_cast<const uint64_t *>(ring + (a & (SLOTS * 16 - 1)));
    return load_now(reinterpret_cast<const uint64_t *>(arena) + (a >> 3));
}

// Loads, for every lane of the wave, the units of [want.lo, want.hi) that
// its ring does not hold.  Must be called with all 64 lanes active.  The
// loads complete asynchronously into LDS; the caller waits on vmcnt.
// Word source of the bit readers: the lane's ring, memory past it.
template <int SLOTS>
struct RingSrc {
    const uint8_t *ring;
    Staged s;
    const uint8_t *arena;
    __device__ __forceinline__ uint64_t word(uint32_t a) const {
        return ring_word<SLOTS>(ring, s, a, arena);
    }
};

template <int SLOTS>
__device__ __forceinline__ void stage_ring(uint4 *wave_ring, const uint8_t *arena, Staged want,
                                           Staged have) {
    constexpr int kOwners = 64 / SLOTS;  // lanes whose rings one load instruction fills
    const int lane = threadIdx.x & 63;
    const int sub = lane % SLOTS, grp = lane / SLOTS;
    const uint4 *src = reinterpret_cast<const uint4 *>(arena);
#pragma unroll
    for (int base = 0; base < 64; base += kOwners) {
        const int owner = base + grp;
        const uint32_t wl = __shfl(want.lo, owner), wh = __shfl(want.hi, owner);
        const uint32_t hl = __shfl(have.lo, owner), hh = __shfl(have.hi, owner);
        // the unit of [wl, wl + SLOTS) that maps to this lane's slot
        const uint32_t u = wl + ((uint32_t)(sub - (int)(wl % SLOTS)) & (SLOTS - 1));
        if (u < wh && !(u >= hl && u < hh))
            __builtin_amdgcn_global_load_lds(src + u, wave_ring + base * SLOTS, 16, 0, 0);
    }
}

// MagSgn top-up for the whole wave: while any active lane holds fewer than
// `need` bits (need <= 33), every active lane with room appends a group.  Lanes refill
// together instead of one at a time at different quads, so the refill runs
// less often and with most lanes active; the bits each lane reads are the
// same (only when its groups are appended changes).
template <class Src>
__device__ __forceinline__ void ms_topup(MsBits &b, Fwd4 &f, const Src &src, uint32_t need) {
    if (!__any(b.bits < need)) return;
    do {
        if (b.bits <= 32) ms_append(b, f, src);
    } while (__any(b.bits < need));
}

// Next MEL run with the refill shared by the active lanes, as ms_topup.
template <class Src>
__device__ __forceinline__ int mel_run_wave(Mel &m, Fwd4 &f, const Src &src) {
    if (__any(m.bits < 6)) {
        do {
            if (m.bits <= 32) mel_fill(m, f, src);
        } while (__any(m.bits < 6));
    }
    return mel_symbol(m);
}

__device__ __forceinline__ uint32_t wave_max(uint32_t v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = max(v, (uint32_t)__shfl_xor(v, o));
    return v;
}


// Decodes the 64 code-blocks jobs[first .. first + 64) (lane l takes
// first + l) with the calling wave.  WAVES waves share the work-group's
// tables; SEG quad pairs are decoded between refills of the rings of MS
// (MagSgn), VS (VLC) and LS (MEL) 16-byte slots per lane.  Must be entered by
// all 64 lanes of the wave.
template <int WAVES, int SEG, int MS, int VS, int LS, typename C>
__device__ __forceinline__ void rows_wave(const CbJob *__restrict__ jobs, uint32_t n,
                                          uint32_t first, const uint8_t *__restrict__ arena,
                                          C *__restrict__ coef,
                                          uint8_t *__restrict__ status, const uint16_t *s_tab,
                                          uint8_t (*s_q)[WAVES * 64], uint4 *s_ms, uint4 *s_vlc,
                                          uint4 *s_mel) {
    constexpr int kSegPairs = SEG, kMsSlots = MS, kVlcSlots = VS, kMelSlots = LS;
    const uint16_t *vlc0 = s_tab, *vlc1 = s_tab + 1024, *uvlc0 = s_tab + 2048,
                   *uvlc1 = s_tab + 2368;
    const int tid = threadIdx.x, wv = tid >> 6;
    const uint32_t i = first + (uint32_t)(tid & 63);
    const bool valid = i < n;
    CbJob j{};
    if (valid) j = jobs[i];
    C *out = coef + j.dst;
    const uint32_t w = j.w, h = j.h, stride = j.stride;
    const bool vec_store = ((j.dst | stride) & 3u) == 0;
    // Rows go out in aligned 32-byte chunks (the HBM write granule): a
    // lane keeps the chunk it is filling in registers, so that no partly
    // written line of its block rows is ever evicted from L2 and written
    // back more than once.
    constexpr uint32_t kCW = kRowChunkBytes / 4;              // words per row chunk
    constexpr uint32_t kChunkCols = kRowChunkBytes / sizeof(C);  // samples per row chunk
    constexpr uint32_t kChunkPairs = kChunkCols / 4;            // quad pairs per row chunk
    const bool chunk_store = ((j.dst | stride) & (kChunkCols - 1)) == 0;
    uint32_t ch0[kCW], ch1[kCW];  // rows y, y+1 of the chunk being filled, packed samples
#pragma unroll
    for (uint32_t t = 0; t < kCW; ++t) ch0[t] = ch1[t] = 0;
    uint8_t st = kCbOk;
    bool run_block = valid;
    uint32_t scup = 0;
    const uint32_t lcup = j.lcup, zblk = j.zblk, kmax = j.kmax;
    if (valid && (j.flags & kCbZero)) {
        for (uint32_t y = 0; y < h; ++y)
            for (uint32_t x = 0; x < w; ++x) out[y * stride + x] = 0;
        run_block = false;
    } else if (valid) {
        if (lcup >= 2) scup = ((uint32_t)arena[j.src + lcup - 1] << 4) + (arena[j.src + lcup - 2] & 0x0F);
        if (lcup < 2 || zblk >= 30 || scup < 2 || scup > lcup || scup > 4079) {
            for (uint32_t y = 0; y < h; ++y)
                for (uint32_t x = 0; x < w; ++x) out[y * stride + x] = 0;
            st = kCbBad;
            run_block = false;
        }
    }
    // stream geometry (arena byte addresses)
    const uint32_t ms_off = j.src, ms_len = run_block ? lcup - scup : 0;
    const uint32_t mel_off = j.src + lcup - scup, mel_len = run_block ? scup - 1 : 0;
    const uint32_t vlc_start = j.src + lcup - 3, vlc_len = run_block ? scup - 2 : 0;
    const uint8_t *ring_ms = reinterpret_cast<const uint8_t *>(s_ms + tid * kMsSlots);
    static_assert(kMsSlots >= kSegPairs, "MagSgn ring smaller than a segment's typical need");
    const uint8_t *ring_vlc = reinterpret_cast<const uint8_t *>(s_vlc + tid * kVlcSlots);
    const uint8_t *ring_mel = reinterpret_cast<const uint8_t *>(s_mel + tid * kMelSlots);
    uint4 *wave_ms = s_ms + wv * 64 * kMsSlots;
    uint4 *wave_vlc = s_vlc + wv * 64 * kVlcSlots;
    uint4 *wave_mel = s_mel + wv * 64 * kMelSlots;
    // words of each segment, counted from its first fetch address
    auto fwd_words = [](uint32_t off, uint32_t len) {
        return len ? (int32_t)((((off + len - 1) & ~7u) - (off & ~7u)) >> 3) + 1 : 0;
    };
    const int32_t rev_words0 =
        vlc_len ? (int32_t)(((vlc_start & ~7u) - ((vlc_start + 1 - vlc_len) & ~7u)) >> 3) + 1 : 0;

    // Units wanted in a ring: from the unit holding the next word to fetch,
    // as many as the ring holds, limited to the segment.
    auto want_fwd = [](uint32_t next, int32_t words, uint32_t slots) {
        Staged r;
        r.lo = next >> 4;
        r.hi = r.lo;
        if (words > 0) r.hi = min(r.lo + slots, ((next + 8u * (uint32_t)(words - 1)) >> 4) + 1);
        return r;
    };
    auto want_rev = [](uint32_t next, int32_t words, uint32_t slots) {
        Staged r;
        r.lo = r.hi = 0;
        if (words > 0) {
            r.hi = (next >> 4) + 1;
            const uint32_t first_u = (next - 8u * (uint32_t)(words - 1)) >> 4;
            r.lo = r.hi - min(r.hi - first_u, slots);
        }
        return r;
    };

    Staged have_ms{0, 0}, have_vlc{0, 0}, have_mel{0, 0};
    {
        const Staged wm = want_fwd(ms_off & ~7u, fwd_words(ms_off, ms_len), kMsSlots);
        const Staged wl = want_fwd(mel_off & ~7u, fwd_words(mel_off, mel_len), kMelSlots);
        const Staged wr = want_rev(vlc_start & ~7u, rev_words0, kVlcSlots);
        stage_ring<kMsSlots>(wave_ms, arena, wm, have_ms);
        stage_ring<kVlcSlots>(wave_vlc, arena, wr, have_vlc);
        stage_ring<kMelSlots>(wave_mel, arena, wl, have_mel);
        asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
        have_ms = wm;
        have_vlc = wr;
        have_mel = wl;
    }

    Fwd4 fms, fmel;
    Rev4 rvlc;
    fwd4_init(fms, ms_off, ms_len, 0xFF, RingSrc<kMsSlots>{ring_ms, have_ms, arena});
    fwd4_init(fmel, mel_off, mel_len, 0xFF, RingSrc<kMelSlots>{ring_mel, have_mel, arena});
    rev4_init(rvlc, vlc_start, vlc_len, RingSrc<kVlcSlots>{ring_vlc, have_vlc, arena});
    MsBits ms{0, 0, 0};
    VlcBits vlc;
    Mel mel = mel_make(mel_len);
    {
        const uint32_t b0 = run_block ? (uint32_t)arena[j.src + lcup - 2] : 0;
        vlc.tmp = b0 >> 4;
        vlc.bits = 4 - (((vlc.tmp & 7) == 7) ? 1u : 0u);
        vlc.unstuff = ((b0 | 0x0F) > 0x8F) ? 1u : 0u;
    }
    const uint32_t p = 30 - zblk, mmsbp2 = zblk + 2;
    const uint32_t kshift = kmax >= 31 ? 0u : 31u - kmax;  // sign-magnitude word -> sample
    const uint32_t nq = (w + 1) >> 1;
    const uint32_t npairs = (w + 3) >> 2, nrows = (h + 1) >> 1;
    const uint32_t nseg = (npairs + kSegPairs - 1) / kSegPairs;
    const uint32_t total = run_block ? nrows * nseg : 0;
    if (run_block)
        for (uint32_t q = 0; q < nq + 2; ++q) {
            s_q[q][tid] = 0;
        }
    int run = 0;
    if (run_block) run = mel_run_wave(mel, fmel, RingSrc<kMelSlots>{ring_mel, have_mel, arena});
    uint32_t c_q = 0, prev_e3 = 0;
    const uint32_t segs = wave_max(total);

    for (uint32_t g = 0; g < segs; ++g) {
        if (g) {
            // uniform point: top up all rings of the wave
            const Staged wm = want_fwd(fms.p, fms.words, kMsSlots);
            const Staged wl = want_fwd(fmel.p, fmel.words, kMelSlots);
            const Staged wr = want_rev(rvlc.p, rvlc.words, kVlcSlots);
            stage_ring<kMsSlots>(wave_ms, arena, wm, have_ms);
            stage_ring<kVlcSlots>(wave_vlc, arena, wr, have_vlc);
            stage_ring<kMelSlots>(wave_mel, arena, wl, have_mel);
            asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
            have_ms = wm;
            have_vlc = wr;
            have_mel = wl;
        }
        if (g >= total || st != kCbOk) continue;
        const uint32_t row = g / nseg, seg = g - row * nseg;
        const uint32_t y = 2 * row;
        const bool first_row = y == 0;
        const bool second = y + 1 < h;
        if (seg == 0) {
            c_q = 0;
            prev_e3 = 0;
        }
        const uint32_t pa = seg * kSegPairs, pb = min(pa + kSegPairs, npairs);
        for (uint32_t pr = pa; pr < pb; ++pr) {
            const uint32_t x = 4 * pr, qj = 2 * pr;
            // ---- cleanup symbols of the quad pair ----
            // quad state of the row above: significance bits 1 and 3 of
            // the quad at 0x20 / 0x80 (the only ones the contexts use) and
            // the exponent bound of its column pair in bits 0..4
            const uint32_t pt0 = s_q[qj][tid], pt1 = s_q[qj + 1][tid], pt2 = s_q[qj + 2][tid];
            uint32_t t0, t1, u0, u1;
            // one fill per pair: a pair takes at most 7 + 7 + 3 + 10 bits
            vlc_fill(vlc, rvlc, RingSrc<kVlcSlots>{ring_vlc, have_vlc, arena});
            if (first_row) {
                t0 = vlc0[c_q + (vlc.tmp & 0x7F)];
                if (c_q == 0) {
                    run -= 2;
                    t0 = (run == -1) ? t0 : 0;
                    if (run < 0) run = mel_run_wave(mel, fmel, RingSrc<kMelSlots>{ring_mel, have_mel, arena});
                }
                c_q = ((t0 & 0x10) << 3) | ((t0 & 0xE0) << 2);
                vlc.tmp >>= t0 & 7;
                vlc.bits -= t0 & 7;
                t1 = vlc0[c_q + (vlc.tmp & 0x7F)];
                if (c_q == 0 && x + 2 < w) {
                    run -= 2;
                    t1 = (run == -1) ? t1 : 0;
                    if (run < 0) run = mel_run_wave(mel, fmel, RingSrc<kMelSlots>{ring_mel, have_mel, arena});
                }
                if (x + 2 >= w) t1 = 0;
                c_q = ((t1 & 0x10) << 3) | ((t1 & 0xE0) << 2);
                vlc.tmp >>= t1 & 7;
                vlc.bits -= t1 & 7;
                uint32_t mode = ((t0 & 8) << 3) | ((t1 & 8) << 4);
                if (mode == 0xC0) {
                    run -= 2;
                    if (run == -1) mode += 0x40;
                    if (run < 0) run = mel_run_wave(mel, fmel, RingSrc<kMelSlots>{ring_mel, have_mel, arena});
                }
                uint32_t e = uvlc0[mode + (vlc.tmp & 0x3F)];
                vlc.tmp >>= e & 7;
                vlc.bits -= e & 7;
                e >>= 3;
                uint32_t len = e & 0xF;
                const uint32_t tmp = (uint32_t)vlc.tmp & ((1u << len) - 1);
                vlc.tmp >>= len;
                vlc.bits -= len;
                e >>= 4;
                len = e & 7;
                e >>= 3;
                u0 = 1 + (e & 7) + (tmp & ~(0xFFu << len));
                u1 = 1 + (e >> 3) + (tmp >> len);
            } else {
                c_q |= (pt0 & 0xA0) << 2;
                c_q |= (pt1 & 0x20) << 4;
                t0 = vlc1[c_q + (vlc.tmp & 0x7F)];
                if (c_q == 0) {
                    run -= 2;
                    t0 = (run == -1) ? t0 : 0;
                    if (run < 0) run = mel_run_wave(mel, fmel, RingSrc<kMelSlots>{ring_mel, have_mel, arena});
                }
                c_q = ((t0 & 0x40) << 2) | ((t0 & 0x80) << 1);
                c_q |= pt0 & 0x80;
                c_q |= (pt1 & 0xA0) << 2;
                c_q |= (pt2 & 0x20) << 4;
                vlc.tmp >>= t0 & 7;
                vlc.bits -= t0 & 7;
                t1 = vlc1[c_q + (vlc.tmp & 0x7F)];
                if (c_q == 0 && x + 2 < w) {
                    run -= 2;
                    t1 = (run == -1) ? t1 : 0;
                    if (run < 0) run = mel_run_wave(mel, fmel, RingSrc<kMelSlots>{ring_mel, have_mel, arena});
                }
                if (x + 2 >= w) t1 = 0;
                c_q = ((t1 & 0x40) << 2) | ((t1 & 0x80) << 1);
                c_q |= pt1 & 0x80;
                vlc.tmp >>= t1 & 7;
                vlc.bits -= t1 & 7;
                const uint32_t mode = ((t0 & 8) << 3) | ((t1 & 8) << 4);
                uint32_t e = uvlc1[mode + (vlc.tmp & 0x3F)];
                vlc.tmp >>= e & 7;
                vlc.bits -= e & 7;
                e >>= 3;
                uint32_t len = e & 0xF;
                const uint32_t tmp = (uint32_t)vlc.tmp & ((1u << len) - 1);
                vlc.tmp >>= len;
                vlc.bits -= len;
                e >>= 4;
                len = e & 7;
                e >>= 3;
                u0 = (e & 7) + (tmp & ~(0xFFu << len));
                u1 = (e >> 3) + (tmp >> len);
            }

            // ---- magnitude / sign of both quads ----
            uint32_t r0[4] = {0, 0, 0, 0}, r1[4] = {0, 0, 0, 0};  // rows y, y+1 at x..x+3
#pragma unroll
            for (uint32_t k = 0; k < 2; ++k) {
                const uint32_t xq = x + 2 * k;
                if (xq >= w) break;
                const uint32_t q = qj + k;
                const uint32_t inf = k ? t1 : t0;
                uint32_t uq = k ? u1 : u0;
                if (!first_row) {
                    uint32_t gamma = inf & 0xF0;
                    gamma &= gamma - 0x10;
                    uint32_t emax = max((uint32_t)s_q[q][tid] & 0x1Fu, (uint32_t)s_q[q + 1][tid] & 0x1Fu);
                    emax = max(emax, 1u);
                    uq += gamma ? emax : 1u;
                }
                if (uq > mmsbp2) {
                    st = kCbBad;
                    break;
                }
                // Bits the quad's significant samples take (m_n = U_q or
                // U_q - 1 each).  Up to 32 of them come in one top-up (to one
                // bit more: a sample reads its sign from the buffer's next bit
                // even when m_n is 0), and the per-sample top-ups below stay
                // idle; larger quads top up to 32 bits before every sample.
                const uint32_t sig = (inf >> 4) & (xq + 1 < w ? 0xFu : 0x3u);
                const uint32_t qbits = (uint32_t)__builtin_popcount(sig) * uq -
                                       (uint32_t)__builtin_popcount(sig & (inf >> 12));
                const bool one_fill = qbits <= 32;
                const uint32_t need = sig ? qbits + 1u : 0u;
                ms_topup(ms, fms, RingSrc<kMsSlots>{ring_ms, have_ms, arena}, one_fill ? need : 0u);
                uint32_t v0, v1, v2, v3;
                if (!one_fill) ms_fill(ms, fms, RingSrc<kMsSlots>{ring_ms, have_ms, arena});
                const int32_t a0 = ms_sample(ms, inf, 0, uq, p, kshift, v0);
                if (!one_fill) ms_fill(ms, fms, RingSrc<kMsSlots>{ring_ms, have_ms, arena});
                const int32_t a1 = ms_sample(ms, inf, 1, uq, p, kshift, v1);
                r0[2 * k] = (uint32_t)a0;
                r1[2 * k] = (uint32_t)a1;
                s_q[q][tid] = (uint8_t)((inf & 0xA0u) | max(prev_e3, flog2r(v1)));
                prev_e3 = 0;
                if (xq + 1 >= w) break;
                if (!one_fill) ms_fill(ms, fms, RingSrc<kMsSlots>{ring_ms, have_ms, arena});
                const int32_t a2 = ms_sample(ms, inf, 2, uq, p, kshift, v2);
                if (!one_fill) ms_fill(ms, fms, RingSrc<kMsSlots>{ring_ms, have_ms, arena});
                const int32_t a3 = ms_sample(ms, inf, 3, uq, p, kshift, v3);
                r0[2 * k + 1] = (uint32_t)a2;
                r1[2 * k + 1] = (uint32_t)a3;
                prev_e3 = flog2r(v3);
            }
            C *o = out + y * stride + x;
            if (chunk_store && (x | (kChunkCols - 1)) < w) {
                // the pair's samples join the row chunks; a full chunk goes
                // out as two 16-byte stores per row
                constexpr int kW = kCW / kChunkPairs;  // words per row per pair
#pragma unroll
                for (int t = 0; t < (int)kCW - kW; ++t) {
                    ch0[t] = ch0[t + kW];
                    ch1[t] = ch1[t + kW];
                }
                if constexpr (sizeof(C) == 4) {
#pragma unroll
                    for (int t = 0; t < 4; ++t) {
                        ch0[kCW - 4 + t] = r0[t];
                        ch1[kCW - 4 + t] = r1[t];
                    }
                } else {
                    ch0[kCW - 2] = (r0[0] & 0xFFFFu) | (r0[1] << 16);
                    ch0[kCW - 1] = (r0[2] & 0xFFFFu) | (r0[3] << 16);
                    ch1[kCW - 2] = (r1[0] & 0xFFFFu) | (r1[1] << 16);
                    ch1[kCW - 1] = (r1[2] & 0xFFFFu) | (r1[3] << 16);
                }
                if ((pr & (kChunkPairs - 1)) == kChunkPairs - 1) {
                    uint4 *c0 = reinterpret_cast<uint4 *>(out + y * stride + (x & ~(kChunkCols - 1)));
#pragma unroll
                    for (uint32_t t = 0; t < kCW / 4; ++t)
                        c0[t] = make_uint4(ch0[4 * t], ch0[4 * t + 1], ch0[4 * t + 2], ch0[4 * t + 3]);
                    if (second) {
                        uint4 *c1 = reinterpret_cast<uint4 *>(reinterpret_cast<C *>(c0) + stride);
#pragma unroll
                        for (uint32_t t = 0; t < kCW / 4; ++t)
                            c1[t] = make_uint4(ch1[4 * t], ch1[4 * t + 1], ch1[4 * t + 2], ch1[4 * t + 3]);
                    }
                }
            } else if (vec_store && x + 4 <= w) {
                if constexpr (sizeof(C) == 4) {
                    *reinterpret_cast<uint4 *>(o) = make_uint4(r0[0], r0[1], r0[2], r0[3]);
                    if (second)
                        *reinterpret_cast<uint4 *>(o + stride) = make_uint4(r1[0], r1[1], r1[2], r1[3]);
                } else {
                    *reinterpret_cast<uint2 *>(o) = make_uint2((r0[0] & 0xFFFFu) | (r0[1] << 16),
                                                               (r0[2] & 0xFFFFu) | (r0[3] << 16));
                    if (second)
                        *reinterpret_cast<uint2 *>(o + stride) =
                            make_uint2((r1[0] & 0xFFFFu) | (r1[1] << 16), (r1[2] & 0xFFFFu) | (r1[3] << 16));
                }
            } else {
#pragma unroll
                for (uint32_t c = 0; c < 4; ++c) {
                    if (x + c >= w) break;
                    o[c] = (C)(int32_t)r0[c];
                    if (second) o[stride + c] = (C)(int32_t)r1[c];
                }
            }
            if (st != kCbOk) break;
        }
        if (pb == npairs) s_q[nq][tid] = (uint8_t)prev_e3;
    }
    if (valid) status[i] = st;
}

// The lane kernel.  PERSIST: a grid sized to the resident capacity whose
// waves take the next CHUNK code-blocks (in the host's cost order, a
// multiple of 64: one block per lane, CHUNK / 64 rounds) from the device
// work counter *queue until none are left; otherwise one wave per 64
// code-blocks.  *queue must be zero at launch.
template <int WAVES, int SEG, int MS, int VS, int LS, bool PERSIST, typename C, int MINW = 1,
          int CHUNK = 64>
__global__ void __launch_bounds__(WAVES * 64, MINW)
sgn_ht_rows(const CbJob *__restrict__ jobs, uint32_t n, const uint8_t *__restrict__ arena,
            C *__restrict__ coef, const uint16_t *__restrict__ tables,
            uint8_t *__restrict__ status, uint32_t *__restrict__ queue) {
    __shared__ uint16_t s_tab[1024 + 1024 + 320 + 256];
    __shared__ uint8_t s_q[kRowsMaxQ][WAVES * 64];  // quad state of the row above
    __shared__ uint4 s_ms[WAVES * 64 * MS];
    __shared__ uint4 s_vlc[WAVES * 64 * VS];
    __shared__ uint4 s_mel[WAVES * 64 * LS];
    for (int t = threadIdx.x; t < 2624; t += WAVES * 64) s_tab[t] = tables[t];
    __syncthreads();
    if constexpr (!PERSIST) {
        const uint32_t first = (blockIdx.x * WAVES + (threadIdx.x >> 6)) * 64;
        rows_wave<WAVES, SEG, MS, VS, LS, C>(jobs, n, first, arena, coef, status, s_tab, s_q, s_ms,
                                  s_vlc, s_mel);
    } else {
        static_assert(CHUNK % 64 == 0, "a chunk is whole waves of code-blocks");
        for (;;) {
            uint32_t first = 0;
            if ((threadIdx.x & 63) == 0) first = atomicAdd(queue, (uint32_t)CHUNK);
            first = __shfl(first, 0);
            if (first >= n) break;
            for (uint32_t c = 0; c < (uint32_t)CHUNK && first + c < n; c += 64)
                rows_wave<WAVES, SEG, MS, VS, LS, C>(jobs, n, first + c, arena, coef, status, s_tab,
                                          s_q, s_ms, s_vlc, s_mel);
        }
    }
}

// SigProp / MagRef passes of the code-blocks jobs[0 .. n) (each with more
// than one pass, at most 64x64), one block per lane, after the lane kernel
// wrote their cleanup samples.  A block whose cleanup failed holds zeros or
// partial samples; the passes keep to its rectangle and its refinement
// segment whatever the samples, and its status already reports it.
template <typename C>
__global__ void __launch_bounds__(256)
sgn_ht_refine(const CbJob *__restrict__ jobs, uint32_t n, const uint8_t *__restrict__ arena,
              C *__restrict__ coef) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const CbJob j = jobs[i];
    // pass count rules of decode_impl (ht_block_decode.rs:1578-1592); more
    // than three passes never reach this kernel
    if ((j.flags & kCbZero) || j.passes < 2 || j.passes > 3 || j.lref == 0 || j.zblk >= 29 ||
        j.lcup < 2 || j.w > 64 || j.h > 64)
        return;
    const uint32_t shift = j.kmax >= 31 ? 0u : 31u - j.kmax;
    rows_refine<C>(coef + j.dst, j.w, j.h, j.stride, ((j.dst | j.stride) & 3u) == 0, arena,
                   j.src + j.lcup, j.lref, j.passes, (j.flags & kCbStripeCausal) != 0,
                   30u - j.zblk, shift);
}

// Timing run for the decode-cost model: of each wave's 64 jobs only lane
// 0's is a code-block (the others are empty), and the wave records the
// shader-clock cycles it spent decoding it.
template <typename C>
__global__ void __launch_bounds__(256)
sgn_ht_rows_timed(const CbJob *__restrict__ jobs, uint32_t n, const uint8_t *__restrict__ arena,
                  C *__restrict__ coef, const uint16_t *__restrict__ tables,
                  uint8_t *__restrict__ status, uint64_t *__restrict__ cycles) {
    constexpr int W = 4, S = 8, M = 8, VS = 4, LS = 2;
    __shared__ uint16_t s_tab[1024 + 1024 + 320 + 256];
    __shared__ uint8_t s_q[kRowsMaxQ][W * 64];
    __shared__ uint4 s_ms[W * 64 * M];
    __shared__ uint4 s_vlc[W * 64 * VS];
    __shared__ uint4 s_mel[W * 64 * LS];
    for (int t = threadIdx.x; t < 2624; t += W * 64) s_tab[t] = tables[t];
    __syncthreads();
    const uint32_t wave = blockIdx.x * W + (threadIdx.x >> 6);
    const uint64_t t0 = clock64();
    rows_wave<W, S, M, VS, LS, C>(jobs, n, wave * 64, arena, coef, status, s_tab, s_q, s_ms, s_vlc,
                       s_mel);
    const uint64_t t1 = clock64();
    if ((threadIdx.x & 63) == 0 && wave * 64 < n) cycles[wave] = t1 - t0;
}

}  // namespace sgn
