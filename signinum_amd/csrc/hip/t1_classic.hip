// Classic (EBCOT) code-block decode: the MQ arithmetic decoder (T.800
// Annex C) and the significance-propagation, magnitude-refinement and
// cleanup passes of Annex D, with the code-block styles selective bypass,
// context reset, termination on each pass, vertically causal context and
// segmentation symbols.  Follows
//   crates/signinum-j2k-native/src/j2c/bitplane.rs:90-317, 903-1116,
//   1228-1370 and arithmetic_decoder.rs:42-250, 329-386.
//
// Decomposition: one code-block per lane.  The MQ decoder is one serial
// chain per block (every decision updates the interval and the context
// the next decision reads), so the parallelism is across blocks.  A block's
// coding state is one 16-bit word per stripe column (4 samples:
// significance, sign, visited-in-this-bit-plane and refined-before bits),
// kept in a row store in HBM (one row of w + 2 words per stripe, zero
// columns at both ends).  A pass walks the stripes with three rows in LDS
// (the stripe above, this one, the one below: a ring, the row below loaded
// when a stripe starts and the current row written back when it ends), so
// a lane needs a few hundred bytes of LDS and several work-groups share a
// CU; within a stripe the 3x3 words around the current column sit in
// registers (three LDS loads per column step, one store).  Magnitude bits
// go to the block's int32 coefficients in HBM by atomic OR (no read in the
// decode chain); the block is zeroed first and the signs are applied in a
// final sweep.  Each context state is its packed Qe-table entry plus the
// MPS bit, so the common decision (MPS, no renormalisation) is one LDS
// read.
#include <hip/hip_runtime.h>

#include "../host/plan_types.h"

namespace sgn {

// Table C.2: Qe << 16 | index after an MPS << 8 | index after an LPS << 1 |
// MPS switch.
__constant__ uint32_t kMqTable[47] = {
    0x56010103u, 0x3401020Cu, 0x18010312u, 0x0AC10418u, 0x0521053Au, 0x02212642u, 0x5601070Du,
    0x5401081Cu, 0x4801091Cu, 0x38010A1Cu, 0x30010B22u, 0x24010C24u, 0x1C010D28u, 0x16011D2Au,
    0x56010F1Du, 0x5401101Cu, 0x5101111Eu, 0x48011220u, 0x38011322u, 0x34011424u, 0x30011526u,
    0x28011626u, 0x24011728u, 0x2201182Au, 0x1C01192Cu, 0x18011A2Eu, 0x16011B30u, 0x14011C32u,
    0x12011D34u, 0x11011E36u, 0x0AC11F38u, 0x09C1203Au, 0x08A1213Cu, 0x0521223Eu, 0x04412340u,
    0x02A12442u, 0x02212544u, 0x01412646u, 0x01112748u, 0x0085284Au, 0x0049294Cu, 0x00252A4Eu,
    0x00152B50u, 0x00092C52u, 0x00052D54u, 0x00012D56u, 0x56012E5Cu,
};

constexpr int kClassicContexts = 19;
constexpr int kCxRunLength = 17, kCxUniform = 18;
constexpr uint32_t kMpsBit = 1u << 15;  // MPS bit of a context word (bits 14-15 of an entry are 0)

// Dwords of one stripe row of a block w wide: w + 2 words, padded to even.
__host__ __device__ inline uint32_t classic_row_dwords(uint32_t w) { return (w + 3) >> 1; }
// Dwords of a block's row store: one row per stripe.
__host__ __device__ inline uint32_t classic_state_dwords(uint32_t w, uint32_t h) {
    return classic_row_dwords(w) * ((h + 3) >> 2);
}
// Row-store dwords reserved per lane: the largest block of any shape
// (4 x 1024: 3 x 256 dwords).
constexpr uint32_t kClassicStateDwords = 768;

// Kernel shapes: dwords of one LDS ring row (blocks up to kMaxW wide) and
// the active lanes of the 64-thread work-group.  64 lanes fill a wave; with
// fewer, a batch with few blocks spreads them over more waves (every SIMD
// gets some, and a SIMD's waves hide each other's LDS and memory waits).
template <int ROWDW, int LANES, int MAXW>
struct ClassicShape {
    static constexpr int kRowDw = ROWDW, kLanes = LANES, kMaxW = MAXW;
};
template <int LANES>
using ClassicW32 = ClassicShape<17, LANES, 32>;
template <int LANES>
using ClassicW64 = ClassicShape<33, LANES, 64>;
template <int LANES>
using ClassicW256 = ClassicShape<129, LANES, 256>;
using ClassicW1024 = ClassicShape<513, 16, 1024>;

// Reads of one codeword segment: absolute arena positions [p, end), bytes
// past `end` read as 0xFF (the MQ decoder's fill, arithmetic_decoder.rs:
// 250-270).  A window of three aligned dwords in registers; the third is
// fetched one dword ahead of use.
struct SegBytes {
    const uint8_t *arena;
    uint32_t end, wpos;  // wpos: arena offset of w0 (multiple of 4)
    uint32_t w0, w1, w2;

    __device__ void start(const uint8_t *a, uint32_t p, uint32_t e) {
        arena = a;
        end = e;
        wpos = p & ~3u;
        const uint32_t *q = reinterpret_cast<const uint32_t *>(a + wpos);
        w0 = q[0];
        w1 = q[1];
        w2 = q[2];
    }
    // byte at p (p >= wpos); slides the window so that p + 1 stays in w0..w1
    __device__ uint32_t at(uint32_t p) {
        if (p + 1 >= wpos + 8) {
            w0 = w1;
            w1 = w2;
            wpos += 4;
            w2 = *reinterpret_cast<const uint32_t *>(arena + wpos + 8);
        }
        const uint32_t rel = p - wpos;
        const uint32_t w = rel < 4 ? w0 : w1;
        const uint32_t b = (w >> ((rel & 3u) * 8u)) & 0xFFu;
        return p < end ? b : 0xFFu;
    }
};

// MQ decoder, register C kept inverted (ITU-T T.88 Annex G conventions).
struct Mq {
    uint32_t c, a, ct, p;
    SegBytes in;

    __device__ void bytein() {
        const uint32_t b = in.at(p);
        const uint32_t b1 = in.at(p + 1);
        if (b == 0xFFu) {
            if (b1 > 0x8Fu) {
                ct = 8;  // a marker or the end: 1 bits from here on
            } else {
                ++p;
                c += 0xFE00u - (b1 << 9);
                ct = 7;
            }
        } else {
            ++p;
            c += 0xFF00u - (b1 << 8);
            ct = 8;
        }
    }
    __device__ void init(const uint8_t *arena, uint32_t start, uint32_t len) {
        in.start(arena, start, start + len);
        p = start;
        c = (in.at(p) ^ 0xFFu) << 16;
        bytein();
        c <<= 7;
        ct -= 7;
        a = 0x8000u;
    }
    // decision with context word *cx (its Qe-table entry | MPS bit)
    __device__ uint32_t decode(uint32_t *cx, const uint32_t *qtab) {
        const uint32_t e = *cx;
        const uint32_t qe = e >> 16;
        uint32_t mps = (e >> 15) & 1u;
        a -= qe;
        if ((c >> 16) < a && (a & 0x8000u)) return mps;
        const bool lps = (c >> 16) >= a;
        if (lps) c -= a << 16;
        // conditional exchange (C.3.2): the decision is the LPS, and the
        // state moves on the LPS side, exactly when (a < qe) differs from
        // the sub-interval the code value fell in
        const bool took_lps = (a < qe) != lps;
        const uint32_t d = took_lps ? 1u - mps : mps;
        if (lps) a = qe;
        if (took_lps && (e & 1u)) mps ^= 1u;
        const uint32_t next = took_lps ? (e >> 1) & 0x7Fu : (e >> 8) & 0x3Fu;
        *cx = qtab[next] | (mps << 15);
        do {
            if (ct == 0) bytein();
            const uint32_t sh = min(ct, uint32_t(__builtin_clz(a) - 16));
            a <<= sh;
            c <<= sh;
            ct -= sh;
        } while (!(a & 0x8000u));
        return d;
    }
};

// Raw (bypass) bits with B.10.1 stuffing; past the segment, or a stuffed
// bit that is not 0, yields 1 (bitplane.rs:1339-1370, non-strict).
struct RawBits {
    SegBytes in;
    uint32_t p, bit;  // byte position, bits consumed in it

    __device__ void init(const uint8_t *arena, uint32_t start, uint32_t len) {
        in.start(arena, start, start + len);
        p = start;
        bit = 0;
    }
    __device__ uint32_t get() {
        if (p >= in.end) return 1;
        const uint32_t byte = in.at(p);
        const uint32_t b = (byte >> (7 - bit)) & 1u;
        if (++bit == 8) {
            bit = 0;
            ++p;
        }
        if (bit == 0 && byte == 0xFFu) {  // the next byte's MSB is a stuffed 0
            if (p >= in.end) return 1;
            const uint32_t s = in.at(p) >> 7;
            bit = 1;
            if (s) return 1;
        }
        return b;
    }
};

// Table D.1 from neighbour counts (h, v, d; `swap` for HL, where the roles
// of horizontal and vertical swap; `diag` for HH).
__device__ inline uint32_t zc_label(uint32_t h, uint32_t v, uint32_t d, bool swap, bool diag) {
    if (diag) {
        const uint32_t hv = h + v;
        if (d >= 3) return 8;
        if (d == 2) return hv ? 7 : 6;
        if (d == 1) return hv >= 2 ? 5 : (hv ? 4 : 3);
        return hv >= 2 ? 2 : hv;
    }
    if (swap) {
        const uint32_t t = h;
        h = v;
        v = t;
    }
    if (h == 2) return 8;
    if (h == 1) return v ? 7 : (d ? 6 : 5);
    if (v == 2) return 4;
    if (v == 1) return 3;
    return d >= 2 ? 2 : d;
}

// Stripe-column word bits: significance rows 0-3, sign (negative) 4-7,
// visited in this bit-plane's significance pass 8-11, refined before 12-15.
constexpr uint32_t kSig = 0, kNeg = 4, kVisit = 8, kRef = 12;

// Six-entry column vector of one flag: the row above the stripe (bit 0),
// the stripe's rows (1-4), the row below (5).
__device__ inline uint32_t col6(uint32_t up, uint32_t mid, uint32_t dn, uint32_t f) {
    return ((up >> (f + 3)) & 1u) | (((mid >> f) & 0xFu) << 1) | (((dn >> f) & 1u) << 5);
}

template <class G>
struct ClassicLane {
    uint32_t *ring;     // this lane's three LDS rows (lane-interleaved dwords: stride kLanes)
    uint32_t *cx;       // this lane's 19 context words (stride kLanes)
    uint32_t *rows;     // this lane's row store in HBM (ndw dwords per stripe)
    const uint32_t *qtab;
    int32_t *coef;
    uint32_t dst, stride, w, h, ns, ndw, band;
    bool causal, raw;
    uint32_t bit;
    Mq mq;
    RawBits rb;

    __device__ uint16_t *word_at(uint32_t slot, uint32_t x1) {
        return reinterpret_cast<uint16_t *>(ring + (slot * G::kRowDw + (x1 >> 1)) * G::kLanes) +
               (x1 & 1u);
    }
    __device__ uint32_t word(uint32_t slot, uint32_t x1) { return *word_at(slot, x1); }
    __device__ uint32_t &ctx(uint32_t label) { return cx[label * G::kLanes]; }

    // ring slot of stripe s (-1 .. ns)
    __device__ static uint32_t slot_of(int s) { return uint32_t(s + 3) % 3u; }
    __device__ void load_row(int s) {
        uint32_t *r = ring + slot_of(s) * G::kRowDw * G::kLanes;
        if (s < 0 || uint32_t(s) >= ns) {
            for (uint32_t k = 0; k < ndw; ++k) r[k * G::kLanes] = 0;
            return;
        }
        const uint32_t *src = rows + uint32_t(s) * ndw;
        uint32_t k = 0;
        for (; k + 8 <= ndw; k += 8) {
            uint32_t v[8];
#pragma unroll
            for (int q = 0; q < 8; ++q) v[q] = src[k + q];
#pragma unroll
            for (int q = 0; q < 8; ++q) r[(k + q) * G::kLanes] = v[q];
        }
        for (; k < ndw; ++k) r[k * G::kLanes] = src[k];
    }
    __device__ void store_row(int s) {
        const uint32_t *r = ring + slot_of(s) * G::kRowDw * G::kLanes;
        uint32_t *d = rows + uint32_t(s) * ndw;
        for (uint32_t k = 0; k < ndw; ++k) d[k] = r[k * G::kLanes];
    }

    __device__ void reset_contexts() {
        const uint32_t z = qtab[0];
        for (int k = 0; k < kClassicContexts; ++k) ctx(k) = z;
        ctx(0) = qtab[4];
        ctx(kCxRunLength) = qtab[3];
        ctx(kCxUniform) = qtab[46];
    }
    __device__ uint32_t decide(uint32_t label) {
        return raw ? rb.get() : mq.decode(&ctx(label), qtab);
    }
    __device__ void set_mag(uint32_t x, uint32_t y) {
        atomicOr(reinterpret_cast<unsigned int *>(coef + dst + y * stride + x), 1u << bit);
    }

    // Sign of the sample at row r (column vectors of significance and sign
    // of the left, centre and right columns), Table D.3.
    __device__ uint32_t decode_sign(uint32_t sL, uint32_t sC, uint32_t sR, uint32_t nL, uint32_t nC,
                                    uint32_t nR, uint32_t r) {
        auto contrib = [](uint32_t sv, uint32_t nv, uint32_t at) -> int {
            return ((sv >> at) & 1u) ? (((nv >> at) & 1u) ? -1 : 1) : 0;
        };
        const int hs = contrib(sL, nL, r + 1) + contrib(sR, nR, r + 1);
        const int vs = contrib(sC, nC, r) + contrib(sC, nC, r + 2);
        const int H = hs > 0 ? 1 : (hs < 0 ? -1 : 0), V = vs > 0 ? 1 : (vs < 0 ? -1 : 0);
        uint32_t label, flip;
        if (H == 0) {
            label = V == 0 ? 9u : 10u;
            flip = V < 0;
        } else {
            label = uint32_t(12 + H * V);
            flip = H < 0;
        }
        const uint32_t s = decide(label);
        return raw ? s : (s ^ flip);  // a raw sign bit is the sign itself
    }

    // One coding pass over the block.  KIND 0 cleanup (also clears the
    // visited bits once a stripe is done), 1 significance propagation,
    // 2 magnitude refinement.
    template <int KIND>
    __device__ void pass() {
        const bool diag = band == 3, swap = band == 1;
        load_row(-1);
        load_row(0);
        for (uint32_t s = 0; s < ns; ++s) {
            load_row(int(s) + 1);
            const uint32_t su = slot_of(int(s) - 1), sm = slot_of(int(s)), sd = slot_of(int(s) + 1);
            const uint32_t rows4 = min(4u, h - 4 * s);
            const uint32_t y0 = 4 * s;
            // words of the previous (L), current (C) and next (R) column in
            // the stripes above, this and below
            uint32_t uL = 0, mL = 0, dL = 0;
            uint32_t uC = word(su, 1), mC = word(sm, 1), dC = word(sd, 1);
            for (uint32_t x = 0; x < w; ++x) {
                const uint32_t uR = word(su, x + 2), mR = word(sm, x + 2), dR = word(sd, x + 2);
                const uint32_t dnL = causal ? 0u : dL, dnC = causal ? 0u : dC, dnR = causal ? 0u : dR;
                const uint32_t sL = col6(uL, mL, dnL, kSig), sR = col6(uR, mR, dnR, kSig);
                const uint32_t before = mC;
                uint32_t r = 0;
                if (KIND == 0 && rows4 == 4 && ((mC >> kSig) & 0xFu) == 0 &&
                    ((mC >> kVisit) & 0xFu) == 0 && sL == 0 && sR == 0 &&
                    col6(uC, 0, dnC, kSig) == 0) {
                    // run-length mode (D.3.4)
                    if (!decide(kCxRunLength)) {
                        r = 4;
                    } else {
                        r = decide(kCxUniform) << 1;
                        r |= decide(kCxUniform);
                        set_mag(x, y0 + r);
                        const uint32_t sC = col6(uC, mC, dnC, kSig);
                        const uint32_t neg = decode_sign(sL, sC, sR, col6(uL, mL, dnL, kNeg),
                                                         col6(uC, mC, dnC, kNeg),
                                                         col6(uR, mR, dnR, kNeg), r);
                        mC |= (1u << (kSig + r)) | (neg << (kNeg + r));
                        ++r;
                    }
                }
                for (; r < rows4; ++r) {
                    const uint32_t sC = col6(uC, mC, dnC, kSig);
                    const uint32_t sig = (mC >> (kSig + r)) & 1u, vis = (mC >> (kVisit + r)) & 1u;
                    const uint32_t hc = ((sL >> (r + 1)) & 1u) + ((sR >> (r + 1)) & 1u);
                    const uint32_t vc = ((sC >> r) & 1u) + ((sC >> (r + 2)) & 1u);
                    const uint32_t dc = ((sL >> r) & 1u) + ((sL >> (r + 2)) & 1u) +
                                        ((sR >> r) & 1u) + ((sR >> (r + 2)) & 1u);
                    if (KIND == 2) {
                        if (!sig || vis) continue;
                        const uint32_t label =
                            ((mC >> (kRef + r)) & 1u) ? 16u : (hc + vc + dc ? 15u : 14u);
                        if (decide(label)) set_mag(x, y0 + r);
                        mC |= 1u << (kRef + r);
                        continue;
                    }
                    if (sig) continue;
                    if (KIND == 1) {
                        if (hc + vc + dc == 0) continue;
                        mC |= 1u << (kVisit + r);
                    } else if (vis) {
                        continue;
                    }
                    if (!decide(zc_label(hc, vc, dc, swap, diag))) continue;
                    set_mag(x, y0 + r);
                    const uint32_t neg =
                        decode_sign(sL, sC, sR, col6(uL, mL, dnL, kNeg), col6(uC, mC, dnC, kNeg),
                                    col6(uR, mR, dnR, kNeg), r);
                    mC |= (1u << (kSig + r)) | (neg << (kNeg + r));
                }
                // the word as the next column sees it; the visited bits end
                // with the cleanup pass (only this sample reads them)
                if (mC != before) *word_at(sm, x + 1) = uint16_t(mC);
                uL = uC;
                mL = mC;
                dL = dC;
                uC = uR;
                mC = mR;
                dC = dR;
            }
            if (KIND == 0)
                for (uint32_t k = 0; k < ndw; ++k) ring[(sm * G::kRowDw + k) * G::kLanes] &= ~0x0F000F00u;
            store_row(int(s));
        }
    }

    // passes [p0, p1) of one segment (bitplane.rs:206-317)
    __device__ void run(uint32_t planes, uint32_t p0, uint32_t p1, uint32_t style) {
        for (uint32_t p = p0; p < p1; ++p) {
            bit = planes - 1 - (p + 2) / 3;
            const uint32_t kind = p % 3;
            // the previous pass's row stores land before this pass reads them
            __threadfence();
            if (kind == 0) {
                pass<0>();
                if (style & kCbsSegSym)  // four UNIFORM symbols, not checked (non-strict)
                    for (int k = 0; k < 4; ++k) (void)decide(kCxUniform);
            } else if (kind == 1) {
                pass<1>();
            } else {
                pass<2>();
            }
            if (style & kCbsReset) reset_contexts();
        }
    }
};

template <class G>
__device__ void classic_block(const ClassicJob &j, const uint8_t *arena, const ClassicSeg *segs,
                              int32_t *coef, uint32_t *ring, uint32_t *cx, uint32_t *rows,
                              const uint32_t *qtab) {
    const uint32_t w = j.w, h = j.h;
    // zero the block's coefficients (magnitude bits arrive by OR)
    for (uint32_t y = 0; y < h; ++y) {
        int32_t *row = coef + j.dst + y * j.stride;
        for (uint32_t x = 0; x < w; ++x) row[x] = 0;
    }
    if (!j.planes || !j.passes) return;
    ClassicLane<G> L;
    L.ring = ring;
    L.cx = cx;
    L.rows = rows;
    L.qtab = qtab;
    L.coef = coef;
    L.dst = j.dst;
    L.stride = j.stride;
    L.w = w;
    L.h = h;
    L.ns = (h + 3) >> 2;
    L.ndw = classic_row_dwords(w);
    L.band = j.band;
    L.causal = (j.style & kCbsCausal) != 0;
    L.raw = false;
    for (uint32_t k = 0; k < L.ns * L.ndw; ++k) rows[k] = 0;
    L.reset_contexts();
    const uint32_t maxp = 1 + 3 * (uint32_t(j.planes) - 1);
    if (j.nseg == 0) {
        L.mq.init(arena, j.src, j.len);
        L.run(j.planes, 0, min(uint32_t(j.passes), maxp), j.style);
    } else {
        for (uint32_t q = 0; q < j.nseg; ++q) {
            const ClassicSeg sg = segs[j.seg + q];
            L.raw = sg.raw != 0;
            if (L.raw)
                L.rb.init(arena, j.src + sg.off, sg.len);
            else
                L.mq.init(arena, j.src + sg.off, sg.len);
            L.run(j.planes, sg.start, min(uint32_t(sg.end), maxp), j.style);
        }
    }
    // signs: the magnitudes' OR results and the last row stores must be
    // visible to this lane's loads
    __threadfence();
    for (uint32_t y = 0; y < h; ++y) {
        const uint32_t *srow = rows + (y >> 2) * L.ndw;
        const uint32_t r = y & 3u;
        int32_t *row = coef + j.dst + y * j.stride;
        for (uint32_t x = 0; x < w; ++x) {
            const uint32_t wd = (srow[(x + 1) >> 1] >> (((x + 1) & 1u) * 16u)) & 0xFFFFu;
            if ((wd >> (kNeg + r)) & 1u) row[x] = -row[x];
        }
    }
}

// One code-block per lane; a grid of at most the resident work-groups
// walks the job groups (one group of kLanes blocks per work-group turn),
// each lane keeping its row store at rows + (blockIdx.x * kLanes + lane) *
// kClassicStateDwords.
template <class G>
__global__ void __launch_bounds__(64)
sgn_t1_classic(const ClassicJob *__restrict__ jobs, uint32_t n, const uint8_t *__restrict__ arena,
               const ClassicSeg *__restrict__ segs, int32_t *__restrict__ coef,
               uint32_t *__restrict__ rows) {
    __shared__ uint32_t ring[3 * G::kRowDw * G::kLanes];
    __shared__ uint32_t cx[kClassicContexts * G::kLanes];
    __shared__ uint32_t qtab[48];
    const uint32_t lane = threadIdx.x;
    if (lane < 47) qtab[lane] = kMqTable[lane];
    __syncthreads();
    if (lane >= uint32_t(G::kLanes)) return;
    uint32_t *my_rows = rows + (size_t(blockIdx.x) * G::kLanes + lane) * kClassicStateDwords;
    for (uint32_t g = blockIdx.x; g * G::kLanes < n; g += gridDim.x) {
        const uint32_t i = g * G::kLanes + lane;
        if (i >= n) break;
        const ClassicJob j = jobs[i];
        classic_block<G>(j, arena, segs, coef, ring + lane, cx + lane, my_rows, qtab);
    }
}

}  // namespace sgn
