// Inverse DWT kernels for gfx950, LDS-tiled.
//
// A work-group synthesises one 64x64 output block of one level.  It loads
// the block plus a halo (2 samples for 5/3, 4 for 9/7) straight from the
// four sub-bands, interleaving them by absolute parity (F.3.3) and applying
// whole-sample symmetric extension by reflection of the coordinates (F.3.7).
// Lifting on that window gives, in the block interior, exactly the values of
// the reference's in-place lifting over whole rows and columns
// (crates/signinum-j2k-native/src/j2c/idwt.rs:535-1143): rows first, then
// columns; 5/3 in integers (exact), 9/7 in f32 with fused multiply-add.
// Dequantisation is applied to the integer sub-band samples as they are
// loaded, so it is fused into the first synthesis level that reads them.
//
// sgn_final_kernel fuses the last synthesis level of every component with
// the inverse colour transform (mct.rs:71-156), the sign shift
// (decode.rs:1720-1752), rounding/saturation and interleaved packing
// (lib.rs:2202-2295, 1772-1835).
#include <hip/hip_runtime.h>

#include "../host/plan_types.h"

namespace sgn {

constexpr int kBlk = 64;
constexpr int kIdwtThreads = 256;

// Output block and work-group shape of a synthesis kernel: Big for every
// level and the final stage, Small (32x32 blocks, one wave) for level
// launches whose outputs are all at most 32x32, whose work-groups are
// latency-bound: more of them fit a CU.
template <int BLK, int THREADS>
struct Geo {
    static constexpr int kBlk = BLK, kThreads = THREADS;
    static_assert(BLK % 8 == 0 && THREADS % BLK == 0 && THREADS % 64 == 0, "block shape");
};
using Big = Geo<kBlk, kIdwtThreads>;
using Small = Geo<32, 64>;
// 64x64 blocks with two waves: half the threads per block, twice the
// work-groups per CU (a level tile shape the plan's autotuner can pick)
using Mid = Geo<kBlk, 128>;
constexpr float kKappa = 1.2301741f;
constexpr float kNegAlpha = 1.5861343f;
constexpr float kNegBeta = 0.052980117f;
constexpr float kNegGamma = -0.8829111f;
constexpr float kNegDelta = -0.44350687f;

__device__ __forceinline__ int reflect(int i, int n) {
    // one mirror at each end covers every halo of a level at least H wide
    int m = i < 0 ? -i : i;
    m = m >= n ? 2 * (n - 1) - m : m;
    if ((unsigned)m < (unsigned)n) return m;
    // narrower levels: periodic reflection
    if (n == 1) return 0;
    const int period = 2 * (n - 1);
    m = i % period;
    if (m < 0) m += period;
    return m < n ? m : period - m;
}

template <bool REV>
struct Sample;
template <>
struct Sample<true> {
    using T = int32_t;
};
template <>
struct Sample<false> {
    using T = float;
};

// Per-level constants.  The per-band tables live in LDS (indexed by the
// band of each sample); the rest are wave-uniform scalars.  Nothing is
// copied into a by-value struct that per-lane indexing would spill.
// Band slots are addressed in bytes from `LevelView::origin`, the lowest
// slot of the level (a tile-component's slots span less than 2 GiB, which
// the planner checks): relb = slot start, pitchb = row pitch, both in bytes.
struct BandLds {
    uint32_t relb[4];
    uint32_t pitchb[4];
    float delta[4];
};

struct LevelView {
    const BandLds *bands;
    const uint8_t *origin;
    uint32_t x0, y0, fin, narrow;
    int w, h, levels;
    bool aligned;  // every used slot and row pitch 16-byte aligned
    bool l53;      // f32 samples lifted with the 5/3 steps (LevelJob::lift53)
};

// Stages the band table of `lj` into `sb`; ends with a barrier.
__device__ __forceinline__ LevelView level_view(const LevelJob *__restrict__ lj, const uint32_t *coef,
                                               BandLds *sb) {
    const int t = threadIdx.x;
    const uint32_t narrow = lj->narrow;
    const uint32_t used = lj->levels ? 0xFu : 0x1u;
    auto slot_byte = [&](int b) {
        return (uint64_t)lj->band[b] << (((narrow >> b) & 1u) ? 1 : 2);
    };
    uint64_t lo = slot_byte(0);
#pragma unroll
    for (int b = 1; b < 4; ++b)
        if ((used >> b) & 1u) lo = min(lo, slot_byte(b));
    if (t < 4) {
        sb->relb[t] = ((used >> t) & 1u) ? (uint32_t)(slot_byte(t) - lo) : 0u;
        sb->pitchb[t] = lj->bw[t] << (((narrow >> t) & 1u) ? 1 : 2);
        sb->delta[t] = lj->delta[t];
    }
    uint64_t mis = 0;
#pragma unroll
    for (int b = 0; b < 4; ++b)
        if ((used >> b) & 1u)
            mis |= (reinterpret_cast<uintptr_t>(coef) + slot_byte(b)) |
                   ((uint64_t)lj->bw[b] << (((narrow >> b) & 1u) ? 1 : 2));
    LevelView v;
    v.aligned = (mis & 15u) == 0;
    v.bands = sb;
    v.origin = reinterpret_cast<const uint8_t *>(coef) + lo;
    v.x0 = lj->x0;
    v.y0 = lj->y0;
    v.fin = lj->float_in;
    v.narrow = narrow;
    v.w = (int)(lj->x1 - lj->x0);
    v.h = (int)(lj->y1 - lj->y0);
    v.levels = lj->levels;
    v.l53 = lj->lift53 != 0;
    __syncthreads();
    return v;
}

// Lifting over the window.  HORIZ: lines are window rows (all nh of them),
// elements run along x; otherwise lines are the 64 block columns and
// elements run along y.  Window pitch WN and the pair count are
// compile-time constants, so index arithmetic needs no runtime division.
template <bool REV, int WN, bool HORIZ, typename T, class G = Big>
__device__ __forceinline__ void lift(T *win, int nlines, int n, uint32_t origin, bool single,
                                     bool odd_single) {
    constexpr int H = REV ? 2 : 4;
    constexpr int PAIRS = WN / 2;
    constexpr int kBlk = G::kBlk, kIdwtThreads = G::kThreads;
    const int tid = threadIdx.x;
    auto at = [&](int line, int e) -> T & {
        return HORIZ ? win[line * WN + e] : win[e * WN + H + line];
    };
    if (single) {
        // length-1 signal: identity, or halved at an odd coordinate
        // (idwt.rs:726-731,942-957)
        if (!REV && odd_single)
            for (int k = tid; k < nlines * WN; k += kIdwtThreads) {
                const int l = HORIZ ? k / WN : k % kBlk, e = HORIZ ? k % WN : k / kBlk;
                if (l < nlines && e < n) at(l, e) = at(l, e) * (T)0.5f;
            }
        __syncthreads();
        return;
    }
    const int items = HORIZ ? nlines * PAIRS : PAIRS * kBlk;
    auto step = [&](uint32_t parity, auto f) {
        for (int k = tid; k < items; k += kIdwtThreads) {
            int l, e;
            if (HORIZ) {
                l = k / PAIRS;
                e = 2 * (k - l * PAIRS);
            } else {
                l = k % kBlk;
                e = 2 * (k / kBlk);
                if (l >= nlines) continue;
            }
            if (((origin + (uint32_t)e) & 1u) != parity) e += 1;
            if (e < 1 || e > n - 2) continue;
            T &c = at(l, e);
            c = f(c, at(l, e - 1), at(l, e + 1));
        }
        __syncthreads();
    };
    if constexpr (REV) {
        step(0u, [](int32_t s, int32_t a, int32_t b) { return s - ((a + b + 2) >> 2); });
        step(1u, [](int32_t s, int32_t a, int32_t b) { return s + ((a + b) >> 1); });
    } else {
        const float inv_k = 1.0f / kKappa;
        // scaling (steps 1-2 of F.3.8.2) on every element of the lines
        const int all = HORIZ ? nlines * WN : WN * kBlk;
        for (int k = tid; k < all; k += kIdwtThreads) {
            const int l = HORIZ ? k / WN : k % kBlk, e = HORIZ ? k % WN : k / kBlk;
            if (l >= nlines || e >= n) continue;
            T &c = at(l, e);
            c = c * (((origin + (uint32_t)e) & 1u) ? inv_k : kKappa);
        }
        __syncthreads();
        step(0u, [](float s, float a, float b) { return fmaf(a + b, kNegDelta, s); });
        step(1u, [](float s, float a, float b) { return fmaf(a + b, kNegGamma, s); });
        step(0u, [](float s, float a, float b) { return fmaf(a + b, kNegBeta, s); });
        step(1u, [](float s, float a, float b) { return fmaf(a + b, kNegAlpha, s); });
    }
}

// Window geometry: halo H (2 for 5/3, 4 for 9/7), N = 64 + 2H samples per
// side.  A window row holds the 80 positions ox - 8 .. ox + 72 of its block
// (ox: the block's first column), so that each sub-band contributes 40
// samples that start on a 4-sample boundary (fetch_chunks); lifting position
// 0 (ox - H) is row position X0 = 8 - H, and the lifting functions take the
// window from that column on (`wb` = window + X0).  Only positions
// [X0, X0 + N) are read; the row pitch P is the smallest odd number above
// X0 + N (a wave walking down 64 rows of one column, or along 64 columns of
// one row, touches 64 distinct LDS banks), and the positions a row's chunks
// write past P land on the next row's unread positions below X0.  kWords:
// the array, whose last row takes all 80 positions.
template <bool REV, class G = Big>
struct Win {
    static constexpr int H = REV ? 2 : 4;
    static constexpr int N = G::kBlk + 2 * H;
    static constexpr int X0 = 8 - H;
    static constexpr int P = (X0 + N) | 1;
    static constexpr int R = 8 + 2 * H;    // inputs of one 8-sample lifting segment
    static constexpr int kRow = G::kBlk + 16;  // row positions the chunk fetch writes
    static constexpr int kWords = (N - 1) * P + kRow;
    static_assert(P >= X0 + N && kRow - P <= X0, "row spill stays in unread positions");
};

// Both lifting passes of an 8-sample segment in registers: v[0..R) holds the
// signal at positions e0 - H .. e0 + 8 + H, v[0] at a position of parity Q
// (0: low-pass / even absolute coordinate).  Afterwards v[H .. H + 8) are
// the synthesised samples.  Each output depends only on inputs within H of
// it, so segment results equal whole-line lifting (idwt.rs:741-764 for 5/3,
// :767-835 for 9/7, same operation order).
template <bool REV, int Q, typename T>
__device__ __forceinline__ void lift_segment(T (&v)[Win<REV>::R], bool l53) {
    // keep every rounding of the line-wise lifting: no contraction of the
    // scaling products into the neighbour sums
#pragma clang fp contract(off)
    constexpr int R = Win<REV>::R;
    if constexpr (REV) {
#pragma unroll
        for (int i = 1; i < R - 1; ++i)
            if (((Q + i) & 1) == 0) v[i] = v[i] - ((v[i - 1] + v[i + 1] + 2) >> 2);
#pragma unroll
        for (int i = 2; i < R - 2; ++i)
            if (((Q + i) & 1) == 1) v[i] = v[i] + ((v[i - 1] + v[i + 1]) >> 1);
    } else if (l53) {
        // the 5/3 steps on f32 samples, as the reference computes them
        // (idwt.rs:741-764): floor((l + r) * 0.25 + 0.5), floor((l + r) * 0.5)
#pragma unroll
        for (int i = 1; i < R - 1; ++i)
            if (((Q + i) & 1) == 0) v[i] = v[i] - floorf(fmaf(v[i - 1] + v[i + 1], 0.25f, 0.5f));
#pragma unroll
        for (int i = 2; i < R - 2; ++i)
            if (((Q + i) & 1) == 1) v[i] = v[i] + floorf((v[i - 1] + v[i + 1]) * 0.5f);
    } else {
        const float inv_k = 1.0f / kKappa;
#pragma unroll
        for (int i = 0; i < R; ++i) v[i] = v[i] * (((Q + i) & 1) ? inv_k : kKappa);
#pragma unroll
        for (int i = 1; i < R - 1; ++i)
            if (((Q + i) & 1) == 0) v[i] = fmaf(v[i - 1] + v[i + 1], kNegDelta, v[i]);
#pragma unroll
        for (int i = 2; i < R - 2; ++i)
            if (((Q + i) & 1) == 1) v[i] = fmaf(v[i - 1] + v[i + 1], kNegGamma, v[i]);
#pragma unroll
        for (int i = 3; i < R - 3; ++i)
            if (((Q + i) & 1) == 0) v[i] = fmaf(v[i - 1] + v[i + 1], kNegBeta, v[i]);
#pragma unroll
        for (int i = 4; i < R - 4; ++i)
            if (((Q + i) & 1) == 1) v[i] = fmaf(v[i - 1] + v[i + 1], kNegAlpha, v[i]);
    }
}

// Destinations of the synthesised block (block coordinates x, y < 64).
// In-place sinks write the window itself; the others let the vertical pass
// write its results straight to their consumer.
template <bool REV, class G = Big>
struct SinkWin {
    typename Sample<REV>::T *win;
    static constexpr bool kInPlace = true;
    __device__ __forceinline__ void operator()(int x, int y, typename Sample<REV>::T v) const {
        win[(y + Win<REV, G>::H) * Win<REV, G>::P + x + Win<REV, G>::H] = v;
    }
};

template <bool REV>
struct SinkPlane {  // a 64x64 LDS plane of samples
    typename Sample<REV>::T *plane;
    static constexpr bool kInPlace = false;
    __device__ __forceinline__ void operator()(int x, int y, typename Sample<REV>::T v) const {
        plane[y * kBlk + x] = v;
    }
};

// A 64x64 LDS plane of int16 samples (5/3 only); `ovf` records a sample
// outside the int16 range (the plane then holds a wrong value).
struct SinkPlane16 {
    int16_t *plane;
    mutable bool ovf;
    static constexpr bool kInPlace = false;
    __device__ __forceinline__ void operator()(int x, int y, int32_t v) const {
        plane[y * kBlk + x] = (int16_t)v;
        ovf |= v != (int32_t)(int16_t)v;
    }
};

template <bool REV>
struct SinkGlobal {  // the level's output in the coefficient arena
    uint32_t *coef;
    uint32_t base, pitch;
    int bh;
    static constexpr bool kInPlace = false;
    __device__ __forceinline__ void operator()(int x, int y, typename Sample<REV>::T v) const {
        if (y >= bh) return;
        uint32_t bits;
        if constexpr (REV)
            bits = (uint32_t)v;
        else
            bits = __float_as_uint(v);
        coef[base + (uint32_t)y * pitch + (uint32_t)x] = bits;
    }
};

// One lifting direction over the window with 8-sample register segments.
// HORIZ: every window row in [0, nh), output columns [H, H + bw), written
// back into the window; else every block column in [0, bw), output rows
// [H, H + bh), handed to `sink`.  q: parity of window position 0 along the
// direction.  All segments are read before any in-place write.
template <bool REV, bool HORIZ, int Q, class Sink, class G = Big>
__device__ __forceinline__ void lift_window_q(typename Sample<REV>::T *win, int nh, int bw, int bh,
                                              const Sink &sink, bool l53) {
    using T = typename Sample<REV>::T;
    constexpr int kBlk = G::kBlk, kIdwtThreads = G::kThreads;
    constexpr int H = Win<REV, G>::H, N = Win<REV, G>::N, P = Win<REV, G>::P, R = Win<REV, G>::R;
    // items: HORIZ (row, segment) with the row fastest; else (column, segment)
    constexpr int kLines = HORIZ ? N : kBlk;
    constexpr int kItems = kLines * (kBlk / 8);
    constexpr int kPer = (kItems + kIdwtThreads - 1) / kIdwtThreads;
    constexpr bool kInPlace = HORIZ || Sink::kInPlace;
    const int nlines = HORIZ ? nh : bw;
    const int nseg = ((HORIZ ? bw : bh) + 7) >> 3;
    T out[kPer][8];
#pragma unroll
    for (int it = 0; it < kPer; ++it) {
        const int i = threadIdx.x + it * kIdwtThreads;
        const int line = i % kLines, seg = i / kLines;
        if (i >= kItems || line >= nlines || seg >= nseg) continue;
        T v[R];
#pragma unroll
        for (int j = 0; j < R; ++j)
            v[j] = HORIZ ? win[line * P + 8 * seg + j] : win[(8 * seg + j) * P + H + line];
        lift_segment<REV, Q>(v, l53);
#pragma unroll
        for (int j = 0; j < 8; ++j) out[it][j] = v[H + j];
    }
    if (kInPlace) __syncthreads();
#pragma unroll
    for (int it = 0; it < kPer; ++it) {
        const int i = threadIdx.x + it * kIdwtThreads;
        const int line = i % kLines, seg = i / kLines;
        if (i >= kItems || line >= nlines || seg >= nseg) continue;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            if (HORIZ)
                win[line * P + H + 8 * seg + j] = out[it][j];
            else
                sink(line, 8 * seg + j, out[it][j]);
        }
    }
    if (kInPlace) __syncthreads();
}

// The window parity q is uniform; dispatching on it outside the segment loop
// keeps each instance's lifting in fixed registers (a merged, select-based
// form would index the segment registers at run time).
template <bool REV, bool HORIZ, class Sink, class G = Big>
__device__ __forceinline__ void lift_window(typename Sample<REV>::T *win, int nh, int bw, int bh,
                                            uint32_t q, const Sink &sink, bool l53) {
    if (__builtin_amdgcn_readfirstlane(q))
        lift_window_q<REV, HORIZ, 1, Sink, G>(win, nh, bw, bh, sink, l53);
    else
        lift_window_q<REV, HORIZ, 0, Sink, G>(win, nh, bw, bh, sink, l53);
}

// Window-row table of block row `by`: for every window row after extension
// and both column parities xp, the byte offset of the row inside band slot
// xp | 2 * row parity, with the row parity in bit 0 (offsets are even).
// rows holds 2 * N words.  No barrier.
template <bool REV, class G = Big>
__device__ __forceinline__ void window_rows(const LevelView &L, int by, uint32_t *rows) {
    constexpr int kBlk = G::kBlk, kIdwtThreads = G::kThreads;
    constexpr int H = Win<REV, G>::H;
    const int oy = by * kBlk;
    const int nh = min(kBlk, L.h - oy) + 2 * H;
    for (int e = threadIdx.x; e < 2 * nh; e += kIdwtThreads) {
        const int yy = e >> 1, xp = e & 1;
        const int my = reflect(oy - H + yy, L.h);
        uint32_t row = (uint32_t)my, par = 0;
        if (L.levels) {
            const uint32_t ay = L.y0 + (uint32_t)my;
            par = ay & 1u;
            row = par ? (ay >> 1) - (L.y0 >> 1) : (ay >> 1) - ((L.y0 + 1) >> 1);
        }
        rows[e] = row * L.bands->pitchb[xp | (par << 1)] + par;
    }
}

// The window's raw coefficient words held in registers between the loads
// and their conversion into LDS, so that the loads of one component can be
// in flight while the previous component is lifted.  Each thread owns one
// window column and walks down it: pass 0 covers the block's columns with
// kGroups threads per column (each every kGroups-th row); pass 1 the 2H halo
// columns, threads / 2H per column.
template <bool REV, class G = Big>
struct WinFetch {
    static constexpr int kGroups = G::kThreads / G::kBlk;  // threads per window column
    static constexpr int kHalo = 2 * Win<REV, G>::H;
    static constexpr int kRows0 = (Win<REV, G>::N + kGroups - 1) / kGroups;
    static constexpr int kStep1 = G::kThreads / kHalo;
    static constexpr int kRows1 = (Win<REV, G>::N + kStep1 - 1) / kStep1;
    static_assert(kGroups - 1 + kGroups * (kRows0 - 1) < Win<REV, G>::N,
                  "pass-0 rows stay inside the window");
    uint32_t raw0[kRows0];
    uint32_t raw1[kRows1];
    uint32_t y1m0, y1m1;  // bit k: row k of the pass read the row-parity-1 slots
};

// Column `xx` of the window: its parity and, for row parity 0 and 1, the
// byte offset of the column in the slot (cb1 less one, which cancels the
// parity bit of the row entry) and the right shift that sign-extends an
// int16 sample (16) or keeps a 32-bit word (0).
struct WinColumn {
    uint32_t xp, cb0, cb1, d0, d1;
};

__device__ __forceinline__ WinColumn window_column(const LevelView &L, int ox, int xx, int H) {
    const int mx = reflect(ox - H + xx, L.w);
    uint32_t col = (uint32_t)mx, xpar = 0;
    if (L.levels) {
        const uint32_t ax = L.x0 + (uint32_t)mx;
        xpar = ax & 1u;
        col = xpar ? (ax >> 1) - (L.x0 >> 1) : (ax >> 1) - ((L.x0 + 1) >> 1);
    }
    const uint32_t n0 = (L.narrow >> xpar) & 1u, n1 = (L.narrow >> (xpar | 2)) & 1u;
    WinColumn c;
    c.xp = xpar;
    c.cb0 = L.bands->relb[xpar] + (col << (n0 ? 1 : 2));
    c.cb1 = L.bands->relb[xpar | 2] + (col << (n1 ? 1 : 2)) - 1u;
    c.d0 = n0 ? 16u : 0u;
    c.d1 = n1 ? 16u : 0u;
    return c;
}

// Sample word of row-table entry e in column c: an aligned 32-bit load,
// shifted so that an int16 sample comes out sign-extended and a 32-bit word
// unchanged.  ok == false loads the level's first word (a value that only
// reaches outputs outside the block).
__device__ __forceinline__ uint32_t window_word(const LevelView &L, const WinColumn &c, uint32_t e,
                                                bool ok) {
    const bool y1 = (e & 1u) != 0;
    const uint32_t off = ok ? (y1 ? c.cb1 : c.cb0) + e : 0u;
    const uint32_t w = *reinterpret_cast<const uint32_t *>(L.origin + (off & ~3u));
    const uint32_t down = y1 ? c.d1 : c.d0;
    const uint32_t up = down - ((off & 2u) << 3);
    return (uint32_t)((int32_t)(w << up) >> down);
}

// Issues every load of the window of block (bx, by); `rows` holds the
// block row's window-row table.  Positions outside the window (xx >= nw or
// yy >= nh) load the level's first word instead: every lane takes the same
// path, and those values only reach outputs outside the block.
template <bool REV, class G = Big>
__device__ __forceinline__ void fetch_window(const LevelView &L, int bx, int by,
                                             const uint32_t *rows, WinFetch<REV, G> &f) {
    constexpr int kBlk = G::kBlk;
    constexpr int H = Win<REV, G>::H, N = Win<REV, G>::N;
    using F = WinFetch<REV, G>;
    const int ox = bx * kBlk, oy = by * kBlk;
    const int nw = min(kBlk, L.w - ox) + 2 * H, nh = min(kBlk, L.h - oy) + 2 * H;
    const int tid = threadIdx.x;
    {
        const int xx = tid % kBlk, y0 = tid / kBlk;
        const bool colok = xx < nw;
        const WinColumn c = window_column(L, ox, colok ? xx : 0, H);
        uint32_t rr[F::kRows0];
#pragma unroll
        for (int k = 0; k < F::kRows0; ++k) rr[k] = rows[2 * (y0 + F::kGroups * k) + c.xp];
        uint32_t m = 0;
#pragma unroll
        for (int k = 0; k < F::kRows0; ++k) {
            f.raw0[k] = window_word(L, c, rr[k], colok && y0 + F::kGroups * k < nh);
            m |= (rr[k] & 1u) << k;
        }
        f.y1m0 = m;
    }
    {
        const int xx = kBlk + tid % F::kHalo, y0 = tid / F::kHalo;
        const bool colok = xx < nw;
        const WinColumn c = window_column(L, ox, colok ? xx : 0, H);
        uint32_t m = 0;
#pragma unroll
        for (int k = 0; k < F::kRows1; ++k) {
            const int yy = y0 + F::kStep1 * k;
            const uint32_t e = rows[2 * min(yy, N - 1) + c.xp];
            f.raw1[k] = window_word(L, c, e, colok && yy < nh);
            m |= (e & 1u) << k;
        }
        f.y1m1 = m;
    }
}

// Loads the window of block (bx, by) of a level with an even origin x0,
// width above H and 16-byte aligned slots, four samples of a sub-band row at
// a time (one 8-byte int16 or 16-byte int32 load), and writes it into `win`
// (the window from column X0 on).  With x0 even, row position x of the
// window is absolute column x0 + ox - 8 + x: even positions hold low-pass
// column ox/2 - 4 + x/2, odd ones high-pass column ox/2 - 4 + (x - 1)/2, so
// chunk k of either band (columns ox/2 - 4 + 4k .. + 3) fills positions
// 8k + 2j + parity, j < 4.  An item is one window row and one k, both
// bands.  Lanes 2m and 2m + 1 take chunks 2q and 2q + 1 of window row r0 + m:
// a half-wave covers 16 consecutive rows, reading 16 bytes (int16) or 32
// bytes (int32) of each per band, and its LDS writes, 8 positions apart
// along a row and P (odd) apart across rows, fall on distinct banks.  (Lanes
// walking the chunks of one row write 8 words apart, three to a bank; lanes
// walking the rows of one chunk read a separate 8 bytes per row.)  Window
// rows past the block's last (yy >= nh) repeat the last row of their parity.
// Positions left of the level (block column 0) and right of it (any block
// whose right halo passes the level's last column) are then mirrored from
// inside (whole-sample symmetric extension, F.3.7); chunks outside the band
// rows load the level's first word instead.
// No barrier at the end.
template <bool REV, class G = Big>
__device__ __forceinline__ void fetch_chunks(const LevelView &L, int bx, int by,
                                             const uint32_t *rows,
                                             typename Sample<REV>::T *win) {
    using T = typename Sample<REV>::T;
    constexpr int kBlk = G::kBlk, kIdwtThreads = G::kThreads;
    constexpr int H = Win<REV, G>::H, N = Win<REV, G>::N, P = Win<REV, G>::P, X0 = Win<REV, G>::X0;
    constexpr int kChunks = Win<REV, G>::kRow / 8;
    constexpr int kItems = N * kChunks;
    constexpr int kPer = (kItems + kIdwtThreads - 1) / kIdwtThreads;
    const int ox = bx * kBlk, oy = by * kBlk;
    const int bw = min(kBlk, L.w - ox), nh = min(kBlk, L.h - oy) + 2 * H;
    const int c0 = (ox >> 1) - 4;
    const int wl = (L.w + 1) >> 1, wh = L.w >> 1;  // low- and high-pass columns
    const int tid = threadIdx.x;
    uint32_t raw[kPer][2][4];
    int dst[kPer];
    uint32_t slot[kPer];  // bit b: band slot of the item's bands, for the conversion
#pragma unroll
    for (int it = 0; it < kPer; ++it) {
        const int i = tid + it * kIdwtThreads;
        dst[it] = -1;
        if (i >= kItems) continue;
        const int t = i >> 1;
        const int yy = t % N, k = 2 * (t / N) + (i & 1);
        // the last window row < nh of yy's parity
        const int yl = yy < nh ? yy : nh - 1 - ((nh - 1 - yy) & 1);
        const uint32_t e0 = rows[2 * yl], e1 = rows[2 * yl + 1];
        const uint32_t par = e0 & 1u;
        const int c = c0 + 4 * k;
        dst[it] = yy * P + 8 * k - X0;
        slot[it] = par << 1;
#pragma unroll
        for (int xp = 0; xp < 2; ++xp) {
            const uint32_t b = (uint32_t)xp | (par << 1);
            const bool nar = ((L.narrow >> b) & 1u) != 0;
            const bool ok = (unsigned)c < (unsigned)(xp ? wh : wl);
            const uint32_t off =
                ok ? L.bands->relb[b] + ((xp ? e1 : e0) - par) + ((uint32_t)c << (nar ? 1 : 2)) : 0u;
            const uint8_t *src = L.origin + off;
            if (nar) {
                const uint2 v = *reinterpret_cast<const uint2 *>(src);
                raw[it][xp][0] = v.x;
                raw[it][xp][1] = v.y;
            } else {
                const uint4 v = *reinterpret_cast<const uint4 *>(src);
                raw[it][xp][0] = v.x;
                raw[it][xp][1] = v.y;
                raw[it][xp][2] = v.z;
                raw[it][xp][3] = v.w;
            }
        }
    }
#pragma unroll
    for (int it = 0; it < kPer; ++it) {
        if (dst[it] < 0) continue;
        T *d = win + dst[it];
#pragma unroll
        for (int xp = 0; xp < 2; ++xp) {
            const uint32_t b = (uint32_t)xp | slot[it];
            const bool nar = ((L.narrow >> b) & 1u) != 0;
            int32_t s[4];
            if (nar) {
                s[0] = (int32_t)(raw[it][xp][0] << 16) >> 16;
                s[1] = (int32_t)raw[it][xp][0] >> 16;
                s[2] = (int32_t)(raw[it][xp][1] << 16) >> 16;
                s[3] = (int32_t)raw[it][xp][1] >> 16;
            } else {
#pragma unroll
                for (int j = 0; j < 4; ++j) s[j] = (int32_t)raw[it][xp][j];
            }
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                if constexpr (REV) {
                    d[2 * j + xp] = s[j];
                } else {
                    const bool isf = !nar && ((L.fin >> b) & 1u);
                    d[2 * j + xp] = isf ? __int_as_float(s[j]) : (float)s[j] * L.bands->delta[b];
                }
            }
        }
    }
    // the right halo (H columns past the block) crosses the level's end in
    // the last block column and also in the one before it when the level
    // ends fewer than H columns past that block (widths 64k + 1 .. 64k + H - 1)
    const bool left = ox == 0, right = L.w - ox - bw < H;
    if (left || right) {
        __syncthreads();
        const int xl = 8 + (L.w - 1 - ox);  // row position of the level's last column
        for (int i = tid; i < N * 2 * H; i += kIdwtThreads) {
            const int yy = i % N, s = i / N;
            const bool lft = s < H;
            const int m = lft ? s + 1 : s - H + 1;
            if (lft ? !left : !right) continue;
            const int to = lft ? 8 - m : xl + m, from = lft ? 8 + m : xl - m;
            win[yy * P + to - X0] = win[yy * P + from - X0];
        }
    }
}

// Converts the fetched words (dequantising with the band step for 9/7) and
// writes them into the window.  5/3 levels hold integers only; 9/7 slots may
// hold f32 samples (float_in).  No barrier.
template <bool REV, class G = Big>
__device__ __forceinline__ void commit_window(const LevelView &L, int bx, int by,
                                              const uint32_t *rows, const WinFetch<REV, G> &f,
                                              typename Sample<REV>::T *win) {
    using T = typename Sample<REV>::T;
    using F = WinFetch<REV, G>;
    constexpr int kBlk = G::kBlk, kIdwtThreads = G::kThreads;
    (void)kIdwtThreads;
    constexpr int H = Win<REV, G>::H, N = Win<REV, G>::N, P = Win<REV, G>::P;
    const int ox = bx * kBlk, oy = by * kBlk;
    const int nh = min(kBlk, L.h - oy) + 2 * H;
    const int tid = threadIdx.x;
    (void)rows;
    // per column: band-slot flags and steps for row parity 0 and 1 (9/7)
    struct Col {
        bool f0, f1;
        float d0, d1;
    };
    auto column = [&](int xx) {
        Col c{false, false, 1.0f, 1.0f};
        if constexpr (!REV) {
            uint32_t xpar = 0;
            if (L.levels) xpar = (L.x0 + (uint32_t)reflect(ox - H + xx, L.w)) & 1u;
            c.f0 = (L.fin >> xpar) & 1u;
            c.f1 = (L.fin >> (xpar | 2)) & 1u;
            c.d0 = L.bands->delta[xpar];
            c.d1 = L.bands->delta[xpar | 2];
        }
        return c;
    };
    auto convert = [&](const Col &c, uint32_t v, bool y1) -> T {
        if constexpr (REV) {
            return (int32_t)v;
        } else {
            const bool isf = y1 ? c.f1 : c.f0;
            return isf ? __uint_as_float(v) : (float)(int32_t)v * (y1 ? c.d1 : c.d0);
        }
    };
    {
        const int xx = tid % kBlk, y0 = tid / kBlk;
        const Col c = column(xx);
#pragma unroll
        for (int k = 0; k < F::kRows0; ++k)
            win[(y0 + F::kGroups * k) * P + xx] = convert(c, f.raw0[k], (f.y1m0 >> k) & 1u);
    }
    {
        const int xx = kBlk + tid % F::kHalo, y0 = tid / F::kHalo;
        const Col c = column(xx);
#pragma unroll
        for (int k = 0; k < F::kRows1; ++k) {
            const int yy = y0 + F::kStep1 * k;
            if (yy < nh && yy < N) win[yy * P + xx] = convert(c, f.raw1[k], (f.y1m1 >> k) & 1u);
        }
    }
}

// Lifts the committed window of block (bx, by) and hands the result to
// `sink`.  Starts with a barrier (the window writes of every thread must be
// visible); the sink's writes are complete for the calling thread only.
template <bool REV, class Sink, class G = Big>
__device__ __forceinline__ void lift_block(const LevelView &L, int bx, int by,
                                           typename Sample<REV>::T *win, const Sink &sink,
                                           int &bw_out, int &bh_out) {
    constexpr int kBlk = G::kBlk, kIdwtThreads = G::kThreads;
    constexpr int H = Win<REV, G>::H, P = Win<REV, G>::P;
    using T = typename Sample<REV>::T;
    const int ox = bx * kBlk, oy = by * kBlk;
    const int bw = min(kBlk, L.w - ox), bh = min(kBlk, L.h - oy);
    const int nw = bw + 2 * H, nh = bh + 2 * H;
    __syncthreads();
    bool in_window = true;  // result still in the window, not yet in the sink
    if (L.levels) {
        const uint32_t qx = (L.x0 + (uint32_t)(ox - H)) & 1u, qy = (L.y0 + (uint32_t)(oy - H)) & 1u;
        if (L.w == 1)
            lift<REV, P, true, T, G>(win, nh, nw, L.x0 + (uint32_t)(ox - H), true, (L.x0 & 1u) != 0);
        else
            lift_window<REV, true, Sink, G>(win, nh, bw, bh, qx, sink, L.l53);
        if (L.h == 1) {
            lift<REV, P, false, T, G>(win, bw, nh, L.y0 + (uint32_t)(oy - H), true, (L.y0 & 1u) != 0);
        } else {
            lift_window<REV, false, Sink, G>(win, nh, bw, bh, qy, sink, L.l53);
            in_window = false;
        }
    }
    if (in_window && !Sink::kInPlace) {
        for (int k = threadIdx.x; k < kBlk * bh; k += kIdwtThreads) {
            const int y = k / kBlk, x = k % kBlk;
            if (x < bw) sink(x, y, win[(y + H) * P + x + H]);
        }
    }
    bw_out = bw;
    bh_out = bh;
}

// Synthesises block (bx, by) of one level: window-row table, loads, window,
// lifting.  win: the window from column X0 on.  Result element (x, y) of an
// in-place sink's block is at win[(y + H) * P + x + H].  rows: LDS scratch
// of 2 N words.
template <bool REV, class Sink, class G = Big>
__device__ __forceinline__ void synth_block(const LevelView &L, const uint32_t *__restrict__ coef,
                                            int bx, int by, typename Sample<REV>::T *win,
                                            uint32_t *rows, const Sink &sink, int &bw_out,
                                            int &bh_out) {
    window_rows<REV, G>(L, by, rows);
    __syncthreads();
    if (L.levels && L.aligned && (L.x0 & 1u) == 0 && L.w > Win<REV, G>::H) {
        fetch_chunks<REV, G>(L, bx, by, rows, win);
    } else {
        WinFetch<REV, G> f;
        fetch_window<REV, G>(L, bx, by, rows, f);
        commit_window<REV, G>(L, bx, by, rows, f, win);
    }
    lift_block<REV, Sink, G>(L, bx, by, win, sink, bw_out, bh_out);
}

// Work-groups of a launch go to the 8 XCDs round-robin (work-group w runs on
// XCD w % 8, each XCD with its own L2).  The remap hands each XCD one
// contiguous range of logical blocks, so neighbouring 64x64 blocks of a tile,
// which share halo lines, are read through the same L2.
__device__ __forceinline__ uint32_t xcd_remap(uint32_t w, uint32_t n) {
    constexpr uint32_t kXcds = 8;
    const uint32_t x = w % kXcds, k = w / kXcds;
    const uint32_t m = n / kXcds, r = n % kXcds;
    return x * m + min(x, r) + k;
}

template <typename J>
__device__ __forceinline__ uint32_t find_job(const J *jobs, uint32_t njobs, uint32_t block) {
    uint32_t lo = 0, hi = njobs;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (jobs[mid].block_base <= block)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

// Job of work-group `blk`: from the block table when there is one, else by
// binary search over the jobs' block bases.
template <typename J>
__device__ __forceinline__ const J *job_of(const J *jobs, uint32_t njobs,
                                           const uint32_t *__restrict__ block_job, uint32_t blk) {
    return jobs + (block_job ? block_job[blk] : find_job(jobs, njobs, blk));
}

// One output block of one level.  G = Small serves launches whose levels
// are all at most 32x32 (one block each; the plan's block tables count
// 64x64 blocks, which are then the same single block).
template <bool REV, class G = Big>
__device__ __forceinline__ void level_body(const LevelJob *jobs, uint32_t njobs, uint32_t block_offset,
                                           const uint32_t *block_job, uint32_t *coef) {
    constexpr int kBlk = G::kBlk;
    constexpr int N = Win<REV, G>::N;
    __shared__ typename Sample<REV>::T win_lds[Win<REV, G>::kWords];
    typename Sample<REV>::T *const win = win_lds + Win<REV, G>::X0;
    __shared__ uint32_t rows[2 * N];
    const uint32_t blk = xcd_remap(blockIdx.x, gridDim.x) + block_offset;
    const LevelJob *lj = job_of(jobs, njobs, block_job, blk);
    __shared__ BandLds sb;
    const LevelView L = level_view(lj, coef, &sb);
    const uint32_t local = blk - lj->block_base;
    const uint32_t nbx = lj->nbx, out = lj->out;
    const int bx = (int)(lj->bx0 + local % nbx), by = (int)(lj->by0 + local / nbx);
    int bw, bh;
    const uint32_t opitch = coef_pitch((uint32_t)L.w);
    const int bh_blk = min(kBlk, L.h - by * kBlk);
    const SinkGlobal<REV> sink{coef, out + (uint32_t)(by * kBlk) * opitch + (uint32_t)(bx * kBlk),
                               opitch, bh_blk};
    synth_block<REV, SinkGlobal<REV>, G>(L, coef, bx, by, win, rows, sink, bw, bh);
}

extern "C" __global__ void __launch_bounds__(kIdwtThreads)
sgn_idwt_level_rev(const LevelJob *jobs, uint32_t njobs, uint32_t block_offset,
                   const uint32_t *block_job, uint32_t *coef) {
    level_body<true>(jobs, njobs, block_offset, block_job, coef);
}

extern "C" __global__ void __launch_bounds__(kIdwtThreads)
sgn_idwt_level_irr(const LevelJob *jobs, uint32_t njobs, uint32_t block_offset,
                   const uint32_t *block_job, uint32_t *coef) {
    level_body<false>(jobs, njobs, block_offset, block_job, coef);
}

// The same for launches whose levels are all at most 32x32: one wave per
// level, a window of a fifth of the size, so several times more
// work-groups fit a CU (these launches are bound by each work-group's
// chain of memory round trips, not by bandwidth).
extern "C" __global__ void __launch_bounds__(Small::kThreads)
sgn_idwt_level_rev_small(const LevelJob *jobs, uint32_t njobs, uint32_t block_offset,
                         const uint32_t *block_job, uint32_t *coef) {
    level_body<true, Small>(jobs, njobs, block_offset, block_job, coef);
}

extern "C" __global__ void __launch_bounds__(Small::kThreads)
sgn_idwt_level_irr_small(const LevelJob *jobs, uint32_t njobs, uint32_t block_offset,
                         const uint32_t *block_job, uint32_t *coef) {
    level_body<false, Small>(jobs, njobs, block_offset, block_job, coef);
}

// 64x64 blocks with two waves (Mid): the same blocks, each work-group's
// window walked by half as many threads.
extern "C" __global__ void __launch_bounds__(Mid::kThreads)
sgn_idwt_level_rev_w2(const LevelJob *jobs, uint32_t njobs, uint32_t block_offset,
                      const uint32_t *block_job, uint32_t *coef) {
    level_body<true, Mid>(jobs, njobs, block_offset, block_job, coef);
}

extern "C" __global__ void __launch_bounds__(Mid::kThreads)
sgn_idwt_level_irr_w2(const LevelJob *jobs, uint32_t njobs, uint32_t block_offset,
                      const uint32_t *block_job, uint32_t *coef) {
    level_body<false, Mid>(jobs, njobs, block_offset, block_job, coef);
}

// ---- final stage -------------------------------------------------------

__device__ __forceinline__ uint8_t sat_u8(float v) {
    // Rust `as u8`: NaN -> 0, saturating
    if (!(v == v) || v <= 0.0f) return 0;
    if (v >= 255.0f) return 255;
    return (uint8_t)v;
}

struct FinalView {
    uint8_t *dst;
    uint64_t pitch;
    int width, height, nc, fmt, ch, bps, mct;
    uint32_t prec0, prec1, prec2, same8;
    bool rct_f32;  // f32 planes under the reversible colour transform
};

// Converts one pixel's component values (after MCT and sign shift) into
// output sample k of component c.
__device__ __forceinline__ uint32_t out_u8(float v, bool same8, uint32_t prec) {
    if (same8) return sat_u8(roundf(v));
    const float mx = (float)((1u << prec) - 1);
    return sat_u8(roundf((v / mx) * 255.0f));
}

__device__ __forceinline__ uint32_t out_u16(float v, uint32_t bd) {
    const float mx = (float)((1ull << bd) - 1);
    const float r = roundf(v);
    const float cl = r < 0.0f ? 0.0f : (r > mx ? mx : r);
    if (bd <= 8) {
        uint32_t m8 = (1u << bd) - 1;
        if (m8 == 0) m8 = 1;
        return ((uint32_t)cl * 65535u + m8 / 2) / m8;
    }
    return (uint32_t)(cl > 65535.0f ? 65535.0f : cl);
}

// Final stage of block `blk`.  PT: sample type of the two LDS planes that
// hold all but the last component (the last stays in the window).  With
// int16 planes (5/3) the work-group's LDS drops from 52 to 36 KB (four
// work-groups per CU instead of three); a block with a sample outside the
// int16 range writes no output and appends itself to `redo` (count in
// redo_n), which the int32-plane pass then decodes.
template <bool REV, typename PT>
__device__ __forceinline__ void final_body(const FinalJob *jobs, uint32_t njobs, uint32_t blk,
                                           const uint32_t *block_job, const LevelJob *lv,
                                           const uint32_t *coef, uint32_t *redo, uint32_t *redo_n) {
    constexpr int H = Win<REV>::H, N = Win<REV>::N, P = Win<REV>::P;
    constexpr bool kNarrowPlanes = sizeof(PT) == 2;
    using T = typename Sample<REV>::T;
    __shared__ T win_lds[Win<REV>::kWords];
    T *const win = win_lds + Win<REV>::X0;  // lifting position 0
    __shared__ uint32_t rows[2 * N];  // window-row table
    __shared__ __attribute__((aligned(16))) PT res[2][kBlk * kBlk];
    const FinalJob *fj = job_of(jobs, njobs, block_job, blk);
    const uint32_t local = blk - fj->block_base;
    const uint32_t nbx = fj->nbx;
    const int bx = (int)(fj->bx0 + local % nbx), by = (int)(fj->by0 + local / nbx);
    // pixels kept, relative to the block: a region decode writes only
    // [cx0, cx1) x [cy0, cy1) of the tile, pixel by pixel
    const int kx0 = (int)fj->cx0 - bx * kBlk, kx1 = (int)fj->cx1 - bx * kBlk;
    const int ky0 = (int)fj->cy0 - by * kBlk, ky1 = (int)fj->cy1 - by * kBlk;
    const bool cropped = kx0 > 0 || ky0 > 0 || kx1 < kBlk || ky1 < kBlk;
    FinalView F;
    F.nc = fj->ncomp > 3 ? 3 : fj->ncomp;
    F.fmt = fj->fmt;
    F.ch = (F.fmt == 0 || F.fmt == 3) ? 1 : (F.fmt == 2 ? 4 : 3);
    F.bps = (F.fmt == 3 || F.fmt == 4) ? 2 : 1;
    F.mct = fj->mct;
    F.dst = reinterpret_cast<uint8_t *>(fj->dst);
    F.pitch = fj->pitch;
    F.prec0 = fj->precision[0];
    F.prec1 = fj->precision[1];
    F.prec2 = fj->precision[2];
    F.same8 = fj->same8;
    F.rct_f32 = fj->rct_f32 != 0;
    const uint32_t lvl0 = fj->level[0], lvl1 = fj->level[1], lvl2 = fj->level[2];
    int bw = 0, bh = 0;
    __shared__ BandLds sb;
    bool ovf = false;
    for (int c = 0; c < F.nc; ++c) {
        // level_view ends with a barrier: the previous component's vertical
        // pass, the last reader of the window, is over before it is rewritten
        const LevelView L = level_view(lv + (c == 0 ? lvl0 : c == 1 ? lvl1 : lvl2), coef, &sb);
        if (c + 1 < F.nc) {
            if constexpr (kNarrowPlanes) {
                const SinkPlane16 sink{reinterpret_cast<int16_t *>(res[c]), false};
                synth_block<REV>(L, coef, bx, by, win, rows, sink, bw, bh);
                ovf |= sink.ovf;
            } else {
                synth_block<REV>(L, coef, bx, by, win, rows,
                                 SinkPlane<REV>{reinterpret_cast<T *>(res[c])}, bw, bh);
            }
        } else {
            synth_block<REV>(L, coef, bx, by, win, rows, SinkWin<REV>{win}, bw, bh);
        }
    }
    if constexpr (kNarrowPlanes) {
        const bool any = __syncthreads_or(ovf) != 0;
        if (any) {  // the int32 pass decodes this block
            if (threadIdx.x == 0) redo[atomicAdd(redo_n, 1u)] = blk;
            return;
        }
    } else {
        __syncthreads();
    }
    // Each thread converts and stores four horizontally adjacent pixels.
    const bool words =
        !cropped && F.bps == 1 && ((reinterpret_cast<uintptr_t>(F.dst) | F.pitch) & 3) == 0;
    const bool int8_out = REV && F.bps == 1 && F.same8;
    // 5/3 with RCT into 8-bit RGB (the common case), integers throughout:
    // R = Cr + I, G = I, B = Cb + I with I = Y - ((Cb + Cr) >> 2), each
    // shifted by 2^(P-1) and clamped (mct.rs:71-156, decode.rs:1720-1752,
    // lib.rs:2202-2295), three words per four pixels.
    const bool rct_rgb8 = REV && int8_out && words && F.nc == 3 && F.mct && F.ch == 3;
    if (rct_rgb8) {
        const int32_t a0 = 1 << (F.prec0 - 1), a1 = 1 << (F.prec1 - 1), a2 = 1 << (F.prec2 - 1);
        for (int k = threadIdx.x; k < (kBlk / 4) * bh; k += kIdwtThreads) {
            const int y = k / (kBlk / 4), x4 = (k % (kBlk / 4)) * 4;
            if (x4 + 4 > bw) continue;  // partial groups: the general loop below
            int32_t c0[4], c1[4], c2[4];
            if constexpr (kNarrowPlanes) {
                const uint2 q0 = *reinterpret_cast<const uint2 *>(&res[0][y * kBlk + x4]);
                const uint2 q1 = *reinterpret_cast<const uint2 *>(&res[1][y * kBlk + x4]);
                c0[0] = (int32_t)(q0.x << 16) >> 16;
                c0[1] = (int32_t)q0.x >> 16;
                c0[2] = (int32_t)(q0.y << 16) >> 16;
                c0[3] = (int32_t)q0.y >> 16;
                c1[0] = (int32_t)(q1.x << 16) >> 16;
                c1[1] = (int32_t)q1.x >> 16;
                c1[2] = (int32_t)(q1.y << 16) >> 16;
                c1[3] = (int32_t)q1.y >> 16;
            } else {
                const int4 q0 = *reinterpret_cast<const int4 *>(&res[0][y * kBlk + x4]);
                const int4 q1 = *reinterpret_cast<const int4 *>(&res[1][y * kBlk + x4]);
                c0[0] = q0.x, c0[1] = q0.y, c0[2] = q0.z, c0[3] = q0.w;
                c1[0] = q1.x, c1[1] = q1.y, c1[2] = q1.z, c1[3] = q1.w;
            }
#pragma unroll
            for (int q = 0; q < 4; ++q) c2[q] = (int32_t)win[(y + H) * P + x4 + q + H];
            uint32_t r[4], g[4], b[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int32_t i1 = c0[q] - ((c2[q] + c1[q]) >> 2);
                r[q] = (uint32_t)min(max(c2[q] + i1 + a0, 0), 255);
                g[q] = (uint32_t)min(max(i1 + a1, 0), 255);
                b[q] = (uint32_t)min(max(c1[q] + i1 + a2, 0), 255);
            }
            uint32_t *pw = reinterpret_cast<uint32_t *>(F.dst + (size_t)(by * kBlk + y) * F.pitch +
                                                        (size_t)(bx * kBlk + x4) * 3);
            pw[0] = r[0] | (g[0] << 8) | (b[0] << 16) | (r[1] << 24);
            pw[1] = g[1] | (b[1] << 8) | (r[2] << 16) | (g[2] << 24);
            pw[2] = b[2] | (r[3] << 8) | (g[3] << 16) | (b[3] << 24);
        }
    }
    for (int k = threadIdx.x; k < (kBlk / 4) * bh; k += kIdwtThreads) {
        const int y = k / (kBlk / 4), x4 = (k % (kBlk / 4)) * 4;
        if (x4 >= bw || (rct_rgb8 && x4 + 4 <= bw)) continue;
        if (cropped && (y < ky0 || y >= ky1 || x4 + 4 <= kx0 || x4 >= kx1)) continue;
        const int np = min(4, bw - x4);
        uint32_t o[4][4];
        // the four pixels of each LDS plane in one 16-byte (8-byte) read
        T pl[2][4];
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            if (c + 1 < F.nc) {
                if constexpr (kNarrowPlanes) {
                    const uint2 q2 = *reinterpret_cast<const uint2 *>(&res[c][y * kBlk + x4]);
                    pl[c][0] = (int32_t)(int16_t)(q2.x & 0xFFFFu);
                    pl[c][1] = (int32_t)q2.x >> 16;
                    pl[c][2] = (int32_t)(int16_t)(q2.y & 0xFFFFu);
                    pl[c][3] = (int32_t)q2.y >> 16;
                } else {
                    const int4 q4 = *reinterpret_cast<const int4 *>(&res[c][y * kBlk + x4]);
                    pl[c][0] = __builtin_bit_cast(T, q4.x);
                    pl[c][1] = __builtin_bit_cast(T, q4.y);
                    pl[c][2] = __builtin_bit_cast(T, q4.z);
                    pl[c][3] = __builtin_bit_cast(T, q4.w);
                }
            } else {
                pl[c][0] = pl[c][1] = pl[c][2] = pl[c][3] = T(0);
            }
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int x = x4 + q;
            float v[3];
            if constexpr (REV) {
                int32_t s[3] = {0, 0, 0};
#pragma unroll
                for (int c = 0; c < 3; ++c)
                    if (c < F.nc) s[c] = c + 1 < F.nc ? pl[c < 2 ? c : 1][q] : win[(y + H) * P + x + H];
                if (F.mct) {
                    const int32_t i1 = s[0] - ((s[2] + s[1]) >> 2);
                    const int32_t r = s[2] + i1, b = s[1] + i1;
                    s[0] = r;
                    s[1] = i1;
                    s[2] = b;
                }
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    const uint32_t pc = c == 0 ? F.prec0 : c == 1 ? F.prec1 : F.prec2;
                    s[c] += 1 << (pc - 1);
                    v[c] = (float)s[c];
                }
                if (int8_out) {
                    // integers: rounding is the identity, `as u8` a clamp
#pragma unroll
                    for (int c = 0; c < 3; ++c) o[q][c] = (uint32_t)min(max(s[c], 0), 255);
                    o[q][3] = 255;
                    continue;
                }
            } else {
#pragma unroll
                for (int c = 0; c < 3; ++c)
                    v[c] = c < F.nc ? (c + 1 < F.nc ? pl[c < 2 ? c : 1][q] : win[(y + H) * P + x + H]) : 0.0f;
                if (F.mct && F.rct_f32) {
                    // the reversible transform on f32 samples (mct.rs:71-156)
                    const float y0 = v[0], y1 = v[1], y2 = v[2];
                    const float i1 = y0 - floorf((y2 + y1) * 0.25f);
                    v[0] = y2 + i1;
                    v[1] = i1;
                    v[2] = y1 + i1;
                } else if (F.mct) {
                    const float y0 = v[0], y1 = v[1], y2 = v[2];
                    v[0] = fmaf(y2, 1.402f, y0);
                    v[1] = fmaf(y2, -0.71414f, fmaf(y1, -0.34413f, y0));
                    v[2] = fmaf(y1, 1.772f, y0);
                }
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    const uint32_t pc = c == 0 ? F.prec0 : c == 1 ? F.prec1 : F.prec2;
                    v[c] = v[c] + (float)(1u << (pc - 1));
                }
            }
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const uint32_t pc = c == 0 ? F.prec0 : c == 1 ? F.prec1 : F.prec2;
                o[q][c] = F.bps == 1 ? out_u8(v[c], F.same8, pc) : out_u16(v[c], F.prec0);
            }
            o[q][3] = 255;
        }
        const int gx = bx * kBlk + x4, gy = by * kBlk + y;
        uint8_t *px = F.dst + (size_t)gy * F.pitch + (size_t)gx * F.ch * F.bps;
        if (words && np == 4) {
            // 4 pixels x ch bytes = ch 32-bit words
            uint32_t *pw = reinterpret_cast<uint32_t *>(px);
            if (F.ch == 1) {
                pw[0] = o[0][0] | (o[1][0] << 8) | (o[2][0] << 16) | (o[3][0] << 24);
            } else if (F.ch == 3) {
                pw[0] = o[0][0] | (o[0][1] << 8) | (o[0][2] << 16) | (o[1][0] << 24);
                pw[1] = o[1][1] | (o[1][2] << 8) | (o[2][0] << 16) | (o[2][1] << 24);
                pw[2] = o[2][2] | (o[3][0] << 8) | (o[3][1] << 16) | (o[3][2] << 24);
            } else {
                for (int q = 0; q < 4; ++q)
                    pw[q] = o[q][0] | (o[q][1] << 8) | (o[q][2] << 16) | (o[q][3] << 24);
            }
        } else {
#pragma unroll
            for (int q = 0; q < 4; ++q)
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    if (q >= np || c >= F.ch) continue;
                    if (cropped && (x4 + q < kx0 || x4 + q >= kx1)) continue;
                    if (F.bps == 1) {
                        px[q * F.ch + c] = (uint8_t)o[q][c];
                    } else {
                        px[2 * (q * F.ch + c)] = (uint8_t)(o[q][c] & 0xFF);
                        px[2 * (q * F.ch + c) + 1] = (uint8_t)(o[q][c] >> 8);
                    }
                }
        }
    }
}

extern "C" __global__ void __launch_bounds__(kIdwtThreads)
sgn_final_rev(const FinalJob *jobs, uint32_t njobs, uint32_t block_offset,
              const uint32_t *block_job, const LevelJob *lv, const uint32_t *coef) {
    final_body<true, int32_t>(jobs, njobs, xcd_remap(blockIdx.x, gridDim.x) + block_offset,
                              block_job, lv, coef, nullptr, nullptr);
}

// 5/3 final stage with int16 component planes.  *redo_n must be zero at
// launch; blocks with a sample outside the int16 range are listed in `redo`
// (capacity: the launch's block count) for sgn_final_rev_redo.
extern "C" __global__ void __launch_bounds__(kIdwtThreads)
sgn_final_rev16(const FinalJob *jobs, uint32_t njobs, uint32_t block_offset,
                const uint32_t *block_job, const LevelJob *lv, const uint32_t *coef,
                uint32_t *redo, uint32_t *redo_n) {
    final_body<true, int16_t>(jobs, njobs, xcd_remap(blockIdx.x, gridDim.x) + block_offset,
                              block_job, lv, coef, redo, redo_n);
}

// The blocks sgn_final_rev16 listed, with int32 planes; a fixed grid walks
// the list (usually empty) and every work-group exits when it ends.
extern "C" __global__ void __launch_bounds__(kIdwtThreads)
sgn_final_rev_redo(const FinalJob *jobs, uint32_t njobs, const uint32_t *block_job,
                   const LevelJob *lv, const uint32_t *coef, const uint32_t *redo,
                   const uint32_t *redo_n) {
    const uint32_t n = *redo_n;
    for (uint32_t k = blockIdx.x; k < n; k += gridDim.x)
        final_body<true, int32_t>(jobs, njobs, redo[k], block_job, lv, coef, nullptr, nullptr);
}

extern "C" __global__ void __launch_bounds__(kIdwtThreads)
sgn_final_irr(const FinalJob *jobs, uint32_t njobs, uint32_t block_offset,
              const uint32_t *block_job, const LevelJob *lv, const uint32_t *coef) {
    final_body<false, float>(jobs, njobs, xcd_remap(blockIdx.x, gridDim.x) + block_offset,
                             block_job, lv, coef, nullptr, nullptr);
}

// Byte fill used instead of memset nodes (safe under graph capture).
// Output of tile-components on different sampling grids (StoreJob): one
// work-group per 64x64 block of a job's output rectangle; each pixel takes
// the component sample its grid position falls on (decode.rs:1858-1898),
// shifted, converted like the fused final stage and written as its channel
// (with alpha from channel 0 in RGBA8).
extern "C" __global__ void __launch_bounds__(kIdwtThreads)
sgn_store_sampled(const StoreJob *jobs, uint32_t njobs, uint32_t block_offset,
                  const uint32_t *block_job, const uint32_t *coef) {
    const uint32_t blk = blockIdx.x + block_offset;
    const StoreJob *sj = job_of(jobs, njobs, block_job, blk);
    const uint32_t local = blk - sj->block_base;
    const uint32_t x0 = sj->ox0 + (local % sj->nbx) * kBlk, y0 = sj->oy0 + (local / sj->nbx) * kBlk;
    const uint32_t x1 = min(x0 + (uint32_t)kBlk, sj->ox1), y1 = min(y0 + (uint32_t)kBlk, sj->oy1);
    const bool rev = (sj->flags & 1u) != 0, same8 = (sj->flags & 2u) != 0;
    const bool alpha = (sj->flags & 4u) != 0;
    const float shift = (float)(1u << (sj->precision - 1));
    uint8_t *const out = reinterpret_cast<uint8_t *>(sj->dst);
    for (uint32_t k = threadIdx.x; k < (uint32_t)(kBlk * kBlk); k += kIdwtThreads) {
        const uint32_t X = x0 + k % kBlk, Y = y0 + k / kBlk;
        if (X >= x1 || Y >= y1) continue;
        const uint32_t u = (X + sj->xo) / sj->xr - sj->cx0, v = (Y + sj->yo) / sj->yr - sj->cy0;
        const uint32_t w = coef[sj->plane + v * sj->pitch + u];
        const float f = (rev ? (float)(int32_t)w : __uint_as_float(w)) + shift;
        uint8_t *px = out + (size_t)Y * sj->opitch + (size_t)X * sj->nchan * sj->bps;
        if (sj->bps == 1) {
            px[sj->chan] = (uint8_t)out_u8(f, same8, sj->precision);
            if (alpha) px[3] = 255;
        } else {
            const uint32_t o = out_u16(f, sj->prec0);
            px[2 * sj->chan] = (uint8_t)(o & 0xFFu);
            px[2 * sj->chan + 1] = (uint8_t)(o >> 8);
        }
    }
}

extern "C" __global__ void sgn_fill_bytes(uint8_t *p, size_t bytes, size_t words, uint32_t v4) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    uint32_t *w = reinterpret_cast<uint32_t *>(p);
    const bool aligned = (reinterpret_cast<uintptr_t>(p) & 3) == 0;
    if (aligned) {
        for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < words; i += stride)
            w[i] = v4;
        for (size_t i = words * 4 + (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < bytes;
             i += stride)
            p[i] = (uint8_t)v4;
    } else {
        for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < bytes; i += stride)
            p[i] = (uint8_t)v4;
    }
}

// Device-to-device byte copy: 16-byte words when both ends are aligned.
extern "C" __global__ void sgn_copy_bytes(uint8_t *__restrict__ dst, const uint8_t *__restrict__ src,
                                          size_t bytes) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t head = 0;
    if (((reinterpret_cast<uintptr_t>(dst) | reinterpret_cast<uintptr_t>(src)) & 15) == 0) {
        const size_t n16 = bytes / 16;
        const uint4 *s4 = reinterpret_cast<const uint4 *>(src);
        uint4 *d4 = reinterpret_cast<uint4 *>(dst);
        for (size_t i = t; i < n16; i += stride) d4[i] = s4[i];
        head = n16 * 16;
    }
    for (size_t i = head + t; i < bytes; i += stride) dst[i] = src[i];
}

// Inverse MCT on f32 planes (J2kInverseMctJob hook, lib.rs:785-800).
extern "C" __global__ void sgn_inverse_mct_planes(float *p0, float *p1, float *p2, uint32_t n,
                                                  int reversible, float a0, float a1, float a2) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float y0 = p0[i], y1 = p1[i], y2 = p2[i];
    float r, g, b;
    if (reversible) {
        const float i1 = y0 - floorf((y2 + y1) * 0.25f);
        r = y2 + i1;
        g = i1;
        b = y1 + i1;
    } else {
        r = fmaf(y2, 1.402f, y0);
        g = fmaf(y2, -0.71414f, fmaf(y1, -0.34413f, y0));
        b = fmaf(y1, 1.772f, y0);
    }
    p0[i] = r + a0;
    p1[i] = g + a1;
    p2[i] = b + a2;
}

// Component store (J2kStoreComponentJob hook, lib.rs:805-828; the copy of
// decode.rs:1792-1810): out[(oy + y) * ow + ox + x] =
// in[(sy + y) * iw + sx + x] + addend over a copy_w x copy_h rectangle.
// One thread per sample, rows along x so loads and stores coalesce.
extern "C" __global__ void sgn_store_plane(const float *in, uint32_t iw, uint32_t sx, uint32_t sy,
                                           uint32_t copy_w, uint32_t copy_h, float *out,
                                           uint32_t ow, uint32_t ox, uint32_t oy, float addend) {
    const uint32_t x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= copy_w) return;
    for (uint32_t y = blockIdx.y; y < copy_h; y += gridDim.y)
        out[size_t(oy + y) * ow + ox + x] = in[size_t(sy + y) * iw + sx + x] + addend;
}

}  // namespace sgn
