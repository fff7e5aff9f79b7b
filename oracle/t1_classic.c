/* Classic (EBCOT) code-block decode: the MQ arithmetic decoder (Annex C)
 * and the significance-propagation, magnitude-refinement and cleanup passes
 * of Annex D, with the code-block styles selective bypass, context reset,
 * termination on each pass, vertically causal context and segmentation
 * symbols.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  Restates
 *   crates/signinum-j2k-native/src/j2c/bitplane.rs:90-317 (segments and pass
 *     loop), :903-1116 (the three passes, run-length mode), :1228-1301 (sign
 *     context), :1303-1330 (zero-coding and refinement contexts), :1339-1370
 *     (raw bypass reads with bit stuffing, padded with 1 bits);
 *   crates/signinum-j2k-native/src/j2c/arithmetic_decoder.rs:42-250 (INITDEC,
 *     BYTEIN, DECODE; bytes past the segment read as 0xFF), :329-386 (Qe).
 * The zero-coding and sign contexts are computed here from neighbour counts
 * and signs as Tables D.1 and D.3 state them, not from the reference's
 * packed lookup tables.
 */
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

/* Table C.2: Qe, index after an MPS, index after an LPS, MPS switch. */
static const struct {
    uint16_t qe;
    uint8_t nmps, nlps, sw;
} QE[47] = {
    {0x5601, 1, 1, 1},   {0x3401, 2, 6, 0},   {0x1801, 3, 9, 0},   {0x0AC1, 4, 12, 0},
    {0x0521, 5, 29, 0},  {0x0221, 38, 33, 0}, {0x5601, 7, 6, 1},   {0x5401, 8, 14, 0},
    {0x4801, 9, 14, 0},  {0x3801, 10, 14, 0}, {0x3001, 11, 17, 0}, {0x2401, 12, 18, 0},
    {0x1C01, 13, 20, 0}, {0x1601, 29, 21, 0}, {0x5601, 15, 14, 1}, {0x5401, 16, 14, 0},
    {0x5101, 17, 15, 0}, {0x4801, 18, 16, 0}, {0x3801, 19, 17, 0}, {0x3401, 20, 18, 0},
    {0x3001, 21, 19, 0}, {0x2801, 22, 19, 0}, {0x2401, 23, 20, 0}, {0x2201, 24, 21, 0},
    {0x1C01, 25, 22, 0}, {0x1801, 26, 23, 0}, {0x1601, 27, 24, 0}, {0x1401, 28, 25, 0},
    {0x1201, 29, 26, 0}, {0x1101, 30, 27, 0}, {0x0AC1, 31, 28, 0}, {0x09C1, 32, 29, 0},
    {0x08A1, 33, 30, 0}, {0x0521, 34, 31, 0}, {0x0441, 35, 32, 0}, {0x02A1, 36, 33, 0},
    {0x0221, 37, 34, 0}, {0x0141, 38, 35, 0}, {0x0111, 39, 36, 0}, {0x0085, 40, 37, 0},
    {0x0049, 41, 38, 0}, {0x0025, 42, 39, 0}, {0x0015, 43, 40, 0}, {0x0009, 44, 41, 0},
    {0x0005, 45, 42, 0}, {0x0001, 45, 43, 0}, {0x5601, 46, 46, 0},
};

enum { CX_RL = 17, CX_UNI = 18, NCX = 19 };

/* ---- MQ decoder: register C kept inverted, as in the software
 * conventions of ITU-T T.88 Annex G ---------------------------------- */
typedef struct {
    const uint8_t *d;
    size_t n, bp;
    uint32_t c, a, ct;
} mq_t;

static uint32_t mq_at(const mq_t *m, size_t i) { return i < m->n ? m->d[i] : 0xFFu; }

static void mq_bytein(mq_t *m)
{
    if (mq_at(m, m->bp) == 0xFF) {
        if (mq_at(m, m->bp + 1) > 0x8F) {
            m->ct = 8; /* a marker (or the end): feed 1 bits, stay put */
        } else {
            m->bp++;
            m->c += 0xFE00u - (mq_at(m, m->bp) << 9);
            m->ct = 7;
        }
    } else {
        m->bp++;
        m->c += 0xFF00u - (mq_at(m, m->bp) << 8);
        m->ct = 8;
    }
}

static void mq_init(mq_t *m, const uint8_t *d, size_t n)
{
    m->d = d;
    m->n = n;
    m->bp = 0;
    m->c = (mq_at(m, 0) ^ 0xFFu) << 16;
    mq_bytein(m);
    m->c <<= 7;
    m->ct -= 7;
    m->a = 0x8000;
}

/* cx: bit 7 the MPS, bits 0-6 the Qe index */
static uint32_t mq_decode(mq_t *m, uint8_t *cx)
{
    uint32_t i = *cx & 0x7Fu, mps = *cx >> 7, d;
    const uint32_t qe = QE[i].qe;
    m->a -= qe;
    if ((m->c >> 16) < m->a) {
        if (m->a & 0x8000u) return mps;
        if (m->a < qe) { /* conditional exchange: the LPS took the larger part */
            d = 1 - mps;
            if (QE[i].sw) mps ^= 1;
            i = QE[i].nlps;
        } else {
            d = mps;
            i = QE[i].nmps;
        }
    } else {
        m->c -= m->a << 16;
        if (m->a < qe) {
            d = mps;
            i = QE[i].nmps;
        } else {
            d = 1 - mps;
            if (QE[i].sw) mps ^= 1;
            i = QE[i].nlps;
        }
        m->a = qe;
    }
    *cx = (uint8_t)(mps << 7 | i);
    do {
        if (m->ct == 0) mq_bytein(m);
        m->a <<= 1;
        m->c <<= 1;
        m->ct--;
    } while (!(m->a & 0x8000u));
    return d;
}

/* ---- raw (bypass) bits: B.10.1 stuffing after 0xFF; a read past the
 * segment or a stuffed bit that is not 0 yields 1 (non-strict) ---------- */
typedef struct {
    const uint8_t *d;
    size_t n, pos; /* pos: bits consumed */
} raw_t;

static uint32_t raw_bit(raw_t *r)
{
    const size_t byte = r->pos >> 3;
    if (byte >= r->n) return 1;
    const int stuffed_next = (r->pos & 7) == 7 && r->d[byte] == 0xFF;
    const uint32_t b = (r->d[byte] >> (7 - (r->pos & 7))) & 1u;
    r->pos++;
    if (stuffed_next) {
        if ((r->pos >> 3) >= r->n) return 1;
        const uint32_t s = (r->d[r->pos >> 3] >> 7) & 1u;
        r->pos++;
        if (s) return 1;
    }
    return b;
}

/* ---- block state --------------------------------------------------- */
enum { F_SIG = 1, F_NEG = 2, F_VISIT = 4, F_REFINED = 8 };

typedef struct {
    uint32_t w, h, pw; /* pw: padded row width */
    uint8_t *f;        /* (w + 2) x (h + 2) flags, one sample of zero border */
    uint32_t *mag;     /* magnitudes, same layout */
    int band, causal, bit;
    uint8_t cx[NCX];
    int raw; /* the current segment is raw */
    mq_t mq;
    raw_t rw;
} blk_t;

static void reset_contexts(blk_t *b)
{
    memset(b->cx, 0, sizeof(b->cx));
    b->cx[0] = 4;
    b->cx[CX_RL] = 3;
    b->cx[CX_UNI] = 46;
}

static uint32_t rd(blk_t *b, int cx) { return b->raw ? raw_bit(&b->rw) : mq_decode(&b->mq, &b->cx[cx]); }

/* Significance of the 8 neighbours of sample (x, y) (padded index i); with
 * the vertically causal style the three below a stripe's last row count as
 * insignificant unless they are past the block (bitplane.rs:688-703). */
static int below_hidden(const blk_t *b, uint32_t y)
{
    return b->causal && y + 1 < b->h && ((y + 1) >> 2) > (y >> 2);
}

static void counts(const blk_t *b, size_t i, uint32_t y, int *hc, int *vc, int *dc)
{
    const uint8_t *f = b->f;
    const size_t pw = b->pw;
    const int hide = below_hidden(b, y);
    *hc = (f[i - 1] & F_SIG) + (f[i + 1] & F_SIG);
    *vc = (f[i - pw] & F_SIG) + (hide ? 0 : (f[i + pw] & F_SIG));
    *dc = (f[i - pw - 1] & F_SIG) + (f[i - pw + 1] & F_SIG) +
          (hide ? 0 : (f[i + pw - 1] & F_SIG) + (f[i + pw + 1] & F_SIG));
}

/* Table D.1 */
static int zc_context(int band, int hc, int vc, int dc)
{
    if (band == 3) {
        const int hv = hc + vc;
        if (dc >= 3) return 8;
        if (dc == 2) return hv >= 1 ? 7 : 6;
        if (dc == 1) return hv >= 2 ? 5 : (hv == 1 ? 4 : 3);
        return hv >= 2 ? 2 : (hv == 1 ? 1 : 0);
    }
    if (band == 1) { /* HL: horizontally high-pass, the roles of h and v swap */
        const int t = hc;
        hc = vc;
        vc = t;
    }
    if (hc == 2) return 8;
    if (hc == 1) return vc >= 1 ? 7 : (dc >= 1 ? 6 : 5);
    if (vc == 2) return 4;
    if (vc == 1) return 3;
    return dc >= 2 ? 2 : (dc == 1 ? 1 : 0);
}

/* Table D.3: contribution of a horizontal or vertical neighbour pair */
static int contribution(uint8_t a, uint8_t c)
{
    int s = 0;
    if (a & F_SIG) s += (a & F_NEG) ? -1 : 1;
    if (c & F_SIG) s += (c & F_NEG) ? -1 : 1;
    return s > 0 ? 1 : (s < 0 ? -1 : 0);
}

static void decode_sign(blk_t *b, size_t i, uint32_t y)
{
    const uint8_t *f = b->f;
    const uint8_t below = below_hidden(b, y) ? 0 : f[i + b->pw];
    const int H = contribution(f[i - 1], f[i + 1]);
    const int V = contribution(f[i - b->pw], below);
    int cx, xr;
    if (H == 0 && V == 0) {
        cx = 9;
        xr = 0;
    } else if (H == 0) {
        cx = 10;
        xr = V < 0;
    } else {
        cx = 12 + H * V; /* H V agree: 13, H only: 12, disagree: 11 */
        xr = H < 0;
    }
    uint32_t s = rd(b, cx);
    if (!b->raw) s ^= (uint32_t)xr; /* a raw sign bit is the sign itself */
    if (s) b->f[i] |= F_NEG;
}

static void become_significant(blk_t *b, size_t i, uint32_t y)
{
    b->mag[i] |= 1u << b->bit;
    decode_sign(b, i, y);
    b->f[i] |= F_SIG;
}

static size_t at(const blk_t *b, uint32_t x, uint32_t y) { return (size_t)(y + 1) * b->pw + x + 1; }

static void pass_sigprop(blk_t *b)
{
    for (uint32_t y0 = 0; y0 < b->h; y0 += 4)
        for (uint32_t x = 0; x < b->w; ++x)
            for (uint32_t y = y0; y < y0 + 4 && y < b->h; ++y) {
                const size_t i = at(b, x, y);
                if (b->f[i] & F_SIG) continue;
                int hc, vc, dc;
                counts(b, i, y, &hc, &vc, &dc);
                if (hc + vc + dc == 0) continue;
                const uint32_t bit = rd(b, zc_context(b->band, hc, vc, dc));
                b->f[i] |= F_VISIT;
                if (bit) become_significant(b, i, y);
            }
}

static void pass_refine(blk_t *b)
{
    for (uint32_t y0 = 0; y0 < b->h; y0 += 4)
        for (uint32_t x = 0; x < b->w; ++x)
            for (uint32_t y = y0; y < y0 + 4 && y < b->h; ++y) {
                const size_t i = at(b, x, y);
                if ((b->f[i] & (F_SIG | F_VISIT)) != F_SIG) continue;
                int cx = 16; /* Table D.4 */
                if (!(b->f[i] & F_REFINED)) {
                    int hc, vc, dc;
                    counts(b, i, y, &hc, &vc, &dc);
                    cx = hc + vc + dc ? 15 : 14;
                }
                if (rd(b, cx)) b->mag[i] |= 1u << b->bit;
                b->f[i] |= F_REFINED;
            }
}

static void cleanup_sample(blk_t *b, uint32_t x, uint32_t y)
{
    const size_t i = at(b, x, y);
    if (b->f[i] & (F_SIG | F_VISIT)) return;
    int hc, vc, dc;
    counts(b, i, y, &hc, &vc, &dc);
    if (rd(b, zc_context(b->band, hc, vc, dc))) become_significant(b, i, y);
}

static void pass_cleanup(blk_t *b)
{
    for (uint32_t y0 = 0; y0 < b->h; y0 += 4)
        for (uint32_t x = 0; x < b->w; ++x) {
            uint32_t y = y0;
            if (y0 + 4 <= b->h) {
                /* run-length mode: four untouched samples, all of zero
                 * context (D.3.4) */
                int run = 1;
                for (uint32_t k = 0; k < 4 && run; ++k) {
                    const size_t i = at(b, x, y0 + k);
                    int hc, vc, dc;
                    counts(b, i, y0 + k, &hc, &vc, &dc);
                    run = !(b->f[i] & (F_SIG | F_VISIT)) && hc + vc + dc == 0;
                }
                if (run) {
                    if (!rd(b, CX_RL)) continue;
                    uint32_t r = rd(b, CX_UNI) << 1;
                    r |= rd(b, CX_UNI);
                    become_significant(b, at(b, x, y0 + r), y0 + r);
                    y = y0 + r + 1;
                }
            }
            for (; y < y0 + 4 && y < b->h; ++y) cleanup_sample(b, x, y);
        }
}

/* Passes [start, end) of the block from one segment (bitplane.rs:206-317). */
static void run_passes(blk_t *b, uint32_t bitplanes, uint32_t start, uint32_t end, int style)
{
    for (uint32_t p = start; p < end; ++p) {
        b->bit = (int)(bitplanes - 1 - (p + 2) / 3);
        switch (p % 3) {
        case 0:
            pass_cleanup(b);
            if (style & ORACLE_CB_SEGSYM) /* four UNIFORM symbols; non-strict: not checked */
                for (int k = 0; k < 4; ++k) (void)rd(b, CX_UNI);
            for (uint32_t y = 0; y < b->h; ++y)
                for (uint32_t x = 0; x < b->w; ++x) b->f[at(b, x, y)] &= (uint8_t)~F_VISIT;
            break;
        case 1:
            pass_sigprop(b);
            break;
        default:
            pass_refine(b);
            break;
        }
        if (style & ORACLE_CB_RESET) reset_contexts(b);
    }
}

/* Segment of pass p under selective bypass (Table D.9, segment.rs:583-593) */
uint32_t oracle_bypass_segment(uint32_t p)
{
    if (p < 10) return 0;
    return 1 + 2 * ((p - 10) / 3) + ((p - 10) % 3 == 2);
}

int oracle_classic_decode_block(const uint8_t *data, size_t len, const oracle_seg_t *segs,
                                size_t nseg, uint32_t w, uint32_t h, uint32_t bitplanes,
                                uint32_t passes, int band, int style, int32_t *out, size_t stride)
{
    if (w == 0 || h == 0) return ORACLE_OK;
    for (uint32_t y = 0; y < h; ++y) memset(out + (size_t)y * stride, 0, w * sizeof(int32_t));
    if (bitplanes == 0 || passes == 0) return ORACLE_OK;
    if (bitplanes > 31 || w > 1024 || h > 1024) return ORACLE_E_BLOCK;
    blk_t b;
    memset(&b, 0, sizeof(b));
    b.w = w;
    b.h = h;
    b.pw = w + 2;
    const size_t cells = (size_t)b.pw * (h + 2);
    b.f = (uint8_t *)calloc(cells, 1);
    b.mag = (uint32_t *)calloc(cells, sizeof(uint32_t));
    if (!b.f || !b.mag) {
        free(b.f);
        free(b.mag);
        return ORACLE_E_NOMEM;
    }
    b.band = band;
    b.causal = (style & ORACLE_CB_CAUSAL) != 0;
    reset_contexts(&b);
    const uint32_t maxp = 1 + 3 * (bitplanes - 1);
    const uint32_t endp = passes < maxp ? passes : maxp;
    int rc = ORACLE_OK;
    if (!(style & (ORACLE_CB_BYPASS | ORACLE_CB_TERMALL))) {
        /* one terminated segment: every pass from one MQ decoder */
        mq_init(&b.mq, data, len);
        run_passes(&b, bitplanes, 0, endp, style);
    } else {
        uint32_t expect = 0;
        for (size_t s = 0; s < nseg; ++s) {
            const oracle_seg_t *sg = &segs[s];
            if (sg->start != expect || sg->end < sg->start || (size_t)sg->off + sg->len > len) {
                rc = ORACLE_E_BLOCK;
                break;
            }
            expect = sg->end;
            const uint32_t e = sg->end < maxp ? sg->end : maxp;
            b.raw = !sg->arith;
            if (b.raw) {
                b.rw.d = data + sg->off;
                b.rw.n = sg->len;
                b.rw.pos = 0;
            } else {
                mq_init(&b.mq, data + sg->off, sg->len);
            }
            run_passes(&b, bitplanes, sg->start, e, style);
        }
        if (!rc && expect != passes) rc = ORACLE_E_BLOCK;
    }
    if (!rc)
        for (uint32_t y = 0; y < h; ++y)
            for (uint32_t x = 0; x < w; ++x) {
                const size_t i = at(&b, x, y);
                const int32_t m = (int32_t)(b.mag[i] & 0x7FFFFFFFu);
                out[(size_t)y * stride + x] = (b.f[i] & F_NEG) ? -m : m;
            }
    free(b.f);
    free(b.mag);
    return rc;
}
