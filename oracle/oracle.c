/* CPU oracle: HTJ2K codestream decode.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  Restates the reference CPU
 * decoder of frames-sg/signinum:
 *   main header          crates/signinum-j2k-native/src/j2c/codestream.rs:38-272,700-1000
 *   tiles / geometry     j2c/tile.rs:131-320,385-489,512-685
 *   code-block partition j2c/build.rs:113-258
 *   tag trees            j2c/tag_tree.rs
 *   packet headers (HT)  j2c/segment.rs:17-133,304-581; reader.rs:129-160
 *   progression orders   j2c/progression.rs
 *   dequantisation       j2c/decode.rs:1117-1187,1554-1718
 *   inverse DWT          j2c/idwt.rs:54-193,535-1143
 *   store / sign shift   j2c/decode.rs:1720-1902
 *   inverse MCT          j2c/mct.rs:11-156
 *   output conversion    src/lib.rs:1772-1835,2202-2295;
 *                        crates/signinum-j2k/src/decode.rs:120-170,391-545
 *   tile batch           crates/signinum-j2k/src/batch.rs:110-147,197-226
 *
 * Scope: HTJ2K (Part 15) codestreams with one quality layer and components
 * without sub-sampling; PPM and PPT packed packet headers
 * (codestream.rs:761-780, tile.rs:251-300,697-706, segment.rs:54-133);
 * signed components, whose flag the reference ignores (codestream.rs:
 * 647-649); non-zero image and tile origins at full resolution.  A reduced-
 * resolution decode with a non-zero image origin is refused: the
 * reference's store subtracts the full-resolution image offset from
 * reduced coordinates (decode.rs:1762-1790).  Anything else returns
 * ORACLE_E_UNSUPPORTED.
 */
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"
#include "oracle_internal.h"

#define MAXC 16
#define MAXR 33

typedef struct {
    uint32_t x0, y0, x1, y1;
} rect_t;

static uint32_t rw(rect_t r) { return r.x1 > r.x0 ? r.x1 - r.x0 : 0; }
static uint32_t rh(rect_t r) { return r.y1 > r.y0 ? r.y1 - r.y0 : 0; }
static uint32_t cdiv(uint32_t a, uint32_t b) { return (uint32_t)(((uint64_t)a + b - 1) / b); }

/* ------------------------------------------------------------------ */
/* Thread-local statistics of the last decode.                         */
static __thread uint32_t st_blocks, st_ref_blocks, st_max_passes;
static __thread int zero_refinement;

/* Test double of the reference's ZeroRefinementHtDecoder
 * (crates/signinum-j2k-native/src/lib.rs:2777-2792): when set, refinement
 * bytes are zeroed before each HT code-block is decoded. */
void oracle_set_zero_refinement(int enable) { zero_refinement = enable; }

void oracle_last_stats(uint32_t *blocks, uint32_t *refinement_blocks, uint32_t *max_passes)
{
    if (blocks) *blocks = st_blocks;
    if (refinement_blocks) *refinement_blocks = st_ref_blocks;
    if (max_passes) *max_passes = st_max_passes;
}

/* ------------------------------------------------------------------ */
/* Main header                                                          */

typedef struct {
    uint8_t precision, xr, yr;
    uint8_t scod;
    uint8_t nlevels, xcb, ycb, cbstyle, transform;
    uint8_t ppx[MAXR], ppy[MAXR];
    uint8_t qstyle, guard;
    int nsteps;
    uint16_t sexp[3 * MAXR + 1], smant[3 * MAXR + 1];
} comp_t;

typedef struct {
    uint32_t xsiz, ysiz, xo, yo, xt, yt, xto, yto;
    uint16_t ncomp;
    uint8_t prog;
    uint16_t nlayers;
    uint8_t mct;
    comp_t comp[MAXC];
    uint8_t skipped;
    size_t first_sot; /* offset of the first SOT marker */
    /* PPM packets (codestream.rs:761-780, 164-177): every PPM segment is
     * Zppm followed by (16-bit length, bytes) pairs; empty ones are dropped
     * and the rest, in Zppm order, are the packet headers of tile-part 0,
     * 1, ... of the codestream */
    size_t nppm;
    const uint8_t *ppm_data[4096];
    uint32_t ppm_len[4096];
    uint8_t ppm_z[4096];
} header_t;

typedef struct {
    const uint8_t *d;
    size_t len, pos;
} br_t;

static int rd8(br_t *b, uint32_t *v)
{
    if (b->pos + 1 > b->len) return -1;
    *v = b->d[b->pos++];
    return 0;
}
static int rd16(br_t *b, uint32_t *v)
{
    if (b->pos + 2 > b->len) return -1;
    *v = ((uint32_t)b->d[b->pos] << 8) | b->d[b->pos + 1];
    b->pos += 2;
    return 0;
}
static int rd32(br_t *b, uint32_t *v)
{
    if (b->pos + 4 > b->len) return -1;
    *v = ((uint32_t)b->d[b->pos] << 24) | ((uint32_t)b->d[b->pos + 1] << 16) |
         ((uint32_t)b->d[b->pos + 2] << 8) | b->d[b->pos + 3];
    b->pos += 4;
    return 0;
}
static int peek_marker(const br_t *b)
{
    if (b->pos + 2 > b->len || b->d[b->pos] != 0xFF) return -1;
    return b->d[b->pos + 1];
}
static int skip_segment(br_t *b)
{
    uint32_t l;
    if (rd16(b, &l) || l < 2 || b->pos + (l - 2) > b->len) return -1;
    b->pos += l - 2;
    return 0;
}

/* COD/COC coding-style parameters (codestream.rs:700-747) */
static int read_cs_params(br_t *b, uint8_t scod, comp_t *c)
{
    uint32_t v;
    if (rd8(b, &v) || v > 32) return -1;
    c->nlevels = (uint8_t)v;
    if (rd8(b, &v)) return -1;
    c->xcb = (uint8_t)(v + 2);
    if (rd8(b, &v)) return -1;
    c->ycb = (uint8_t)(v + 2);
    if (rd8(b, &v)) return -1;
    c->cbstyle = (uint8_t)v;
    if (rd8(b, &v) || v > 1) return -1;
    c->transform = (uint8_t)v;
    for (int r = 0; r <= c->nlevels; ++r) {
        if (scod & 1) {
            if (rd8(b, &v)) return -1;
            c->ppx[r] = (uint8_t)(v & 0xF);
            c->ppy[r] = (uint8_t)(v >> 4);
        } else {
            c->ppx[r] = 15;
            c->ppy[r] = 15;
        }
    }
    return 0;
}

/* QCD/QCC quantisation parameters (codestream.rs:880-960) */
static int read_q_params(br_t *b, size_t remaining, comp_t *c)
{
    uint32_t v;
    if (rd8(b, &v)) return -1;
    uint8_t style = v & 0x1F;
    if (style > 2) return -1;
    c->qstyle = style;
    c->guard = (uint8_t)((v >> 5) & 7);
    c->nsteps = 0;
    if (style == 0) {
        for (size_t i = 0; i < remaining; ++i) {
            if (rd8(b, &v)) return -1;
            if (c->nsteps < 3 * MAXR + 1) {
                c->sexp[c->nsteps] = (uint16_t)(v >> 3);
                c->smant[c->nsteps] = 0;
                c->nsteps++;
            }
        }
    } else if (style == 1) {
        if (rd16(b, &v)) return -1;
        c->sexp[0] = (uint16_t)(v >> 11);
        c->smant[0] = (uint16_t)(v & 0x7FF);
        c->nsteps = 1;
        /* any trailing bytes of the segment are ignored */
        if (remaining > 2) b->pos += remaining - 2;
    } else {
        size_t n = remaining / 2;
        for (size_t i = 0; i < n; ++i) {
            if (rd16(b, &v)) return -1;
            if (c->nsteps < 3 * MAXR + 1) {
                c->sexp[c->nsteps] = (uint16_t)(v >> 11);
                c->smant[c->nsteps] = (uint16_t)(v & 0x7FF);
                c->nsteps++;
            }
        }
    }
    return 0;
}

static int read_main_header(const uint8_t *data, size_t len, header_t *h, uint32_t target_w,
                            uint32_t target_h)
{
    br_t b = {data, len, 0};
    memset(h, 0, sizeof(*h));
    if (peek_marker(&b) != 0x4F) return ORACLE_E_MALFORMED;
    b.pos += 2;
    if (peek_marker(&b) != 0x51) return ORACLE_E_MALFORMED;
    b.pos += 2;
    uint32_t v;
    if (rd16(&b, &v) || rd16(&b, &v)) return ORACLE_E_MALFORMED;
    if (rd32(&b, &h->xsiz) || rd32(&b, &h->ysiz) || rd32(&b, &h->xo) || rd32(&b, &h->yo) ||
        rd32(&b, &h->xt) || rd32(&b, &h->yt) || rd32(&b, &h->xto) || rd32(&b, &h->yto))
        return ORACLE_E_MALFORMED;
    uint32_t csiz;
    if (rd16(&b, &csiz) || csiz == 0) return ORACLE_E_MALFORMED;
    if (csiz > MAXC) return ORACLE_E_UNSUPPORTED;
    if (h->xo >= h->xsiz || h->yo >= h->ysiz || h->xt == 0 || h->yt == 0)
        return ORACLE_E_MALFORMED;
    h->ncomp = (uint16_t)csiz;
    for (uint32_t i = 0; i < csiz; ++i) {
        uint32_t ssiz, xr, yr;
        if (rd8(&b, &ssiz) || rd8(&b, &xr) || rd8(&b, &yr)) return ORACLE_E_MALFORMED;
        h->comp[i].precision = (uint8_t)((ssiz & 0x7F) + 1);
        h->comp[i].xr = (uint8_t)xr;
        h->comp[i].yr = (uint8_t)yr;
        if (xr == 0 || yr == 0 || h->comp[i].precision > 32) return ORACLE_E_MALFORMED;
        /* the signed flag (ssiz bit 7) changes nothing (codestream.rs:647-649) */
    }
    /* the first tile holds a reference grid point of the image (B-4,
     * codestream.rs:582-596) */
    if ((uint64_t)h->xto + h->xt <= h->xo || (uint64_t)h->yto + h->yt <= h->yo ||
        h->xto > h->xo || h->yto > h->yo)
        return ORACLE_E_MALFORMED;

    int have_cod = 0, have_qcd = 0;
    int coc_set[MAXC] = {0}, qcc_set[MAXC] = {0};
    comp_t cod_c, qcd_c, coc_c[MAXC], qcc_c[MAXC];
    uint8_t cod_scod = 0;
    for (;;) {
        int m = peek_marker(&b);
        if (m < 0) return ORACLE_E_MALFORMED;
        if (m == 0x90) break;
        b.pos += 2;
        if (m == 0x52) { /* COD */
            uint32_t l, scod, prog, nl, mct;
            if (rd16(&b, &l) || rd8(&b, &scod) || rd8(&b, &prog) || rd16(&b, &nl) ||
                rd8(&b, &mct))
                return ORACLE_E_MALFORMED;
            if (prog > 4 || nl == 0 || nl > 32) return ORACLE_E_MALFORMED;
            h->prog = (uint8_t)prog;
            h->nlayers = (uint16_t)nl;
            h->mct = (mct == 1);
            memset(&cod_c, 0, sizeof(cod_c));
            if (read_cs_params(&b, (uint8_t)scod, &cod_c)) return ORACLE_E_MALFORMED;
            cod_scod = (uint8_t)scod;
            have_cod = 1;
        } else if (m == 0x53) { /* COC */
            uint32_t l, ci, scoc;
            if (rd16(&b, &l)) return ORACLE_E_MALFORMED;
            if ((csiz < 257 ? rd8(&b, &ci) : rd16(&b, &ci)) || rd8(&b, &scoc))
                return ORACLE_E_MALFORMED;
            if (ci >= csiz) return ORACLE_E_MALFORMED;
            memset(&coc_c[ci], 0, sizeof(comp_t));
            if (read_cs_params(&b, (uint8_t)scoc, &coc_c[ci])) return ORACLE_E_MALFORMED;
            coc_c[ci].scod = (uint8_t)scoc;
            coc_set[ci] = 1;
        } else if (m == 0x5C) { /* QCD */
            uint32_t l;
            if (rd16(&b, &l) || l < 3) return ORACLE_E_MALFORMED;
            if (read_q_params(&b, l - 3, &qcd_c)) return ORACLE_E_MALFORMED;
            have_qcd = 1;
        } else if (m == 0x5D) { /* QCC */
            uint32_t l, ci;
            if (rd16(&b, &l)) return ORACLE_E_MALFORMED;
            if (csiz < 257 ? rd8(&b, &ci) : rd16(&b, &ci)) return ORACLE_E_MALFORMED;
            size_t isz = csiz < 257 ? 1 : 2;
            if (ci >= csiz || l < 2 + isz + 1) return ORACLE_E_MALFORMED;
            if (read_q_params(&b, l - 2 - isz - 1, &qcc_c[ci])) return ORACLE_E_MALFORMED;
            qcc_set[ci] = 1;
        } else if (m == 0x60) { /* PPM */
            uint32_t l, z, n;
            if (rd16(&b, &l) || l < 3 || b.pos + (l - 2) > len) return ORACLE_E_MALFORMED;
            size_t end = b.pos + (l - 2);
            if (rd8(&b, &z)) return ORACLE_E_MALFORMED;
            while (b.pos < end) {
                if (rd16(&b, &n) || b.pos + n > end) return ORACLE_E_MALFORMED;
                if (n) {
                    if (h->nppm == sizeof(h->ppm_len) / sizeof(h->ppm_len[0]))
                        return ORACLE_E_UNSUPPORTED;
                    h->ppm_data[h->nppm] = data + b.pos;
                    h->ppm_len[h->nppm] = n;
                    h->ppm_z[h->nppm] = (uint8_t)z;
                    h->nppm++;
                }
                b.pos += n;
            }
        } else if (m == 0x50 || m == 0x59 || m == 0x5E || m == 0x55 || m == 0x64 ||
                   m == 0x63) {
            if (skip_segment(&b)) return ORACLE_E_MALFORMED;
        } else if (m >= 0x30 && m <= 0x3F) {
            /* no parameters */
        } else {
            return ORACLE_E_UNSUPPORTED;
        }
    }
    if (!have_cod || !have_qcd) return ORACLE_E_MALFORMED;
    h->first_sot = b.pos;
    /* PPM packets in Zppm order, stable (codestream.rs:169) */
    for (size_t i = 1; i < h->nppm; ++i)
        for (size_t j = i; j > 0 && h->ppm_z[j - 1] > h->ppm_z[j]; --j) {
            const uint8_t *pd = h->ppm_data[j];
            uint32_t pl = h->ppm_len[j];
            uint8_t pz = h->ppm_z[j];
            h->ppm_data[j] = h->ppm_data[j - 1];
            h->ppm_len[j] = h->ppm_len[j - 1];
            h->ppm_z[j] = h->ppm_z[j - 1];
            h->ppm_data[j - 1] = pd;
            h->ppm_len[j - 1] = pl;
            h->ppm_z[j - 1] = pz;
        }
    uint8_t min_nr = 255;
    for (uint32_t i = 0; i < csiz; ++i) {
        comp_t *c = &h->comp[i];
        uint8_t prec = c->precision, xr = c->xr, yr = c->yr;
        if (coc_set[i]) {
            uint8_t s = coc_c[i].scod | cod_scod;
            *c = coc_c[i];
            c->scod = s;
        } else {
            *c = cod_c;
            c->scod = cod_scod;
        }
        const comp_t *q = qcc_set[i] ? &qcc_c[i] : &qcd_c;
        c->qstyle = q->qstyle;
        c->guard = q->guard;
        c->nsteps = q->nsteps;
        memcpy(c->sexp, q->sexp, sizeof(c->sexp));
        memcpy(c->smant, q->smant, sizeof(c->smant));
        c->precision = prec;
        c->xr = xr;
        c->yr = yr;
        /* validate() at codestream.rs:188-220 */
        if (c->nsteps == 0) return ORACLE_E_MALFORMED;
        if (c->qstyle != 1 && c->nlevels > 0 && 1 + (c->nlevels - 1) * 3 + 2 >= c->nsteps)
            return ORACLE_E_MALFORMED;
        if (c->nlevels + 1 < min_nr) min_nr = (uint8_t)(c->nlevels + 1);
    }
    /* reduced resolution (codestream.rs:137-162) */
    uint32_t skipped = 0;
    if (target_w && target_h) {
        uint32_t wr = (h->xsiz - h->xo) / target_w, hr = (h->ysiz - h->yo) / target_h;
        uint32_t wl = wr ? 31 - (uint32_t)__builtin_clz(wr) : 0;
        uint32_t hl = hr ? 31 - (uint32_t)__builtin_clz(hr) : 0;
        skipped = wl < hl ? wl : hl;
    }
    if (skipped > (uint32_t)(min_nr - 1)) skipped = min_nr - 1;
    h->skipped = (uint8_t)skipped;
    if (skipped && (h->xo || h->yo)) return ORACLE_E_UNSUPPORTED;
    return ORACLE_OK;
}

/* ------------------------------------------------------------------ */
/* Geometry (tile.rs:385-489,512-685; build.rs:113-258)                */

typedef struct tagnode {
    uint32_t value, low;
    int known;
} tagnode_t;

typedef struct {
    uint32_t w, h, levels;
    uint32_t lw[40], lh[40], off[40];
    tagnode_t *nodes;
} tagtree_t;

/* A classic code-block's contribution from one packet: `len` bytes of
 * codeword segment `seg` carrying `passes` coding passes. */
typedef struct {
    const uint8_t *data;
    uint32_t len, seg, passes;
} contrib_t;

typedef struct {
    rect_t rect;
    uint32_t xi, yi;
    int included;
    uint32_t zblk, lblock, passes;
    const uint8_t *data; /* HT: cleanup followed by refinement */
    uint32_t lcup, lref;
    /* classic: contributions in codestream order, layers with any */
    contrib_t *contrib;
    uint32_t ncontrib, capcontrib, layers_in;
} cblk_t;

typedef struct {
    uint32_t ncbx, ncby;
    cblk_t *cbs;
    tagtree_t incl, zbp;
} prec_band_t;

typedef struct {
    int type; /* 0 LL, 1 HL, 2 LH, 3 HH */
    rect_t rect;
    float *coef;
    double *coef64;
} band_t;

typedef struct {
    rect_t rect;
    uint32_t ppx, ppy, npx, npy;
    uint32_t cbw, cbh;
    int nbands;
    band_t bands[3];
    prec_band_t *prec; /* npx*npy*nbands, [p * nbands + b] */
} res_t;

typedef struct {
    rect_t rect;  /* the tile-component, component coordinates (B-12) */
    rect_t trect; /* the tile on the reference grid */
    uint32_t xr, yr;
    int nres;
    res_t res[MAXR];
} tcomp_t;

static void tagtree_init(tagtree_t *t, uint32_t w, uint32_t h)
{
    memset(t, 0, sizeof(*t));
    t->w = w;
    t->h = h;
    if (w == 0 || h == 0) return;
    uint32_t cw = w, ch = h, total = 0, l = 0;
    for (;;) {
        t->lw[l] = cw;
        t->lh[l] = ch;
        t->off[l] = total;
        total += cw * ch;
        l++;
        if (cw == 1 && ch == 1) break;
        cw = (cw + 1) / 2;
        ch = (ch + 1) / 2;
    }
    t->levels = l;
    t->nodes = (tagnode_t *)calloc(total, sizeof(tagnode_t));
    for (uint32_t i = 0; i < total; ++i) t->nodes[i].value = 0xFFFFFFFFu;
}

typedef struct {
    const uint8_t *d;
    size_t len;
    size_t bitpos;
} pbr_t;

static int pbr_bit(pbr_t *r, uint32_t *bit)
{
    size_t byte = r->bitpos >> 3;
    if (byte >= r->len) return -1;
    int stuff = ((r->bitpos & 7) == 7) && r->d[byte] == 0xFF;
    *bit = (r->d[byte] >> (7 - (r->bitpos & 7))) & 1;
    r->bitpos++;
    if (stuff) {
        size_t b2 = r->bitpos >> 3;
        if (b2 >= r->len) return -1;
        uint32_t sb = (r->d[b2] >> (7 - (r->bitpos & 7))) & 1;
        r->bitpos++;
        if (sb) return -1;
    }
    return 0;
}

static int pbr_bits(pbr_t *r, uint32_t n, uint32_t *v)
{
    uint32_t x = 0, bit;
    for (uint32_t i = 0; i < n; ++i) {
        if (pbr_bit(r, &bit)) return -1;
        x = (x << 1) | bit;
    }
    *v = x;
    return 0;
}

/* Tag-tree decode (B.10.2): returns 1 when the leaf value is below
 * `threshold`, 0 otherwise; the leaf value is stored once known. */
static int tagtree_decode(tagtree_t *t, uint32_t x, uint32_t y, uint32_t threshold, pbr_t *r,
                          uint32_t *value)
{
    tagnode_t *path[40];
    uint32_t cx = x, cy = y;
    for (uint32_t l = 0; l < t->levels; ++l) {
        path[l] = &t->nodes[t->off[l] + cy * t->lw[l] + cx];
        cx >>= 1;
        cy >>= 1;
    }
    uint32_t low = 0;
    for (int l = (int)t->levels - 1; l >= 0; --l) {
        tagnode_t *n = path[l];
        if (low > n->low)
            n->low = low;
        else
            low = n->low;
        while (low < threshold && low < n->value) {
            uint32_t bit;
            if (pbr_bit(r, &bit)) return -1;
            if (bit)
                n->value = low;
            else
                low++;
        }
        n->low = low;
    }
    *value = path[0]->value;
    return path[0]->value < threshold ? 1 : 0;
}

static int build_tcomp(const header_t *h, const comp_t *c, rect_t tile, tcomp_t *tc)
{
    memset(tc, 0, sizeof(*tc));
    /* tile-component rectangle (B-12, tile.rs:327-357) */
    rect_t trect = {cdiv(tile.x0, c->xr), cdiv(tile.y0, c->yr), cdiv(tile.x1, c->xr),
                    cdiv(tile.y1, c->yr)};
    tc->rect = trect;
    tc->trect = tile;
    tc->xr = c->xr;
    tc->yr = c->yr;
    tc->nres = c->nlevels + 1;
    uint32_t nl = c->nlevels;
    for (int r = 0; r < tc->nres; ++r) {
        res_t *rs = &tc->res[r];
        uint32_t s = nl - (uint32_t)r;
        uint64_t den = 1ull << s;
        rs->rect.x0 = (uint32_t)(((uint64_t)trect.x0 + den - 1) / den);
        rs->rect.y0 = (uint32_t)(((uint64_t)trect.y0 + den - 1) / den);
        rs->rect.x1 = (uint32_t)(((uint64_t)trect.x1 + den - 1) / den);
        rs->rect.y1 = (uint32_t)(((uint64_t)trect.y1 + den - 1) / den);
        rs->ppx = c->ppx[r];
        rs->ppy = c->ppy[r];
        rs->npx = rw(rs->rect) ? cdiv(rs->rect.x1, 1u << rs->ppx) - (rs->rect.x0 >> rs->ppx) : 0;
        rs->npy = rh(rs->rect) ? cdiv(rs->rect.y1, 1u << rs->ppy) - (rs->rect.y0 >> rs->ppy) : 0;
        uint32_t xcb = c->xcb, ycb = c->ycb;
        if (r > 0) {
            if (rs->ppx == 0 || rs->ppy == 0) {
                if (rs->npx && rs->npy) return ORACLE_E_MALFORMED;
            } else {
                if (xcb > rs->ppx - 1) xcb = rs->ppx - 1;
                if (ycb > rs->ppy - 1) ycb = rs->ppy - 1;
            }
        } else {
            if (xcb > rs->ppx) xcb = rs->ppx;
            if (ycb > rs->ppy) ycb = rs->ppy;
        }
        rs->cbw = 1u << xcb;
        rs->cbh = 1u << ycb;
        /* sub-bands (B-15) */
        rs->nbands = r == 0 ? 1 : 3;
        uint32_t nb = r == 0 ? nl : nl - (uint32_t)r + 1;
        for (int bi = 0; bi < rs->nbands; ++bi) {
            int type = r == 0 ? 0 : bi + 1;
            uint32_t xob = (type == 1 || type == 3) ? 1 : 0;
            uint32_t yob = (type == 2 || type == 3) ? 1 : 0;
            uint64_t nx = nb > 0 ? ((uint64_t)1 << (nb - 1)) * xob : 0;
            uint64_t ny = nb > 0 ? ((uint64_t)1 << (nb - 1)) * yob : 0;
            uint64_t d = 1ull << nb;
#define SBC(v, n) ((uint32_t)((((uint64_t)(v) > (n) ? (uint64_t)(v) - (n) : 0) + d - 1) / d))
            band_t *bd = &rs->bands[bi];
            bd->type = type;
            bd->rect.x0 = SBC(trect.x0, nx);
            bd->rect.x1 = SBC(trect.x1, nx);
            bd->rect.y0 = SBC(trect.y0, ny);
            bd->rect.y1 = SBC(trect.y1, ny);
#undef SBC
        }
        /* precincts and code-blocks (tile.rs:542-645, build.rs:113-258) */
        uint32_t np = rs->npx * rs->npy;
        if (np == 0) continue;
        rs->prec = (prec_band_t *)calloc((size_t)np * rs->nbands, sizeof(prec_band_t));
        if (!rs->prec) return ORACLE_E_NOMEM;
        uint32_t ppx = rs->ppx, ppy = rs->ppy;
        uint32_t xs = (rs->rect.x0 >> ppx) << ppx, ys = (rs->rect.y0 >> ppy) << ppy;
        if (r > 0) {
            ppx -= 1;
            ppy -= 1;
            xs /= 2;
            ys /= 2;
        }
        for (uint32_t py = 0; py < rs->npy; ++py) {
            for (uint32_t px = 0; px < rs->npx; ++px) {
                rect_t pr;
                pr.x0 = px * (1u << ppx) + xs;
                pr.y0 = py * (1u << ppy) + ys;
                pr.x1 = pr.x0 + (1u << ppx);
                pr.y1 = pr.y0 + (1u << ppy);
                for (int bi = 0; bi < rs->nbands; ++bi) {
                    prec_band_t *pb = &rs->prec[(size_t)(py * rs->npx + px) * rs->nbands + bi];
                    rect_t sb = rs->bands[bi].rect;
                    uint32_t ax0 = pr.x0 > sb.x0 ? pr.x0 : sb.x0;
                    uint32_t ay0 = pr.y0 > sb.y0 ? pr.y0 : sb.y0;
                    uint32_t ax1 = pr.x1 < sb.x1 ? pr.x1 : sb.x1;
                    uint32_t ay1 = pr.y1 < sb.y1 ? pr.y1 : sb.y1;
                    uint32_t cx0 = (ax0 / rs->cbw) * rs->cbw, cy0 = (ay0 / rs->cbh) * rs->cbh;
                    uint32_t cx1 = cdiv(ax1, rs->cbw) * rs->cbw, cy1 = cdiv(ay1, rs->cbh) * rs->cbh;
                    uint32_t ncx = (rw(sb) == 0 || cx1 <= cx0) ? 0 : (cx1 - cx0) / rs->cbw;
                    uint32_t ncy = (rh(sb) == 0 || cy1 <= cy0) ? 0 : (cy1 - cy0) / rs->cbh;
                    pb->ncbx = ncx;
                    pb->ncby = ncy;
                    if (ncx && ncy) {
                        pb->cbs = (cblk_t *)calloc((size_t)ncx * ncy, sizeof(cblk_t));
                        if (!pb->cbs) return ORACLE_E_NOMEM;
                    }
                    for (uint32_t j = 0; j < ncy; ++j) {
                        for (uint32_t i = 0; i < ncx; ++i) {
                            cblk_t *cb = &pb->cbs[j * ncx + i];
                            rect_t a = {cx0 + i * rs->cbw, cy0 + j * rs->cbh,
                                        cx0 + (i + 1) * rs->cbw, cy0 + (j + 1) * rs->cbh};
                            cb->rect.x0 = a.x0 > sb.x0 ? a.x0 : sb.x0;
                            cb->rect.y0 = a.y0 > sb.y0 ? a.y0 : sb.y0;
                            cb->rect.x1 = a.x1 < sb.x1 ? a.x1 : sb.x1;
                            cb->rect.y1 = a.y1 < sb.y1 ? a.y1 : sb.y1;
                            if (cb->rect.x1 < cb->rect.x0) cb->rect.x1 = cb->rect.x0;
                            if (cb->rect.y1 < cb->rect.y0) cb->rect.y1 = cb->rect.y0;
                            cb->xi = i;
                            cb->yi = j;
                            cb->lblock = 3;
                        }
                    }
                    tagtree_init(&pb->incl, ncx, ncy);
                    tagtree_init(&pb->zbp, ncx, ncy);
                }
            }
        }
    }
    return ORACLE_OK;
}

static void free_tcomp(tcomp_t *tc)
{
    for (int r = 0; r < tc->nres; ++r) {
        res_t *rs = &tc->res[r];
        if (rs->prec) {
            uint32_t n = rs->npx * rs->npy * (uint32_t)rs->nbands;
            for (uint32_t i = 0; i < n; ++i) {
                for (uint32_t k = 0; rs->prec[i].cbs && k < rs->prec[i].ncbx * rs->prec[i].ncby; ++k)
                    free(rs->prec[i].cbs[k].contrib);
                free(rs->prec[i].cbs);
                free(rs->prec[i].incl.nodes);
                free(rs->prec[i].zbp.nodes);
            }
            free(rs->prec);
            rs->prec = NULL;
        }
        for (int b = 0; b < rs->nbands; ++b) {
            free(rs->bands[b].coef);
            free(rs->bands[b].coef64);
            rs->bands[b].coef = NULL;
            rs->bands[b].coef64 = NULL;
        }
    }
}

/* ------------------------------------------------------------------ */
/* Tiles and tile parts (tile.rs:131-320)                              */

/* One tile-part: its body and, when packet headers are packed (PPT of
 * this tile-part in Zppt order, then the tile-part's PPM packet;
 * tile.rs:291-309), the header pieces read one after another. */
#define MAXHDR 64
typedef struct {
    const uint8_t *body;
    size_t body_len;
    int nhdr;
    const uint8_t *hdr[MAXHDR];
    size_t hdr_len[MAXHDR];
} tpart_t;

typedef struct {
    rect_t rect;
    tpart_t *parts;
    size_t nparts, cap;
    int seen;
    comp_t comp[MAXC];
    uint8_t prog, mct;
    uint16_t nlayers;
} tile_t;

static tpart_t *add_part(tile_t *t)
{
    if (t->nparts == t->cap) {
        size_t cap = t->cap * 2 + 4;
        tpart_t *np = (tpart_t *)realloc(t->parts, cap * sizeof(tpart_t));
        if (!np) return NULL;
        t->parts = np;
        t->cap = cap;
    }
    tpart_t *p = &t->parts[t->nparts++];
    memset(p, 0, sizeof(*p));
    return p;
}

static int parse_tiles(const uint8_t *data, size_t len, const header_t *h, tile_t *tiles,
                       uint32_t ntiles)
{
    uint32_t ntx = cdiv(h->xsiz - h->xto, h->xt);
    for (uint32_t i = 0; i < ntiles; ++i) {
        tile_t *t = &tiles[i];
        uint32_t tx = i % ntx, ty = i / ntx;
        uint32_t x0 = h->xto + tx * h->xt, y0 = h->yto + ty * h->yt;
        t->rect.x0 = x0 > h->xo ? x0 : h->xo;
        t->rect.y0 = y0 > h->yo ? y0 : h->yo;
        t->rect.x1 = x0 + h->xt < h->xsiz ? x0 + h->xt : h->xsiz;
        t->rect.y1 = y0 + h->yt < h->ysiz ? y0 + h->yt : h->ysiz;
        memcpy(t->comp, h->comp, sizeof(t->comp));
        t->prog = h->prog;
        t->mct = h->mct;
        t->nlayers = h->nlayers;
    }
    br_t b = {data, len, h->first_sot};
    size_t tp_index = 0; /* tile-parts seen, in codestream order */
    while (peek_marker(&b) == 0x90) {
        b.pos += 2;
        int nppt = 0;
        const uint8_t *ppt[MAXHDR];
        size_t ppt_len[MAXHDR];
        uint8_t ppt_z[MAXHDR];
        uint32_t lsot, isot, psot, tpsot, tnsot;
        if (rd16(&b, &lsot) || rd16(&b, &isot) || rd32(&b, &psot) || rd8(&b, &tpsot) ||
            rd8(&b, &tnsot))
            return ORACLE_E_MALFORMED;
        if (isot >= ntiles) return ORACLE_E_MALFORMED;
        size_t start = b.pos;
        size_t data_len;
        if (psot == 0)
            data_len = len - b.pos;
        else {
            if (psot < 12) return ORACLE_E_MALFORMED;
            data_len = psot - 12;
        }
        tile_t *t = &tiles[isot];
        for (;;) {
            int m = peek_marker(&b);
            if (m < 0) return ORACLE_OK; /* non-strict: stop at a broken tile part */
            if (m == 0x93) {
                b.pos += 2;
                break;
            }
            b.pos += 2;
            if (m == 0x52) {
                uint32_t l, scod, prog, nl, mct;
                if (rd16(&b, &l) || rd8(&b, &scod) || rd8(&b, &prog) || rd16(&b, &nl) ||
                    rd8(&b, &mct))
                    return ORACLE_E_MALFORMED;
                comp_t cc;
                memset(&cc, 0, sizeof(cc));
                if (prog > 4 || read_cs_params(&b, (uint8_t)scod, &cc)) return ORACLE_E_MALFORMED;
                t->mct = (mct == 1);
                t->nlayers = (uint16_t)nl;
                t->prog = (uint8_t)prog;
                for (int c = 0; c < h->ncomp; ++c) {
                    comp_t *tc = &t->comp[c];
                    tc->scod |= (uint8_t)scod;
                    tc->nlevels = cc.nlevels;
                    tc->xcb = cc.xcb;
                    tc->ycb = cc.ycb;
                    tc->cbstyle = cc.cbstyle;
                    tc->transform = cc.transform;
                    memcpy(tc->ppx, cc.ppx, sizeof(cc.ppx));
                    memcpy(tc->ppy, cc.ppy, sizeof(cc.ppy));
                }
            } else if (m == 0x53) {
                uint32_t l, ci, scoc;
                if (rd16(&b, &l)) return ORACLE_E_MALFORMED;
                if ((h->ncomp < 257 ? rd8(&b, &ci) : rd16(&b, &ci)) || rd8(&b, &scoc))
                    return ORACLE_E_MALFORMED;
                if (ci >= h->ncomp) return ORACLE_E_MALFORMED;
                comp_t cc;
                memset(&cc, 0, sizeof(cc));
                if (read_cs_params(&b, (uint8_t)scoc, &cc)) return ORACLE_E_MALFORMED;
                comp_t *tc = &t->comp[ci];
                tc->scod |= (uint8_t)scoc;
                tc->nlevels = cc.nlevels;
                tc->xcb = cc.xcb;
                tc->ycb = cc.ycb;
                tc->cbstyle = cc.cbstyle;
                tc->transform = cc.transform;
                memcpy(tc->ppx, cc.ppx, sizeof(cc.ppx));
                memcpy(tc->ppy, cc.ppy, sizeof(cc.ppy));
            } else if (m == 0x5C || m == 0x5D) {
                uint32_t l, ci = 0;
                if (rd16(&b, &l)) return ORACLE_E_MALFORMED;
                size_t isz = 0;
                if (m == 0x5D) {
                    if (h->ncomp < 257 ? rd8(&b, &ci) : rd16(&b, &ci)) return ORACLE_E_MALFORMED;
                    isz = h->ncomp < 257 ? 1 : 2;
                    if (ci >= h->ncomp) return ORACLE_E_MALFORMED;
                }
                if (l < 3 + isz) return ORACLE_E_MALFORMED;
                comp_t q;
                if (read_q_params(&b, l - 3 - isz, &q)) return ORACLE_E_MALFORMED;
                for (int c = 0; c < h->ncomp; ++c) {
                    if (m == 0x5D && (uint32_t)c != ci) continue;
                    t->comp[c].qstyle = q.qstyle;
                    t->comp[c].guard = q.guard;
                    t->comp[c].nsteps = q.nsteps;
                    memcpy(t->comp[c].sexp, q.sexp, sizeof(q.sexp));
                    memcpy(t->comp[c].smant, q.smant, sizeof(q.smant));
                }
            } else if (m == 0xD9) {
                break;
            } else if (m == 0x61) { /* PPT (tile.rs:251-258,697-706) */
                uint32_t l, z;
                if (h->nppm) return ORACLE_E_MALFORMED; /* PPM and PPT together */
                if (rd16(&b, &l) || l < 3 || b.pos + (l - 2) > len || rd8(&b, &z))
                    return ORACLE_E_MALFORMED;
                if (nppt == MAXHDR - 1) return ORACLE_E_UNSUPPORTED;
                ppt[nppt] = data + b.pos;
                ppt_len[nppt] = l - 3;
                ppt_z[nppt] = (uint8_t)z;
                nppt++;
                b.pos += l - 3;
            } else if (m == 0x58 || m == 0x64) {
                if (skip_segment(&b)) return ORACLE_E_MALFORMED;
            } else if (m >= 0x30 && m <= 0x3F) {
            } else {
                return ORACLE_E_UNSUPPORTED;
            }
        }
        size_t used = b.pos - start;
        if (used > data_len) return ORACLE_OK;
        size_t remain = data_len - used;
        if (b.pos + remain > len) return ORACLE_E_MALFORMED;
        tpart_t *tp = add_part(t);
        if (!tp) return ORACLE_E_NOMEM;
        tp->body = data + b.pos;
        tp->body_len = remain;
        /* PPT pieces in Zppt order (stable), then the PPM packet */
        for (int i = 1; i < nppt; ++i)
            for (int j = i; j > 0 && ppt_z[j - 1] > ppt_z[j]; --j) {
                const uint8_t *pd = ppt[j];
                size_t pl = ppt_len[j];
                uint8_t pz = ppt_z[j];
                ppt[j] = ppt[j - 1];
                ppt_len[j] = ppt_len[j - 1];
                ppt_z[j] = ppt_z[j - 1];
                ppt[j - 1] = pd;
                ppt_len[j - 1] = pl;
                ppt_z[j - 1] = pz;
            }
        for (int i = 0; i < nppt; ++i) {
            tp->hdr[tp->nhdr] = ppt[i];
            tp->hdr_len[tp->nhdr] = ppt_len[i];
            tp->nhdr++;
        }
        if (tp_index < h->nppm) {
            tp->hdr[tp->nhdr] = h->ppm_data[tp_index];
            tp->hdr_len[tp->nhdr] = h->ppm_len[tp_index];
            tp->nhdr++;
        }
        tp_index++;
        t->seen = 1;
        b.pos += remain;
    }
    return ORACLE_OK;
}

/* ------------------------------------------------------------------ */
/* Progression (progression.rs; B.12)                                   */

typedef struct {
    uint16_t l, r, c;
    uint32_t p;
} prog_t;

typedef struct {
    prog_t *v;
    size_t n, cap;
} prog_list_t;

static int prog_push(prog_list_t *pl, uint32_t l, uint32_t r, uint32_t c, uint32_t p)
{
    if (pl->n == pl->cap) {
        size_t cap = pl->cap ? pl->cap * 2 : 256;
        prog_t *nv = (prog_t *)realloc(pl->v, cap * sizeof(prog_t));
        if (!nv) return -1;
        pl->v = nv;
        pl->cap = cap;
    }
    pl->v[pl->n].l = (uint16_t)l;
    pl->v[pl->n].r = (uint16_t)r;
    pl->v[pl->n].c = (uint16_t)c;
    pl->v[pl->n].p = p;
    pl->n++;
    return 0;
}

/* Reference-grid position of precinct rows / columns (tile.rs:544-645):
 * the first at the tile origin, or at the next multiple of the step
 * XRsiz * 2^(PP + NL - r) when the resolution's origin times 2^(NL - r) is a
 * multiple of it; each next one at the next multiple of the step. */
static uint64_t first_position(uint64_t origin, uint64_t res_origin, uint32_t s, uint64_t step)
{
    if (origin % step != 0 && ((res_origin << s) % step) == 0) return (origin / step + 1) * step;
    return origin;
}

typedef struct {
    uint64_t y, x;
    uint32_t c, r, idx;
} ppos_t;

static int prog_order;
static int ppos_cmp(const void *pa, const void *pb)
{
    const ppos_t *a = (const ppos_t *)pa, *b = (const ppos_t *)pb;
    uint64_t ka[5], kb[5];
    if (prog_order == 2) { /* RPCL: resolution, y, x, component */
        uint64_t t1[5] = {a->r, a->y, a->x, a->c, a->idx}, t2[5] = {b->r, b->y, b->x, b->c, b->idx};
        memcpy(ka, t1, sizeof(ka));
        memcpy(kb, t2, sizeof(kb));
    } else if (prog_order == 3) { /* PCRL: y, x, component, resolution */
        uint64_t t1[5] = {a->y, a->x, a->c, a->r, a->idx}, t2[5] = {b->y, b->x, b->c, b->r, b->idx};
        memcpy(ka, t1, sizeof(ka));
        memcpy(kb, t2, sizeof(kb));
    } else { /* CPRL: component, y, x, resolution */
        uint64_t t1[5] = {a->c, a->y, a->x, a->r, a->idx}, t2[5] = {b->c, b->y, b->x, b->r, b->idx};
        memcpy(ka, t1, sizeof(ka));
        memcpy(kb, t2, sizeof(kb));
    }
    for (int i = 0; i < 5; ++i)
        if (ka[i] != kb[i]) return ka[i] < kb[i] ? -1 : 1;
    /* equal keys: keep the generation order (component, resolution, index) */
    if (a->c != b->c) return a->c < b->c ? -1 : 1;
    if (a->r != b->r) return a->r < b->r ? -1 : 1;
    return 0;
}

static int build_progression(const tile_t *t, int ncomp, const tcomp_t *tcs, prog_list_t *pl)
{
    uint32_t L = t->nlayers, maxr = 0;
    for (int c = 0; c < ncomp; ++c)
        if ((uint32_t)tcs[c].nres > maxr) maxr = (uint32_t)tcs[c].nres;
    if (t->prog == 0 || t->prog == 1) {
        if (t->prog == 0) {
            for (uint32_t l = 0; l < L; ++l)
                for (uint32_t r = 0; r < maxr; ++r)
                    for (int c = 0; c < ncomp; ++c) {
                        if ((int)r >= tcs[c].nres) continue;
                        uint32_t np = tcs[c].res[r].npx * tcs[c].res[r].npy;
                        for (uint32_t p = 0; p < np; ++p)
                            if (prog_push(pl, l, r, (uint32_t)c, p)) return ORACLE_E_NOMEM;
                    }
        } else {
            for (uint32_t r = 0; r < maxr; ++r)
                for (uint32_t l = 0; l < L; ++l)
                    for (int c = 0; c < ncomp; ++c) {
                        if ((int)r >= tcs[c].nres) continue;
                        uint32_t np = tcs[c].res[r].npx * tcs[c].res[r].npy;
                        for (uint32_t p = 0; p < np; ++p)
                            if (prog_push(pl, l, r, (uint32_t)c, p)) return ORACLE_E_NOMEM;
                    }
        }
        return ORACLE_OK;
    }
    /* position-driven orders: every precinct of every component and
     * resolution sorted by its reference-grid position (progression.rs:
     * 257-345) */
    size_t total = 0;
    for (int c = 0; c < ncomp; ++c)
        for (int r = 0; r < tcs[c].nres; ++r) total += (size_t)tcs[c].res[r].npx * tcs[c].res[r].npy;
    ppos_t *v = (ppos_t *)malloc((total + 1) * sizeof(ppos_t));
    if (!v) return ORACLE_E_NOMEM;
    size_t n = 0;
    for (int c = 0; c < ncomp; ++c)
        for (int r = 0; r < tcs[c].nres; ++r) {
            const res_t *rs = &tcs[c].res[r];
            if (rs->npx * rs->npy == 0) continue;
            uint32_t s = (uint32_t)(tcs[c].nres - 1 - r);
            uint64_t stepx = (uint64_t)tcs[c].xr << (rs->ppx + s);
            uint64_t stepy = (uint64_t)tcs[c].yr << (rs->ppy + s);
            uint64_t ry = first_position(tcs[c].trect.y0, rs->rect.y0, s, stepy);
            for (uint32_t py = 0; py < rs->npy; ++py) {
                uint64_t rx = first_position(tcs[c].trect.x0, rs->rect.x0, s, stepx);
                for (uint32_t px = 0; px < rs->npx; ++px) {
                    v[n++] = (ppos_t){ry, rx, (uint32_t)c, (uint32_t)r, py * rs->npx + px};
                    rx = (rx / stepx + 1) * stepx;
                }
                ry = (ry / stepy + 1) * stepy;
            }
        }
    prog_order = t->prog;
    /* merge sort is stable; qsort with the generation order as last key */
    qsort(v, n, sizeof(ppos_t), ppos_cmp);
    for (size_t i = 0; i < n; ++i)
        for (uint32_t l = 0; l < L; ++l)
            if (prog_push(pl, l, v[i].r, v[i].c, v[i].idx)) {
                free(v);
                return ORACLE_E_NOMEM;
            }
    free(v);
    return ORACLE_OK;
}

/* ------------------------------------------------------------------ */
/* Packet headers, HT only (segment.rs:17-133,304-581)                  */

static int read_ht_passes(pbr_t *r, uint32_t *n)
{
    uint32_t b;
    uint32_t np = 1;
    if (pbr_bits(r, 1, &b)) return -1;
    if (b) {
        np = 2;
        if (pbr_bits(r, 1, &b)) return -1;
        if (b) {
            uint32_t e;
            if (pbr_bits(r, 2, &e)) return -1;
            np = 3 + e;
            if (e == 3) {
                if (pbr_bits(r, 5, &e)) return -1;
                np = 6 + e;
                if (e == 31) {
                    if (pbr_bits(r, 7, &e)) return -1;
                    np = 37 + e;
                }
            }
        }
    }
    *n = np;
    return 0;
}

static uint32_t ilog2u(uint32_t v) { return 31 - (uint32_t)__builtin_clz(v); }

/* Table B.4: number of coding passes (segment.rs:481-506). */
static int read_classic_passes(pbr_t *r, uint32_t *n)
{
    uint32_t b;
    if (pbr_bits(r, 1, &b)) return -1;
    if (!b) {
        *n = 1;
        return 0;
    }
    if (pbr_bits(r, 1, &b)) return -1;
    if (!b) {
        *n = 2;
        return 0;
    }
    uint32_t v;
    if (pbr_bits(r, 2, &v)) return -1;
    if (v < 3) {
        *n = 3 + v;
        return 0;
    }
    if (pbr_bits(r, 5, &v)) return -1;
    if (v < 31) {
        *n = 6 + v;
        return 0;
    }
    if (pbr_bits(r, 7, &v)) return -1;
    *n = 37 + v;
    return 0;
}

/* Codeword segment of coding pass p (segment.rs:223-239): every pass with
 * termination on each pass, Table D.9 with selective bypass, otherwise one
 * per layer that contributes. */
static uint32_t classic_segment(const cblk_t *cb, int style, uint32_t p)
{
    if (style & ORACLE_CB_TERMALL) return p;
    if (style & ORACLE_CB_BYPASS) return oracle_bypass_segment(p);
    return cb->layers_in;
}

/* Classic code-block header after inclusion (segment.rs:168-301): passes,
 * Lblock increment, then one length per codeword segment the new passes
 * touch, of Lblock + floor(log2(passes in it)) bits. */
static int read_classic_block(pbr_t *r, cblk_t *cb, int style)
{
    uint32_t added, bit;
    if (read_classic_passes(r, &added)) return -1;
    for (;;) {
        if (pbr_bits(r, 1, &bit)) return -1;
        if (!bit) break;
        cb->lblock++;
    }
    const uint32_t prev = cb->passes, cum = prev + added;
    if (cum > 1 + 3 * 31) return -1;
    uint32_t last = classic_segment(cb, style, prev), count = 0;
    for (uint32_t p = prev; p <= cum; ++p) {
        const uint32_t sg = p < cum ? classic_segment(cb, style, p) : last + 1;
        if (p < cum && sg == last) {
            count++;
            continue;
        }
        if (count) {
            const uint32_t nb = cb->lblock + ilog2u(count);
            uint32_t len;
            if (nb > 32 || pbr_bits(r, nb, &len)) return -1;
            if (cb->ncontrib == cb->capcontrib) {
                uint32_t cap = cb->capcontrib ? 2 * cb->capcontrib : 4;
                contrib_t *v = (contrib_t *)realloc(cb->contrib, cap * sizeof(contrib_t));
                if (!v) return -1;
                cb->contrib = v;
                cb->capcontrib = cap;
            }
            cb->contrib[cb->ncontrib++] = (contrib_t){NULL, len, last, count};
        }
        last = sg;
        count = 1;
    }
    cb->passes = cum;
    cb->layers_in++;
    return 0;
}

/* Packets of one tile-part (segment.rs:54-133): while the header stream is
 * not at its end, the next packet of the progression: an SOP marker in the
 * body is skipped, the header is read from the header stream (the body
 * itself when headers are not packed; with packed headers the next piece
 * once the current one is exhausted, tile.rs:55-67), aligned, followed by
 * an EPH marker when signalled, then the included blocks' bytes come from
 * the body.  Returns -1 on a malformed packet (the caller stops, non-strict
 * segment.rs:36-50). */
static int parse_tile_part(const tpart_t *tp, size_t *next, int ncomp, tcomp_t *tcs,
                           const comp_t *comps, const prog_list_t *pl)
{
    const int packed = tp->nhdr > 0;
    int hi = 0;
    pbr_t hb = {packed ? tp->hdr[0] : tp->body, packed ? tp->hdr_len[0] : tp->body_len, 0};
    size_t body = 0; /* byte position in the body (packed headers) */
    (void)ncomp;
    for (;;) {
        if (packed && (hb.bitpos >> 3) >= hb.len && hi + 1 < tp->nhdr) {
            ++hi;
            hb.d = tp->hdr[hi];
            hb.len = tp->hdr_len[hi];
            hb.bitpos = 0;
        }
        if ((hb.bitpos >> 3) >= hb.len) break;
        if (*next >= pl->n) return -1;
        const prog_t *pg = &pl->v[(*next)++];
        tcomp_t *tc = &tcs[pg->c];
        res_t *rs = &tc->res[pg->r];
        const comp_t *cp = &comps[pg->c];
        /* SOP in the body */
        const uint8_t *bd = tp->body;
        size_t bpos = packed ? body : (hb.bitpos >> 3);
        if ((cp->scod & 2) && bpos + 1 < tp->body_len && bd[bpos] == 0xFF && bd[bpos + 1] == 0x91) {
            if (bpos + 6 > tp->body_len) return -1;
            if (packed)
                body += 6;
            else
                hb.bitpos += 6 * 8;
        }
        pbr_t *r = &hb;
        uint32_t nonzero;
        if (pbr_bits(r, 1, &nonzero)) return -1;
        if (nonzero) {
            for (int bi = 0; bi < rs->nbands; ++bi) {
                prec_band_t *pb = &rs->prec[(size_t)pg->p * rs->nbands + bi];
                for (uint32_t k = 0; k < pb->ncbx * pb->ncby; ++k) {
                    cblk_t *cb = &pb->cbs[k];
                    int first = !cb->included;
                    uint32_t inc, v;
                    if (cb->included) {
                        if (pbr_bits(r, 1, &inc)) return -1;
                    } else {
                        int rc = tagtree_decode(&pb->incl, cb->xi, cb->yi, pg->l + 1u, r, &v);
                        if (rc < 0) return -1;
                        inc = (v <= pg->l);
                    }
                    if (!inc) continue;
                    if (first) {
                        /* zero bit-plane tag tree, read to completion */
                        uint32_t thr = 1;
                        for (;;) {
                            int rc = tagtree_decode(&pb->zbp, cb->xi, cb->yi, thr, r, &v);
                            if (rc < 0) return -1;
                            if (rc) break;
                            thr++;
                            if (thr > 255) return -1;
                        }
                        cb->zblk = v;
                    }
                    cb->included = 1;
                    if (!(cp->cbstyle & ORACLE_CB_HT)) {
                        if (read_classic_block(r, cb, cp->cbstyle)) return -1;
                        continue;
                    }
                    if (!first || cb->passes != 0) return -1;
                    uint32_t raw;
                    if (read_ht_passes(r, &raw) || raw > 94) return -1;
                    uint32_t groups = raw > 0 ? (raw - 1) / 3 : 0;
                    cb->zblk += groups;
                    if (cb->zblk > 255) return -1;
                    uint32_t ph = groups * 3;
                    uint32_t actual = raw - ph;
                    uint32_t bit;
                    for (;;) {
                        if (pbr_bits(r, 1, &bit)) return -1;
                        if (!bit) break;
                        cb->lblock++;
                    }
                    uint32_t nb = cb->lblock + ilog2u(ph + 1);
                    if (nb > 32) return -1;
                    uint32_t lc;
                    if (pbr_bits(r, nb, &lc)) return -1;
                    if (lc < 2 || lc >= 65535) return -1;
                    uint32_t lr = 0;
                    if (actual > 1) {
                        uint32_t nb2 = cb->lblock + (actual > 2 ? 1 : 0);
                        if (nb2 > 32 || pbr_bits(r, nb2, &lr)) return -1;
                        if (lr >= 2047) return -1;
                    }
                    cb->passes = actual;
                    cb->lcup = lc;
                    cb->lref = lr;
                }
            }
        }
        /* align to a byte boundary; EPH follows the header */
        if (r->bitpos & 7) r->bitpos += 8 - (r->bitpos & 7);
        if (cp->scod & 4) {
            size_t bb = r->bitpos >> 3;
            if (bb + 2 > r->len || r->d[bb] != 0xFF || r->d[bb + 1] != 0x92) return -1;
            r->bitpos += 16;
        }
        if (!packed) body = hb.bitpos >> 3;
        if (nonzero) {
            for (int bi = 0; bi < rs->nbands; ++bi) {
                prec_band_t *pb = &rs->prec[(size_t)pg->p * rs->nbands + bi];
                for (uint32_t k = 0; k < pb->ncbx * pb->ncby; ++k) {
                    cblk_t *cb = &pb->cbs[k];
                    if (!(cp->cbstyle & ORACLE_CB_HT)) {
                        /* this packet's contributions: those without data yet */
                        for (uint32_t q = 0; q < cb->ncontrib; ++q) {
                            contrib_t *ct = &cb->contrib[q];
                            if (ct->data) continue;
                            if (body + ct->len > tp->body_len) return -1;
                            ct->data = tp->body + body;
                            body += ct->len;
                        }
                        continue;
                    }
                    if (cb->passes == 0 || cb->data) continue;
                    size_t need = (size_t)cb->lcup + cb->lref;
                    if (body + need > tp->body_len) return -1;
                    cb->data = tp->body + body;
                    body += need;
                }
            }
        }
        if (!packed) hb.bitpos = body * 8;
    }
    return 0;
}

static int parse_packets(tile_t *t, int ncomp, tcomp_t *tcs, const prog_list_t *pl)
{
    /* non-strict: a malformed tile-part ends its own packets only; the next
     * tile-part continues the progression (segment.rs:36-50) */
    size_t next = 0;
    int rc = 0;
    for (size_t i = 0; i < t->nparts; ++i)
        if (parse_tile_part(&t->parts[i], &next, ncomp, tcs, t->comp, pl)) rc = -1;
    return rc;
}

/* ------------------------------------------------------------------ */
/* Sub-band decode (decode.rs:1117-1187,1554-1718)                      */

static int exponent_mantissa(const comp_t *c, int type, uint32_t r, uint32_t *eps, uint32_t *mu)
{
    if (c->qstyle == 1) {
        uint32_t nb = r == 0 ? c->nlevels : c->nlevels + 1 - r;
        int e = (int)c->sexp[0] - (int)c->nlevels;
        if (e < 0) return -1;
        *eps = (uint32_t)e + nb;
        *mu = c->smant[0];
        return 0;
    }
    uint32_t idx = r == 0 ? 0 : 1 + (r - 1) * 3 + (uint32_t)(type - 1);
    if ((int)idx >= c->nsteps) return -1;
    *eps = c->sexp[idx];
    *mu = c->smant[idx];
    return 0;
}

static float pow2i_f32(int e)
{
    /* math.rs:550-556 */
    if (e >= 0) return (float)(1u << e);
    return 1.0f / (float)(1u << (-e));
}

/* Classic code-blocks of one band (decode.rs:1189-1343, bitplane.rs:90-186):
 * each block's contributions joined in codestream order; with selective
 * bypass or termination on each pass the joined bytes split into codeword
 * segments where the segment index changes, the passes of a segment
 * following on from the previous one; a raw segment is one that starts
 * after pass 9 on a significance-propagation or refinement pass. */
static int decode_band_classic(const comp_t *c, res_t *rs, int bi, uint32_t K, float delta,
                               double delta64, int use_f64)
{
    band_t *bd = &rs->bands[bi];
    const uint32_t w = rw(bd->rect);
    const int style = c->cbstyle;
    const uint32_t np = rs->npx * rs->npy;
    int rc = ORACLE_OK;
    if (K > 32) return ORACLE_E_MALFORMED;
    for (uint32_t p = 0; p < np && !rc; ++p) {
        prec_band_t *pb = &rs->prec[(size_t)p * rs->nbands + bi];
        for (uint32_t k = 0; k < pb->ncbx * pb->ncby && !rc; ++k) {
            cblk_t *cb = &pb->cbs[k];
            const uint32_t cw = rw(cb->rect), ch = rh(cb->rect);
            const uint32_t planes = K > cb->zblk ? K - cb->zblk : 0;
            if (cw == 0 || ch == 0 || cb->passes == 0 || planes == 0) continue;
            size_t total = 0;
            for (uint32_t q = 0; q < cb->ncontrib; ++q) total += cb->contrib[q].data ? cb->contrib[q].len : 0;
            uint8_t *joined = (uint8_t *)malloc(total + 1);
            oracle_seg_t *segs = (oracle_seg_t *)calloc(cb->ncontrib + 1, sizeof(oracle_seg_t));
            int32_t *v = (int32_t *)malloc((size_t)cw * ch * sizeof(int32_t));
            if (!joined || !segs || !v) {
                free(joined);
                free(segs);
                free(v);
                return ORACLE_E_NOMEM;
            }
            size_t at = 0, nseg = 0;
            uint32_t passes = 0;
            for (uint32_t q = 0; q < cb->ncontrib; ++q) {
                const contrib_t *ct = &cb->contrib[q];
                if (q == 0 || ct->seg != cb->contrib[q - 1].seg) {
                    oracle_seg_t *sg = &segs[nseg++];
                    sg->off = (uint32_t)at;
                    sg->start = (uint8_t)passes;
                    sg->arith = !(style & ORACLE_CB_BYPASS) || passes <= 9 || passes % 3 == 0;
                }
                oracle_seg_t *sg = &segs[nseg - 1];
                if (ct->data) {
                    memcpy(joined + at, ct->data, ct->len);
                    at += ct->len;
                    sg->len += ct->len;
                }
                passes += ct->passes;
                sg->end = (uint8_t)passes;
            }
            st_blocks++;
            if (cb->passes > st_max_passes) st_max_passes = cb->passes;
            rc = oracle_classic_decode_block(joined, at, segs, nseg, cw, ch, planes, cb->passes,
                                             bd->type, style, v, cw);
            if (!rc) {
                const uint32_t ox = cb->rect.x0 - bd->rect.x0, oy = cb->rect.y0 - bd->rect.y0;
                for (uint32_t y = 0; y < ch; ++y)
                    for (uint32_t x = 0; x < cw; ++x) {
                        const int32_t q = v[(size_t)y * cw + x];
                        const size_t o = (size_t)(oy + y) * w + ox + x;
                        if (use_f64)
                            bd->coef64[o] = (double)q * delta64;
                        else
                            bd->coef[o] = (float)q * delta;
                    }
            }
            free(joined);
            free(segs);
            free(v);
        }
    }
    return rc;
}

static int decode_band(const comp_t *c, res_t *rs, int bi, uint32_t r, int use_f64)
{
    band_t *bd = &rs->bands[bi];
    uint32_t w = rw(bd->rect), h = rh(bd->rect);
    size_t n = (size_t)w * h;
    if (use_f64) {
        bd->coef64 = (double *)calloc(n ? n : 1, sizeof(double));
        if (!bd->coef64) return ORACLE_E_NOMEM;
    } else {
        bd->coef = (float *)calloc(n ? n : 1, sizeof(float));
        if (!bd->coef) return ORACLE_E_NOMEM;
    }
    uint32_t eps, mu;
    if (exponent_mantissa(c, bd->type, r, &eps, &mu)) return ORACLE_E_MALFORMED;
    float delta = 1.0f;
    double delta64 = 1.0;
    if (c->qstyle != 0) {
        int gain = bd->type == 0 ? 0 : (bd->type == 3 ? 2 : 1);
        int rb = (int)c->precision + gain;
        delta = pow2i_f32(rb - (int)eps) * (1.0f + (float)mu / 2048.0f);
        delta64 = ldexp(1.0, rb - (int)eps) * (1.0 + (double)mu / 2048.0);
    }
    if ((int)c->guard + (int)eps - 1 < 0) return ORACLE_E_MALFORMED;
    uint32_t K = c->guard + eps - 1;
    if (K > 32) return ORACLE_E_MALFORMED;
    int stripe_causal = (c->cbstyle & 0x08) != 0;
    uint32_t np = rs->npx * rs->npy;
    uint32_t *buf = NULL;
    size_t bufcap = 0;
    if (!(c->cbstyle & ORACLE_CB_HT)) return decode_band_classic(c, rs, bi, K, delta, delta64, use_f64);
    for (uint32_t p = 0; p < np; ++p) {
        prec_band_t *pb = &rs->prec[(size_t)p * rs->nbands + bi];
        for (uint32_t k = 0; k < pb->ncbx * pb->ncby; ++k) {
            cblk_t *cb = &pb->cbs[k];
            uint32_t cw = rw(cb->rect), ch = rh(cb->rect);
            if (K == 0) continue;
            if (K > 31) {
                free(buf);
                return ORACLE_E_MALFORMED;
            }
            uint32_t actual = K > cb->zblk ? K - cb->zblk : 0;
            if (cb->passes == 0 || actual == 0 || !cb->data) continue;
            st_blocks++;
            if (cb->lref) st_ref_blocks++;
            if (cb->passes > st_max_passes) st_max_passes = cb->passes;
            size_t need = (size_t)cw * ch;
            if (need > bufcap) {
                free(buf);
                buf = (uint32_t *)malloc(need * sizeof(uint32_t));
                bufcap = need;
                if (!buf) return ORACLE_E_NOMEM;
            }
            const uint8_t *src = cb->data;
            uint8_t *zcopy = NULL;
            if (zero_refinement && cb->lref) {
                zcopy = (uint8_t *)malloc((size_t)cb->lcup + cb->lref);
                if (!zcopy) {
                    free(buf);
                    return ORACLE_E_NOMEM;
                }
                memcpy(zcopy, cb->data, cb->lcup);
                memset(zcopy + cb->lcup, 0, cb->lref);
                src = zcopy;
            }
            int rc = oracle_ht_decode_block(src, cb->lcup, cb->lref, cw, ch, cb->zblk,
                                            cb->passes, stripe_causal, 2, buf, cw);
            free(zcopy);
            if (rc) {
                free(buf);
                return rc;
            }
            uint32_t ox = cb->rect.x0 - bd->rect.x0, oy = cb->rect.y0 - bd->rect.y0;
            for (uint32_t y = 0; y < ch; ++y)
                for (uint32_t x = 0; x < cw; ++x) {
                    int32_t v = oracle_coefficient_to_i32(buf[(size_t)y * cw + x], K);
                    size_t o = (size_t)(oy + y) * w + ox + x;
                    if (use_f64)
                        bd->coef64[o] = (double)v * delta64;
                    else
                        bd->coef[o] = (float)v * delta;
                }
        }
    }
    free(buf);
    return ORACLE_OK;
}

/* ------------------------------------------------------------------ */
/* Inverse DWT (idwt.rs:535-1143)                                        */

#define KAPPA_F 1.2301741f
static const float NEG_ALPHA_F = 1.5861343f;
static const float NEG_BETA_F = 0.052980117f;
static const float NEG_GAMMA_F = -0.8829111f;
static const float NEG_DELTA_F = -0.44350687f;

/* Textbook F.3.8.2 constants for the double-precision oracle. */
static const double ALPHA_D = -1.586134342059924;
static const double BETA_D = -0.052980118572961;
static const double GAMMA_D = 0.882911075530934;
static const double DELTA_D = 0.443506852043971;
static const double KAPPA_D = 1.230174104914001;

static size_t ext_left(size_t i) { return i >= 1 ? i - 1 : 1 - i; }
static size_t ext_right(size_t i, size_t n) { return i + 1 < n ? i + 1 : n - 2 - (i + 1 - n); }

/* 1D_SR over `n` samples spaced `step` apart; x0 is the absolute origin. */
static void sr_f32(float *s, size_t n, size_t step, uint32_t x0, int rev)
{
#define S(i) s[(size_t)(i) * step]
    if (n == 1) {
        if (x0 & 1) S(0) *= 0.5f;
        return;
    }
    size_t fe = x0 & 1, fo = 1 - fe;
    if (rev) {
        for (size_t i = fe; i < n; i += 2)
            S(i) = S(i) - floorf(fmaf(S(ext_left(i)) + S(ext_right(i, n)), 0.25f, 0.5f));
        for (size_t i = fo; i < n; i += 2)
            S(i) = S(i) + floorf((S(ext_left(i)) + S(ext_right(i, n))) * 0.5f);
        return;
    }
    const float inv_k = 1.0f / KAPPA_F;
    float k0 = fe == 0 ? KAPPA_F : inv_k, k1 = fe == 0 ? inv_k : KAPPA_F;
    for (size_t i = 0; i + 1 < n; i += 2) {
        S(i) *= k0;
        S(i + 1) *= k1;
    }
    if (n & 1) S(n - 1) *= k0;
    for (size_t i = fe; i < n; i += 2)
        S(i) = fmaf(S(ext_left(i)) + S(ext_right(i, n)), NEG_DELTA_F, S(i));
    for (size_t i = fo; i < n; i += 2)
        S(i) = fmaf(S(ext_left(i)) + S(ext_right(i, n)), NEG_GAMMA_F, S(i));
    for (size_t i = fe; i < n; i += 2)
        S(i) = fmaf(S(ext_left(i)) + S(ext_right(i, n)), NEG_BETA_F, S(i));
    for (size_t i = fo; i < n; i += 2)
        S(i) = fmaf(S(ext_left(i)) + S(ext_right(i, n)), NEG_ALPHA_F, S(i));
#undef S
}

static void sr_f64(double *s, size_t n, size_t step, uint32_t x0, int rev)
{
#define S(i) s[(size_t)(i) * step]
    if (n == 1) {
        if (x0 & 1) S(0) *= 0.5;
        return;
    }
    size_t fe = x0 & 1, fo = 1 - fe;
    if (rev) {
        for (size_t i = fe; i < n; i += 2)
            S(i) = S(i) - floor((S(ext_left(i)) + S(ext_right(i, n))) * 0.25 + 0.5);
        for (size_t i = fo; i < n; i += 2)
            S(i) = S(i) + floor((S(ext_left(i)) + S(ext_right(i, n))) * 0.5);
        return;
    }
    double k0 = fe == 0 ? KAPPA_D : 1.0 / KAPPA_D, k1 = fe == 0 ? 1.0 / KAPPA_D : KAPPA_D;
    for (size_t i = 0; i + 1 < n; i += 2) {
        S(i) *= k0;
        S(i + 1) *= k1;
    }
    if (n & 1) S(n - 1) *= k0;
    for (size_t i = fe; i < n; i += 2) S(i) -= DELTA_D * (S(ext_left(i)) + S(ext_right(i, n)));
    for (size_t i = fo; i < n; i += 2) S(i) -= GAMMA_D * (S(ext_left(i)) + S(ext_right(i, n)));
    for (size_t i = fe; i < n; i += 2) S(i) -= BETA_D * (S(ext_left(i)) + S(ext_right(i, n)));
    for (size_t i = fo; i < n; i += 2) S(i) -= ALPHA_D * (S(ext_left(i)) + S(ext_right(i, n)));
#undef S
}

/* One level of 2D_SR: interleave LL/HL/LH/HH (F.3.3) then rows, then
 * columns.  `ll` has the rect of resolution r-1; output has resolution r. */
#define DEFINE_LEVEL(NAME, T, SR, FIELD)                                                       \
    static T *NAME(const T *ll, const res_t *rs, int rev)                                    \
    {                                                                                          \
        rect_t o = rs->rect;                                                                   \
        size_t w = rw(o), h = rh(o);                                                           \
        T *out = (T *)calloc((w && h) ? w * h : 1, sizeof(T));                                    \
        if (!out) return NULL;                                                                 \
        uint32_t lx0 = (o.x0 + 1) / 2, hx0 = o.x0 / 2, ly0 = (o.y0 + 1) / 2, hy0 = o.y0 / 2;   \
        size_t wl = (o.x1 + 1) / 2 - lx0, wh = o.x1 / 2 - hx0;                                 \
        const T *hl = rs->bands[0].FIELD, *lh = rs->bands[1].FIELD, *hh = rs->bands[2].FIELD;  \
        for (size_t y = 0; y < h; ++y) {                                                       \
            uint32_t ay = o.y0 + (uint32_t)y;                                                  \
            for (size_t x = 0; x < w; ++x) {                                                   \
                uint32_t ax = o.x0 + (uint32_t)x;                                              \
                T v;                                                                           \
                if (!(ay & 1)) {                                                               \
                    size_t row = ay / 2 - ly0;                                                 \
                    v = !(ax & 1) ? ll[row * wl + (ax / 2 - lx0)] : hl[row * wh + (ax / 2 - hx0)]; \
                } else {                                                                       \
                    size_t row = ay / 2 - hy0;                                                 \
                    v = !(ax & 1) ? lh[row * wl + (ax / 2 - lx0)] : hh[row * wh + (ax / 2 - hx0)]; \
                }                                                                              \
                out[y * w + x] = v;                                                            \
            }                                                                                  \
        }                                                                                      \
        if (w > 0 && h > 0) {                                                                  \
            for (size_t y = 0; y < h; ++y) SR(out + y * w, w, 1, o.x0, rev);                   \
            for (size_t x = 0; x < w; ++x) SR(out + x, h, w, o.y0, rev);                       \
        }                                                                                      \
        return out;                                                                            \
    }

DEFINE_LEVEL(idwt_level_f32, float, sr_f32, coef)
DEFINE_LEVEL(idwt_level_f64, double, sr_f64, coef64)

/* ------------------------------------------------------------------ */
/* Full decode to planes                                                 */

typedef struct {
    uint32_t w, h, ncomp;
    double *planes64;
    float *planes;
    uint8_t precision[MAXC];
    int mct, reversible;
} image_t;

/* Shrink factors of the output (codestream.rs:663-683): the components'
 * sub-sampling when every component has the same, else 1. */
static void shrink_of(const header_t *h, uint32_t *sx, uint32_t *sy)
{
    int same = 1;
    for (int c = 1; c < h->ncomp; ++c)
        same &= h->comp[c].xr == h->comp[0].xr && h->comp[c].yr == h->comp[0].yr;
    *sx = same ? h->comp[0].xr : 1;
    *sy = same ? h->comp[0].yr : 1;
}

static int any_subsampled(const header_t *h)
{
    for (int c = 0; c < h->ncomp; ++c)
        if (h->comp[c].xr != 1 || h->comp[c].yr != 1) return 1;
    return 0;
}

static int decode_image(const uint8_t *data, size_t len, uint32_t denominator, int use_f64,
                        image_t *img)
{
    oracle_build_tables();
    st_blocks = st_ref_blocks = st_max_passes = 0;
    memset(img, 0, sizeof(*img));
    header_t *h = (header_t *)calloc(1, sizeof(header_t));
    if (!h) return ORACLE_E_NOMEM;
    uint32_t tw = 0, th = 0;
    /* target = ceil(dims / denominator) (crates/signinum-j2k/src/decode.rs:22-25) */
    int rc;
    {
        header_t probe;
        rc = read_main_header(data, len, &probe, 0, 0);
        if (rc) {
            free(h);
            return rc;
        }
        if (denominator > 1) {
            uint32_t sx, sy;
            shrink_of(&probe, &sx, &sy);
            tw = cdiv(cdiv(probe.xsiz - probe.xo, sx), denominator);
            th = cdiv(cdiv(probe.ysiz - probe.yo, sy), denominator);
        }
    }
    rc = read_main_header(data, len, h, tw, th);
    if (rc) {
        free(h);
        return rc;
    }
    /* the store of a reduced resolution indexes sub-sampled components
     * with full-resolution offsets (decode.rs:1858-1898): full resolution
     * only */
    if (h->skipped && any_subsampled(h)) {
        free(h);
        return ORACLE_E_UNSUPPORTED;
    }
    uint32_t shx, shy;
    shrink_of(h, &shx, &shy);
    uint32_t ntx = cdiv(h->xsiz - h->xto, h->xt), nty = cdiv(h->ysiz - h->yto, h->yt);
    uint32_t ntiles = ntx * nty;
    tile_t *tiles = (tile_t *)calloc(ntiles, sizeof(tile_t));
    tcomp_t *tcs = (tcomp_t *)calloc(h->ncomp, sizeof(tcomp_t));
    if (!tiles || !tcs) {
        free(h);
        free(tiles);
        free(tcs);
        return ORACLE_E_NOMEM;
    }
    rc = parse_tiles(data, len, h, tiles, ntiles);
    uint32_t sk = h->skipped;
    /* image_width / image_height (codestream.rs:544-553); a non-zero image
     * origin only at full resolution (read_main_header) */
    uint32_t iw = cdiv(h->xsiz - h->xo, shx << sk), ih = cdiv(h->ysiz - h->yo, shy << sk);
    img->w = iw;
    img->h = ih;
    img->ncomp = h->ncomp;
    img->mct = tiles[0].mct;
    img->reversible = h->comp[0].transform == 1;
    for (int c = 0; c < h->ncomp; ++c) img->precision[c] = h->comp[c].precision;
    size_t plane = (size_t)iw * ih;
    if (!rc) {
        if (use_f64)
            img->planes64 = (double *)calloc(plane * h->ncomp + 1, sizeof(double));
        else
            img->planes = (float *)calloc(plane * h->ncomp + 1, sizeof(float));
        if (!img->planes64 && !img->planes) rc = ORACLE_E_NOMEM;
    }
    for (uint32_t ti = 0; ti < ntiles && !rc; ++ti) {
        tile_t *t = &tiles[ti];
        /* several quality layers: classic code-blocks only (an HT block
         * included twice is malformed in the packet parser) */
        for (int c = 0; c < h->ncomp; ++c)
            if ((t->comp[c].cbstyle & ORACLE_CB_HT) && t->nlayers > 1) rc = ORACLE_E_UNSUPPORTED;
        if (rc) break;
        for (int c = 0; c < h->ncomp && !rc; ++c) rc = build_tcomp(h, &t->comp[c], t->rect, &tcs[c]);
        if (!rc) {
            for (int c = 0; c < h->ncomp; ++c)
                if (tcs[c].nres <= (int)sk) rc = ORACLE_E_MALFORMED;
        }
        if (!rc) {
            prog_list_t pl = {0};
            rc = build_progression(t, h->ncomp, tcs, &pl);
            if (!rc) parse_packets(t, h->ncomp, tcs, &pl); /* non-strict: errors stop parsing */
            free(pl.v);
        }
        for (int c = 0; c < h->ncomp && !rc; ++c) {
            tcomp_t *tc = &tcs[c];
            int nr = tc->nres - (int)sk;
            for (int r = 0; r < nr && !rc; ++r)
                for (int bi = 0; bi < tc->res[r].nbands && !rc; ++bi)
                    rc = decode_band(&t->comp[c], &tc->res[r], bi, (uint32_t)r, use_f64);
            if (rc) break;
            /* IDWT over the retained levels */
            int rev = t->comp[c].transform == 1;
            void *cur = use_f64 ? (void *)tc->res[0].bands[0].coef64 : (void *)tc->res[0].bands[0].coef;
            int own = 0;
            for (int r = 1; r < nr; ++r) {
                void *nx = use_f64 ? (void *)idwt_level_f64((const double *)cur, &tc->res[r], rev)
                                   : (void *)idwt_level_f32((const float *)cur, &tc->res[r], rev);
                if (own) free(cur);
                if (!nx) {
                    rc = ORACLE_E_NOMEM;
                    cur = NULL;
                    break;
                }
                cur = nx;
                own = 1;
            }
            if (rc) break;
            /* store (decode.rs:1731-1850): copy the resolution rect into
             * the image plane; add the sign shift unless MCT follows. */
            rect_t rr = tc->res[nr - 1].rect;
            float shift = t->mct ? 0.0f : (float)(1u << (t->comp[c].precision - 1));
            if (tc->xr != 1 || tc->yr != 1 || shx != 1 || shy != 1) {
                /* sample by sample (decode.rs:1858-1898): sample x covers
                 * [XRsiz x / shrink, + XRsiz) of the output grid, later
                 * samples overwriting earlier ones */
                uint32_t xoff = cdiv(h->xo, shx), yoff = cdiv(h->yo, shy);
                for (uint32_t y = rr.y0; y < rr.y1; ++y) {
                    uint32_t ry = (tc->yr * y) / shy;
                    for (uint32_t x = rr.x0; x < rr.x1; ++x) {
                        uint32_t rx = (tc->xr * x) / shx;
                        size_t src = (size_t)(y - rr.y0) * rw(rr) + (x - rr.x0);
                        double v = use_f64 ? ((double *)cur)[src] + (double)shift
                                           : (double)(((float *)cur)[src] + shift);
                        uint32_t xa = rx > xoff ? rx : xoff, xb = rx + tc->xr;
                        uint32_t ya = ry > yoff ? ry : yoff, yb = ry + tc->yr;
                        if (xb > iw + xoff) xb = iw + xoff;
                        if (yb > ih + yoff) yb = ih + yoff;
                        for (uint32_t xp = xa; xp < xb; ++xp)
                            for (uint32_t yp = ya; yp < yb; ++yp) {
                                size_t dst = (size_t)c * plane + (size_t)(yp - yoff) * iw + (xp - xoff);
                                if (use_f64)
                                    img->planes64[dst] = v;
                                else
                                    img->planes[dst] = (float)v;
                            }
                    }
                }
                if (own) free(cur);
                continue;
            }
            for (uint32_t y = 0; y < rh(rr); ++y)
                for (uint32_t x = 0; x < rw(rr); ++x) {
                    size_t src = (size_t)y * rw(rr) + x;
                    size_t dst = (size_t)c * plane + (size_t)(rr.y0 + y - h->yo) * iw +
                                 (rr.x0 + x - h->xo);
                    if (use_f64)
                        img->planes64[dst] = ((double *)cur)[src] + (double)shift;
                    else
                        img->planes[dst] = ((float *)cur)[src] + shift;
                }
            if (own) free(cur);
        }
        for (int c = 0; c < h->ncomp; ++c) free_tcomp(&tcs[c]);
    }
    /* MCT + sign shift (decode.rs:114-119; mct.rs:11-156); components of
     * different sub-sampling are not transformed together here */
    if (!rc && tiles[0].mct && any_subsampled(h) && (shx == 1 && shy == 1)) rc = ORACLE_E_UNSUPPORTED;
    if (!rc && tiles[0].mct) {
        if (h->ncomp >= 3) {
            int tr = h->comp[0].transform;
            if (tr != h->comp[1].transform || tr != h->comp[2].transform)
                rc = ORACLE_E_MALFORMED;
            else {
                for (size_t i = 0; i < plane; ++i) {
                    if (use_f64) {
                        double y0 = img->planes64[i], y1 = img->planes64[plane + i],
                               y2 = img->planes64[2 * plane + i];
                        if (tr == 1) {
                            double i1 = y0 - floor((y2 + y1) * 0.25);
                            img->planes64[i] = y2 + i1;
                            img->planes64[plane + i] = i1;
                            img->planes64[2 * plane + i] = y1 + i1;
                        } else {
                            img->planes64[i] = y0 + 1.402 * y2;
                            img->planes64[plane + i] = y0 - 0.34413 * y1 - 0.71414 * y2;
                            img->planes64[2 * plane + i] = y0 + 1.772 * y1;
                        }
                    } else {
                        float y0 = img->planes[i], y1 = img->planes[plane + i],
                              y2 = img->planes[2 * plane + i];
                        if (tr == 1) {
                            float i1 = y0 - floorf((y2 + y1) * 0.25f);
                            img->planes[i] = y2 + i1;
                            img->planes[plane + i] = i1;
                            img->planes[2 * plane + i] = y1 + i1;
                        } else {
                            img->planes[i] = fmaf(y2, 1.402f, y0);
                            img->planes[plane + i] = fmaf(y2, -0.71414f, fmaf(y1, -0.34413f, y0));
                            img->planes[2 * plane + i] = fmaf(y1, 1.772f, y0);
                        }
                    }
                }
            }
        }
        if (!rc) {
            for (int c = 0; c < h->ncomp; ++c) {
                float a = (float)(1u << (h->comp[c].precision - 1));
                for (size_t i = 0; i < plane; ++i) {
                    if (use_f64)
                        img->planes64[(size_t)c * plane + i] += (double)a;
                    else
                        img->planes[(size_t)c * plane + i] += a;
                }
            }
        }
    }
    for (uint32_t ti = 0; ti < ntiles; ++ti) free(tiles[ti].parts);
    free(tiles);
    free(tcs);
    free(h);
    if (rc) {
        free(img->planes);
        free(img->planes64);
        img->planes = NULL;
        img->planes64 = NULL;
    }
    return rc;
}

/* ------------------------------------------------------------------ */
/* Output conversion (lib.rs:1772-1835,2202-2295; signinum-j2k decode.rs) */

static uint8_t sat_u8(float v)
{
    /* Rust `as u8`: NaN -> 0, saturating */
    if (!(v == v)) return 0;
    if (v <= 0.0f) return 0;
    if (v >= 255.0f) return 255;
    return (uint8_t)v;
}

static int fmt_channels(int fmt)
{
    switch (fmt) {
    case ORACLE_GRAY8: case ORACLE_GRAY16: return 1;
    case ORACLE_RGB8: case ORACLE_RGB16: return 3;
    case ORACLE_RGBA8: return 4;
    default: return 0;
    }
}

static int fmt_bps(int fmt) { return (fmt == ORACLE_GRAY16 || fmt == ORACLE_RGB16) ? 2 : 1; }

static int convert_output(const image_t *img, int fmt, uint8_t *out, size_t out_len, size_t stride)
{
    int ch = fmt_channels(fmt);
    if (!ch) return ORACLE_E_ARG;
    size_t row_bytes = (size_t)img->w * ch * fmt_bps(fmt);
    if (stride < row_bytes) return ORACLE_E_BUFFER;
    if (img->h && out_len < stride * (img->h - 1) + row_bytes) return ORACLE_E_BUFFER;
    int gray = img->ncomp == 1, rgb = img->ncomp == 3;
    if ((ch == 1 && !gray) || (ch >= 3 && !rgb)) return ORACLE_E_UNSUPPORTED;
    size_t plane = (size_t)img->w * img->h;
    uint32_t bd = img->precision[0];
    int same8 = 1;
    for (uint32_t c = 0; c < img->ncomp; ++c) same8 &= img->precision[c] == 8;
    for (uint32_t y = 0; y < img->h; ++y) {
        uint8_t *row = out + (size_t)y * stride;
        for (uint32_t x = 0; x < img->w; ++x) {
            size_t i = (size_t)y * img->w + x;
            for (uint32_t c = 0; c < img->ncomp; ++c) {
                float v = img->planes[(size_t)c * plane + i];
                if (fmt_bps(fmt) == 1) {
                    uint8_t o;
                    if (same8) {
                        o = sat_u8(roundf(v));
                    } else {
                        float mx = (float)((1u << img->precision[c]) - 1);
                        o = sat_u8(roundf((v / mx) * 255.0f));
                    }
                    row[(size_t)x * ch + c] = o;
                } else {
                    /* decode_native: round, clamp to [0, 2^bd - 1] */
                    float mx = (float)(bd >= 32 ? 4294967295.0 : (double)((1ull << bd) - 1));
                    float r = roundf(v);
                    float cl = r < 0.0f ? 0.0f : (r > mx ? mx : r);
                    uint16_t s;
                    if (bd <= 8) {
                        uint32_t s8 = (uint32_t)cl;
                        uint32_t m8 = (1u << bd) - 1;
                        if (m8 == 0) m8 = 1;
                        s = (uint16_t)((s8 * 65535u + m8 / 2) / m8);
                    } else {
                        s = (uint16_t)(cl > 65535.0f ? 65535.0f : cl);
                    }
                    row[((size_t)x * ch + c) * 2] = (uint8_t)(s & 0xFF);
                    row[((size_t)x * ch + c) * 2 + 1] = (uint8_t)(s >> 8);
                }
            }
            if (ch == 4) row[(size_t)x * 4 + 3] = 255;
        }
    }
    return ORACLE_OK;
}

/* ------------------------------------------------------------------ */
/* Public API                                                            */

int oracle_inspect(const uint8_t *data, size_t len, oracle_info_t *info)
{
    header_t *h = (header_t *)calloc(1, sizeof(header_t));
    if (!h) return ORACLE_E_NOMEM;
    int rc = read_main_header(data, len, h, 0, 0);
    if (!rc) {
        memset(info, 0, sizeof(*info));
        uint32_t sx, sy;
        shrink_of(h, &sx, &sy);
        info->width = cdiv(h->xsiz - h->xo, sx);
        info->height = cdiv(h->ysiz - h->yo, sy);
        info->num_components = h->ncomp;
        info->bit_depth = h->comp[0].precision;
        uint32_t nl = 255;
        for (int c = 0; c < h->ncomp; ++c)
            if (h->comp[c].nlevels < nl) nl = h->comp[c].nlevels;
        info->num_levels = nl;
        info->reversible = h->comp[0].transform == 1;
        info->mct = h->mct;
        info->high_throughput = (h->comp[0].cbstyle & 0x40) != 0;
        info->tiles_x = cdiv(h->xsiz - h->xto, h->xt);
        info->tiles_y = cdiv(h->ysiz - h->yto, h->yt);
        info->cb_width = 1u << h->comp[0].xcb;
        info->cb_height = 1u << h->comp[0].ycb;
        info->progression = h->prog;
    }
    free(h);
    return rc;
}

int oracle_scaled_dims(const uint8_t *data, size_t len, uint32_t denominator, uint32_t *out_w,
                       uint32_t *out_h)
{
    header_t *h = (header_t *)calloc(1, sizeof(header_t));
    if (!h) return ORACLE_E_NOMEM;
    int rc = read_main_header(data, len, h, 0, 0);
    if (!rc && denominator > 1) {
        uint32_t sx, sy;
        shrink_of(h, &sx, &sy);
        uint32_t tw = cdiv(cdiv(h->xsiz - h->xo, sx), denominator);
        uint32_t th = cdiv(cdiv(h->ysiz - h->yo, sy), denominator);
        rc = read_main_header(data, len, h, tw, th);
    }
    if (!rc && h->skipped && any_subsampled(h)) rc = ORACLE_E_UNSUPPORTED;
    if (!rc) {
        uint32_t sx, sy;
        shrink_of(h, &sx, &sy);
        *out_w = cdiv(h->xsiz - h->xo, sx << h->skipped);
        *out_h = cdiv(h->ysiz - h->yo, sy << h->skipped);
    }
    free(h);
    return rc;
}

int oracle_decode(const uint8_t *data, size_t len, int fmt, uint32_t denominator, uint8_t *out,
                  size_t out_len, size_t stride)
{
    if (!data || !out) return ORACLE_E_ARG;
    if (denominator != 1 && denominator != 2 && denominator != 4 && denominator != 8)
        return ORACLE_E_ARG;
    image_t img;
    int rc = decode_image(data, len, denominator, 0, &img);
    if (rc) return rc;
    rc = convert_output(&img, fmt, out, out_len, stride);
    free(img.planes);
    return rc;
}

int oracle_decode_planes(const uint8_t *data, size_t len, uint32_t denominator, int use_f64,
                         double *planes, size_t plane_len)
{
    image_t img;
    int rc = decode_image(data, len, denominator, use_f64, &img);
    if (rc) return rc;
    size_t n = (size_t)img.w * img.h * img.ncomp;
    if (plane_len < n)
        rc = ORACLE_E_BUFFER;
    else
        for (size_t i = 0; i < n; ++i) planes[i] = use_f64 ? img.planes64[i] : (double)img.planes[i];
    free(img.planes);
    free(img.planes64);
    return rc;
}

typedef struct {
    const uint8_t *const *datas;
    const size_t *lens;
    uint8_t *const *outs;
    const size_t *out_lens, *strides;
    size_t begin, end;
    int fmt;
    uint32_t den;
    int status;
    size_t fail_index;
} batch_job_t;

static void *batch_worker(void *arg)
{
    batch_job_t *j = (batch_job_t *)arg;
    j->status = ORACLE_OK;
    for (size_t i = j->begin; i < j->end; ++i) {
        int rc = oracle_decode(j->datas[i], j->lens[i], j->fmt, j->den, j->outs[i], j->out_lens[i],
                               j->strides[i]);
        if (rc && j->status == ORACLE_OK) {
            j->status = rc;
            j->fail_index = i;
        }
    }
    return NULL;
}

int oracle_decode_batch(const uint8_t *const *datas, const size_t *lens, size_t n, int fmt,
                        uint32_t denominator, uint8_t *const *outs, const size_t *out_lens,
                        const size_t *strides, int workers, size_t *first_fail_index)
{
    oracle_build_tables();
    if (n == 0) return ORACLE_OK;
    if (workers < 1) workers = 1;
    if ((size_t)workers > n) workers = (int)n;
    size_t chunk = (n + (size_t)workers - 1) / (size_t)workers;
    batch_job_t *jobs = (batch_job_t *)calloc((size_t)workers, sizeof(batch_job_t));
    pthread_t *th = (pthread_t *)calloc((size_t)workers, sizeof(pthread_t));
    if (!jobs || !th) {
        free(jobs);
        free(th);
        return ORACLE_E_NOMEM;
    }
    int nj = 0;
    for (size_t b = 0; b < n; b += chunk, ++nj) {
        batch_job_t *j = &jobs[nj];
        j->datas = datas;
        j->lens = lens;
        j->outs = outs;
        j->out_lens = out_lens;
        j->strides = strides;
        j->begin = b;
        j->end = b + chunk < n ? b + chunk : n;
        j->fmt = fmt;
        j->den = denominator;
        if (nj == 0 || workers == 1)
            continue;
        pthread_create(&th[nj], NULL, batch_worker, j);
    }
    if (workers == 1) {
        for (int k = 0; k < nj; ++k) batch_worker(&jobs[k]);
    } else {
        batch_worker(&jobs[0]);
        for (int k = 1; k < nj; ++k) pthread_join(th[k], NULL);
    }
    int rc = ORACLE_OK;
    for (int k = 0; k < nj; ++k) {
        if (jobs[k].status != ORACLE_OK) {
            rc = jobs[k].status;
            if (first_fail_index) *first_fail_index = jobs[k].fail_index;
            break;
        }
    }
    free(jobs);
    free(th);
    return rc;
}

/* ------------------------------------------------------------------ */
/* Stage entry points for hook parity tests                             */

int oracle_idwt_level(int reversible, uint32_t x0, uint32_t y0, uint32_t x1, uint32_t y1,
                      const float *ll, const float *hl, const float *lh, const float *hh,
                      float *out)
{
    res_t rs;
    memset(&rs, 0, sizeof(rs));
    rs.rect.x0 = x0;
    rs.rect.y0 = y0;
    rs.rect.x1 = x1;
    rs.rect.y1 = y1;
    rs.nbands = 3;
    rs.bands[0].coef = (float *)hl;
    rs.bands[1].coef = (float *)lh;
    rs.bands[2].coef = (float *)hh;
    float *r = idwt_level_f32(ll, &rs, reversible);
    if (!r) return ORACLE_E_NOMEM;
    memcpy(out, r, (size_t)rw(rs.rect) * rh(rs.rect) * sizeof(float));
    free(r);
    return ORACLE_OK;
}

int oracle_inverse_mct(int reversible, float *p0, float *p1, float *p2, size_t n, float a0,
                       float a1, float a2)
{
    for (size_t i = 0; i < n; ++i) {
        float y0 = p0[i], y1 = p1[i], y2 = p2[i];
        if (reversible) {
            float i1 = y0 - floorf((y2 + y1) * 0.25f);
            p0[i] = y2 + i1;
            p1[i] = i1;
            p2[i] = y1 + i1;
        } else {
            p0[i] = fmaf(y2, 1.402f, y0);
            p1[i] = fmaf(y2, -0.71414f, fmaf(y1, -0.34413f, y0));
            p2[i] = fmaf(y1, 1.772f, y0);
        }
        p0[i] += a0;
        p1[i] += a1;
        p2[i] += a2;
    }
    return ORACLE_OK;
}

/* Code-block listing of a full-resolution decode, in geometry order
 * (component, resolution, precinct, band, block): rect in band coordinates,
 * data offset within the codestream (or -1), lengths, Zblk and passes. */
int oracle_code_blocks(const uint8_t *data, size_t len, oracle_cb_t *out, size_t cap,
                       size_t *count)
{
    oracle_build_tables();
    header_t *h = (header_t *)calloc(1, sizeof(header_t));
    if (!h) return ORACLE_E_NOMEM;
    int rc = read_main_header(data, len, h, 0, 0);
    if (rc) {
        free(h);
        return rc;
    }
    uint32_t ntx = cdiv(h->xsiz - h->xto, h->xt), nty = cdiv(h->ysiz - h->yto, h->yt);
    tile_t *tiles = (tile_t *)calloc(ntx * nty, sizeof(tile_t));
    tcomp_t *tcs = (tcomp_t *)calloc(h->ncomp, sizeof(tcomp_t));
    size_t n = 0;
    rc = tiles && tcs ? parse_tiles(data, len, h, tiles, ntx * nty) : ORACLE_E_NOMEM;
    for (uint32_t ti = 0; ti < ntx * nty && !rc; ++ti) {
        tile_t *t = &tiles[ti];
        for (int c = 0; c < h->ncomp && !rc; ++c) rc = build_tcomp(h, &t->comp[c], t->rect, &tcs[c]);
        if (rc) break;
        prog_list_t pl = {0};
        rc = build_progression(t, h->ncomp, tcs, &pl);
        if (!rc) parse_packets(t, h->ncomp, tcs, &pl);
        free(pl.v);
        for (int c = 0; c < h->ncomp; ++c) {
            for (int r = 0; r < tcs[c].nres; ++r) {
                res_t *rs = &tcs[c].res[r];
                uint32_t np = rs->npx * rs->npy;
                for (uint32_t p = 0; p < np; ++p)
                    for (int b = 0; b < rs->nbands; ++b) {
                        prec_band_t *pb = &rs->prec[(size_t)p * rs->nbands + b];
                        for (uint32_t k = 0; k < pb->ncbx * pb->ncby; ++k, ++n) {
                            cblk_t *cb = &pb->cbs[k];
                            if (!out || n >= cap) continue;
                            oracle_cb_t *o = &out[n];
                            o->tile = ti;
                            o->comp = (uint32_t)c;
                            o->res = (uint32_t)r;
                            o->band = (uint32_t)rs->bands[b].type;
                            o->x0 = cb->rect.x0;
                            o->y0 = cb->rect.y0;
                            o->x1 = cb->rect.x1;
                            o->y1 = cb->rect.y1;
                            o->offset = cb->data ? (int64_t)(cb->data - data) : -1;
                            o->lcup = cb->lcup;
                            o->lref = cb->lref;
                            o->zblk = cb->zblk;
                            o->passes = cb->data ? cb->passes : 0;
                        }
                    }
            }
            free_tcomp(&tcs[c]);
        }
    }
    for (uint32_t ti = 0; tiles && ti < ntx * nty; ++ti) free(tiles[ti].parts);
    free(tiles);
    free(tcs);
    free(h);
    if (count) *count = n;
    return rc;
}
