/* HT code-block decoder of the CPU oracle.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  Restates
 * crates/signinum-j2k-native/src/j2c/ht_block_decode.rs of the reference:
 * the MEL decoder (:783-879), the forward and backward bit readers
 * (:881-999), cleanup symbol decode (:1056-1204), the magnitude/sign phase
 * (:1435-1560), sigma construction (:1207-1251), significance propagation
 * (:1254-1374), magnitude refinement (:1377-1432) and decode_impl
 * (:1563-1680).
 */
#include <stdlib.h>
#include <string.h>

#include "oracle.h"
#include "oracle_internal.h"

extern uint16_t oracle_vlc_table0[1024];
extern uint16_t oracle_vlc_table1[1024];
extern uint16_t oracle_uvlc_table0[320];
extern uint16_t oracle_uvlc_table1[256];

int32_t oracle_coefficient_to_i32(uint32_t value, uint32_t k_max)
{
    /* ht_block_decode.rs:225-234 */
    uint32_t shift = k_max >= 31 ? 0 : 31 - k_max;
    int32_t magnitude = (int32_t)((value & 0x7FFFFFFFu) >> shift);
    return (value & 0x80000000u) ? -magnitude : magnitude;
}

/* ---- MEL decoder (ht_block_decode.rs:783-879) ---- */
typedef struct {
    const uint8_t *data;
    size_t pos, remaining;
    int unstuff;
    uint8_t cur;
    int bits_left;
    int k;
    int num_runs;
    uint64_t runs;
} mel_t;

static const int MEL_EXP[13] = {0, 0, 0, 1, 1, 1, 2, 2, 2, 3, 3, 4, 5};

static int mel_bit(mel_t *m)
{
    if (m->bits_left == 0) {
        uint8_t b;
        if (m->remaining > 0) {
            b = m->data[m->pos++];
            m->remaining--;
        } else {
            b = 0xFF;
        }
        if (m->remaining == 0)
            b |= 0x0F;
        m->cur = b;
        m->bits_left = 8 - (m->unstuff ? 1 : 0);
        m->unstuff = (b == 0xFF);
    }
    m->bits_left--;
    return (m->cur >> m->bits_left) & 1;
}

static void mel_more_runs(mel_t *m)
{
    while (m->num_runs < 8) {
        int eval = MEL_EXP[m->k];
        int first = mel_bit(m);
        uint64_t run;
        if (first) {
            m->k = m->k + 1 > 12 ? 12 : m->k + 1;
            run = (uint64_t)((1 << eval) - 1) << 1;
        } else {
            uint32_t v = 0;
            for (int i = 0; i < eval; ++i)
                v = (v << 1) | (uint32_t)mel_bit(m);
            m->k = m->k > 0 ? m->k - 1 : 0;
            run = ((uint64_t)v << 1) | 1;
        }
        m->runs |= run << (m->num_runs * 7);
        m->num_runs++;
        if (eval == 5 && first == 0 && m->num_runs >= 8)
            break;
    }
}

static int mel_get_run(mel_t *m)
{
    if (m->num_runs == 0)
        mel_more_runs(m);
    int run = (int)(m->runs & 0x7F);
    m->runs >>= 7;
    m->num_runs--;
    return run;
}

/* ---- forward reader with pad byte (ht_block_decode.rs:881-929) ---- */
typedef struct {
    const uint8_t *data;
    size_t len, pos;
    uint64_t tmp;
    uint32_t bits;
    int unstuff;
    uint8_t pad;
} fwd_t;

static void fwd_init(fwd_t *r, const uint8_t *data, size_t len, uint8_t pad)
{
    r->data = data;
    r->len = len;
    r->pos = 0;
    r->tmp = 0;
    r->bits = 0;
    r->unstuff = 0;
    r->pad = pad;
}

static uint32_t fwd_fetch(fwd_t *r)
{
    if (r->bits < 32) {
        while (r->bits <= 32) {
            uint8_t b = r->pos < r->len ? r->data[r->pos++] : r->pad;
            r->tmp |= (uint64_t)b << r->bits;
            r->bits += 8 - (r->unstuff ? 1 : 0);
            r->unstuff = (b == 0xFF);
        }
    }
    return (uint32_t)r->tmp;
}

static void fwd_advance(fwd_t *r, uint32_t n)
{
    r->tmp >>= n;
    r->bits -= n;
}

/* ---- backward reader (ht_block_decode.rs:931-999) ---- */
typedef struct {
    const uint8_t *data;
    ptrdiff_t pos;
    size_t remaining;
    uint64_t tmp;
    uint32_t bits;
    int unstuff;
} rev_t;

static void rev_init_vlc(rev_t *r, const uint8_t *data, size_t lcup, size_t scup)
{
    uint8_t d = data[lcup - 2];
    r->data = data;
    r->tmp = d >> 4;
    r->bits = 4 - ((r->tmp & 7) == 7 ? 1 : 0);
    r->pos = (ptrdiff_t)lcup - 3;
    r->remaining = scup - 2;
    r->unstuff = (d | 0x0F) > 0x8F;
}

static void rev_init_mrp(rev_t *r, const uint8_t *data, size_t len)
{
    r->data = data;
    r->pos = (ptrdiff_t)len - 1;
    r->remaining = len;
    r->tmp = 0;
    r->bits = 0;
    r->unstuff = 1;
}

static uint32_t rev_fetch(rev_t *r)
{
    if (r->bits < 32) {
        while (r->bits <= 32) {
            uint8_t b = 0;
            if (r->remaining > 0) {
                b = r->data[r->pos--];
                r->remaining--;
            }
            uint32_t d_bits = 8 - ((r->unstuff && (b & 0x7F) == 0x7F) ? 1 : 0);
            r->tmp |= (uint64_t)b << r->bits;
            r->bits += d_bits;
            r->unstuff = b > 0x8F;
        }
    }
    return (uint32_t)r->tmp;
}

static uint32_t rev_advance(rev_t *r, uint32_t n)
{
    r->tmp >>= n;
    r->bits -= n;
    return (uint32_t)r->tmp;
}

static size_t cleanup_stride(uint32_t width) { return (size_t)((width + 2 + 7) & ~7u); }
static size_t sigma_stride(uint32_t width) { return (size_t)((((width + 3) / 4) + 2 + 7) & ~7u); }

/* ---- cleanup symbols: MEL + VLC + UVLC (ht_block_decode.rs:1056-1204) ---- */
static int decode_cleanup_symbols(const uint8_t *d, size_t lcup, size_t scup, uint32_t width,
                                  uint32_t height, size_t sstr, uint16_t *sc)
{
    mel_t mel;
    memset(&mel, 0, sizeof(mel));
    mel.data = d;
    mel.pos = lcup - scup;
    mel.remaining = scup - 1;
    rev_t vlc;
    rev_init_vlc(&vlc, d, lcup, scup);

    int run = mel_get_run(&mel);
    uint32_t c_q = 0;
    size_t ro = 0;
    uint32_t x = 0;
    while (x < width) {
        uint32_t vlc_val = rev_fetch(&vlc);
        uint32_t t0 = oracle_vlc_table0[c_q + (vlc_val & 0x7F)];
        if (c_q == 0) {
            run -= 2;
            t0 = (run == -1) ? t0 : 0;
            if (run < 0)
                run = mel_get_run(&mel);
        }
        sc[ro] = (uint16_t)t0;
        x += 2;
        c_q = ((t0 & 0x10) << 3) | ((t0 & 0xE0) << 2);
        vlc_val = rev_advance(&vlc, t0 & 7);

        uint32_t t1 = oracle_vlc_table0[c_q + (vlc_val & 0x7F)];
        if (c_q == 0 && x < width) {
            run -= 2;
            t1 = (run == -1) ? t1 : 0;
            if (run < 0)
                run = mel_get_run(&mel);
        }
        if (x >= width)
            t1 = 0;
        sc[ro + 2] = (uint16_t)t1;
        x += 2;
        c_q = ((t1 & 0x10) << 3) | ((t1 & 0xE0) << 2);
        vlc_val = rev_advance(&vlc, t1 & 7);

        uint32_t mode = ((t0 & 8) << 3) | ((t1 & 8) << 4);
        if (mode == 0xC0) {
            run -= 2;
            if (run == -1)
                mode += 0x40;
            if (run < 0)
                run = mel_get_run(&mel);
        }
        uint32_t e = oracle_uvlc_table0[mode + (vlc_val & 0x3F)];
        vlc_val = rev_advance(&vlc, e & 7);
        e >>= 3;
        uint32_t len = e & 0xF;
        uint32_t tmp = vlc_val & ((1u << len) - 1);
        vlc_val = rev_advance(&vlc, len);
        e >>= 4;
        len = e & 7;
        e >>= 3;
        sc[ro + 1] = (uint16_t)(1 + (e & 7) + (tmp & ~(0xFFu << len)));
        sc[ro + 3] = (uint16_t)(1 + (e >> 3) + (tmp >> len));
        ro += 4;
    }
    sc[ro] = 0;
    sc[ro + 1] = 0;

    for (uint32_t y = 2; y < height; y += 2) {
        size_t rb = (size_t)(y >> 1) * sstr;
        size_t pb = rb - sstr;
        x = 0;
        c_q = 0;
        ro = rb;
        while (x < width) {
            size_t o = ro - rb;
            c_q |= ((uint32_t)sc[pb + o] & 0xA0) << 2;
            c_q |= ((uint32_t)sc[pb + o + 2] & 0x20) << 4;
            uint32_t vlc_val = rev_fetch(&vlc);
            uint32_t t0 = oracle_vlc_table1[c_q + (vlc_val & 0x7F)];
            if (c_q == 0) {
                run -= 2;
                t0 = (run == -1) ? t0 : 0;
                if (run < 0)
                    run = mel_get_run(&mel);
            }
            sc[ro] = (uint16_t)t0;
            x += 2;
            c_q = ((t0 & 0x40) << 2) | ((t0 & 0x80) << 1);
            c_q |= (uint32_t)sc[pb + o] & 0x80;
            c_q |= ((uint32_t)sc[pb + o + 2] & 0xA0) << 2;
            c_q |= ((uint32_t)sc[pb + o + 4] & 0x20) << 4;
            vlc_val = rev_advance(&vlc, t0 & 7);

            uint32_t t1 = oracle_vlc_table1[c_q + (vlc_val & 0x7F)];
            if (c_q == 0 && x < width) {
                run -= 2;
                t1 = (run == -1) ? t1 : 0;
                if (run < 0)
                    run = mel_get_run(&mel);
            }
            if (x >= width)
                t1 = 0;
            sc[ro + 2] = (uint16_t)t1;
            x += 2;
            c_q = ((t1 & 0x40) << 2) | ((t1 & 0x80) << 1);
            c_q |= (uint32_t)sc[pb + o + 2] & 0x80;
            vlc_val = rev_advance(&vlc, t1 & 7);

            uint32_t mode = ((t0 & 8) << 3) | ((t1 & 8) << 4);
            uint32_t e = oracle_uvlc_table1[mode + (vlc_val & 0x3F)];
            vlc_val = rev_advance(&vlc, e & 7);
            e >>= 3;
            uint32_t len = e & 0xF;
            uint32_t tmp = vlc_val & ((1u << len) - 1);
            vlc_val = rev_advance(&vlc, len);
            e >>= 4;
            len = e & 7;
            e >>= 3;
            sc[ro + 1] = (uint16_t)((e & 7) + (tmp & ~(0xFFu << len)));
            sc[ro + 3] = (uint16_t)((e >> 3) + (tmp >> len));
            ro += 4;
        }
        sc[ro] = 0;
        sc[ro + 1] = 0;
    }
    (void)mel;
    return 0;
}

/* ht_block_decode.rs:1010-1032 */
static uint32_t magsgn_sample(fwd_t *ms, uint32_t inf, uint32_t bit, uint32_t uq, uint32_t p,
                              uint32_t *v_out)
{
    if ((inf & (1u << (4 + bit))) == 0) {
        *v_out = 0;
        return 0;
    }
    uint32_t ms_val = fwd_fetch(ms);
    uint32_t m_n = uq - ((inf >> (12 + bit)) & 1);
    fwd_advance(ms, m_n);
    uint32_t value = ms_val << 31;
    uint32_t mask = m_n == 0 ? 0 : ((m_n >= 32) ? 0xFFFFFFFFu : ((1u << m_n) - 1));
    uint32_t v_n = ms_val & mask;
    v_n |= ((inf >> (8 + bit)) & 1) << m_n;
    v_n |= 1;
    value |= (v_n + 2) << (p - 1);
    *v_out = v_n;
    return value;
}

/* ht_block_decode.rs:1435-1560 */
static int decode_magsgn(const uint8_t *d, size_t lcup, size_t scup, const uint16_t *sc,
                         uint32_t *out, uint32_t zblk, uint32_t width, uint32_t height,
                         uint32_t stride, size_t sstr, uint32_t *v_n)
{
    size_t vn_w = (width + 1) / 2 + 2;
    memset(v_n, 0, vn_w * sizeof(uint32_t));
    uint32_t p = 30 - zblk;
    uint32_t mmsbp2 = zblk + 2;
    fwd_t ms;
    fwd_init(&ms, d, lcup - scup, 0xFF);

    for (uint32_t y = 0; y < height; y += 2) {
        size_t sp = (size_t)(y >> 1) * sstr;
        size_t vp = 0;
        size_t dp = (size_t)y * stride;
        uint32_t prev_v_n = 0;
        uint32_t x = 0;
        int second = (y + 1 < height);
        while (x < width) {
            uint32_t inf = sc[sp];
            uint32_t uq = sc[sp + 1];
            if (y > 0) {
                uint32_t gamma = inf & 0xF0;
                gamma &= gamma - 0x10;
                uint32_t emax = v_n[vp] | v_n[vp + 1];
                emax = 31 - (uint32_t)__builtin_clz(emax | 2);
                uint32_t kappa = gamma ? emax : 1;
                uq += kappa;
            }
            if (uq > mmsbp2)
                return -1;
            uint32_t vv;
            out[dp] = magsgn_sample(&ms, inf, 0, uq, p, &vv);
            uint32_t v1;
            uint32_t val1 = magsgn_sample(&ms, inf, 1, uq, p, &v1);
            if (second)
                out[dp + stride] = val1;
            v_n[vp] = prev_v_n | v1;
            prev_v_n = 0;
            dp++;
            x++;
            if (x >= width) {
                vp++;
                break;
            }
            out[dp] = magsgn_sample(&ms, inf, 2, uq, p, &vv);
            uint32_t v3;
            uint32_t val3 = magsgn_sample(&ms, inf, 3, uq, p, &v3);
            if (second)
                out[dp + stride] = val3;
            prev_v_n = v3;
            dp++;
            x++;
            sp += 2;
            vp++;
        }
        v_n[vp] = prev_v_n;
    }
    return 0;
}

/* ht_block_decode.rs:1207-1251 */
static void build_sigma(const uint16_t *sc, uint16_t *sigma, uint32_t width, uint32_t height,
                        size_t sstr, size_t mstr)
{
    for (uint32_t y = 0; y < height; y += 4) {
        size_t sp = (size_t)(y >> 1) * sstr;
        size_t dp = (size_t)(y >> 2) * mstr;
        for (uint32_t x = 0; x < width; x += 4, sp += 4, dp++) {
            uint32_t t0 = (((uint32_t)sc[sp] & 0x30) >> 4) | (((uint32_t)sc[sp] & 0xC0) >> 2);
            t0 |= (((uint32_t)sc[sp + 2] & 0x30) << 4) | (((uint32_t)sc[sp + 2] & 0xC0) << 6);
            uint32_t t1 = (((uint32_t)sc[sp + sstr] & 0x30) >> 2) | ((uint32_t)sc[sp + sstr] & 0xC0);
            t1 |= (((uint32_t)sc[sp + sstr + 2] & 0x30) << 6) |
                  (((uint32_t)sc[sp + sstr + 2] & 0xC0) << 8);
            sigma[dp] = (uint16_t)(t0 | t1);
        }
        sigma[dp] = 0;
    }
    size_t base = (size_t)((height + 3) / 4) * mstr;
    for (uint32_t x = 0; x <= (width + 3) / 4; ++x)
        sigma[base + x] = 0;
}

static const uint32_t SPREAD[16] = {0x33,    0x76,    0xEC,    0xC8,    0x330,   0x760,
                                    0xEC0,   0xC80,   0x3300,  0x7600,  0xEC00,  0xC800,
                                    0x33000, 0x76000, 0xEC000, 0xC8000};

static uint32_t pair_u32(const uint16_t *v, size_t i)
{
    return (uint32_t)v[i] | ((uint32_t)v[i + 1] << 16);
}

/* ht_block_decode.rs:1254-1374 */
static void apply_sigprop(const uint8_t *ref, size_t lref, const uint16_t *sigma, uint32_t *out,
                          uint32_t width, uint32_t height, uint32_t stride, size_t mstr,
                          int stripe_causal, uint32_t p, uint16_t *prev_row_sig)
{
    size_t prs_len = (width + 3) / 4 + 8;
    memset(prev_row_sig, 0, prs_len * sizeof(uint16_t));
    fwd_t sp;
    fwd_init(&sp, ref, lref, 0);
    for (uint32_t y = 0; y < height; y += 4) {
        uint32_t pattern = 0xFFFF;
        if (height - y < 4) {
            pattern = 0x7777;
            if (height - y < 3) {
                pattern = 0x3333;
                if (height - y < 2)
                    pattern = 0x1111;
            }
        }
        uint32_t prev = 0;
        size_t cur_row = (size_t)(y >> 2) * mstr;
        size_t next_row = cur_row + mstr;
        size_t dpp = (size_t)y * stride;
        for (uint32_t x = 0; x < width; x += 4) {
            uint32_t col_pattern = pattern;
            int s = (int)x + 4 - (int)width;
            if (s < 0)
                s = 0;
            col_pattern >>= (uint32_t)(s * 4);
            size_t idx = x >> 2;
            uint32_t ps = (uint32_t)prev_row_sig[idx] | ((uint32_t)prev_row_sig[idx + 1] << 16);
            uint32_t ns = pair_u32(sigma, next_row + idx);
            uint32_t u = (ps & 0x88888888u) >> 3;
            if (!stripe_causal)
                u |= (ns & 0x11111111u) << 3;
            uint32_t cs = pair_u32(sigma, cur_row + idx);
            uint32_t mbr = cs;
            mbr |= (cs & 0x77777777u) << 1;
            mbr |= (cs & 0xEEEEEEEEu) >> 1;
            mbr |= u;
            uint32_t t = mbr;
            mbr |= t << 4;
            mbr |= t >> 4;
            mbr |= prev >> 12;
            mbr &= col_pattern;
            mbr &= ~cs;
            uint32_t new_sig = 0;
            if (mbr) {
                uint32_t cwd = fwd_fetch(&sp);
                uint32_t cnt = 0;
                uint32_t inv_sig = ~cs & col_pattern;
                uint32_t cand = mbr, processed = 0;
                while (cand) {
                    uint32_t bit = (uint32_t)__builtin_ctz(cand);
                    uint32_t m = 1u << bit;
                    cand &= ~m;
                    processed |= m;
                    if (cwd & 1) {
                        new_sig |= m;
                        cand |= SPREAD[bit] & inv_sig & ~processed;
                    }
                    cwd >>= 1;
                    cnt++;
                }
                if (new_sig) {
                    uint32_t value = 3u << (p - 2);
                    size_t bb = dpp + x;
                    uint32_t sb = new_sig;
                    while (sb) {
                        uint32_t bit = (uint32_t)__builtin_ctz(sb);
                        sb &= ~(1u << bit);
                        size_t off = (bit >> 2) + (size_t)(bit & 3) * stride;
                        out[bb + off] = ((cwd & 1) << 31) | value;
                        cwd >>= 1;
                        cnt++;
                    }
                }
                fwd_advance(&sp, cnt);
            }
            uint32_t comb = new_sig | cs;
            prev_row_sig[idx] = (uint16_t)comb;
            if (idx + 1 < prs_len)
                prev_row_sig[idx + 1] = (uint16_t)(comb >> 16);
            uint32_t np = comb;
            np |= (comb & 0x7777) << 1;
            np |= (comb & 0xEEEE) >> 1;
            prev = (np | u) & 0xF000;
        }
    }
}

/* ht_block_decode.rs:1377-1432 */
static int apply_magref(const uint8_t *ref, size_t lref, const uint16_t *sigma, uint32_t *out,
                        uint32_t width, uint32_t height, uint32_t stride, size_t mstr, uint32_t p)
{
    if (p < 2)
        return -1;
    rev_t mr;
    rev_init_mrp(&mr, ref, lref);
    uint32_t half = 1u << (p - 2);
    for (uint32_t y = 0; y < height; y += 4) {
        size_t si = (size_t)(y >> 2) * mstr;
        size_t dpp = (size_t)y * stride;
        for (uint32_t i = 0; i < width; i += 8) {
            uint32_t cwd = rev_fetch(&mr);
            uint32_t sig = pair_u32(sigma, si);
            si += 2;
            uint32_t col_mask = 0xF;
            if (sig) {
                for (int j = 0; j < 8; ++j) {
                    if (sig & col_mask) {
                        size_t dp = dpp + i + (size_t)j;
                        uint32_t sm = 0x11111111u & col_mask;
                        for (int k = 0; k < 4; ++k) {
                            if (sig & sm) {
                                uint32_t sym = cwd & 1;
                                sym = (1 - sym) << (p - 1);
                                sym |= half;
                                out[dp] ^= sym;
                                cwd >>= 1;
                            }
                            sm <<= 1;
                            dp += stride;
                        }
                    }
                    col_mask <<= 4;
                }
            }
            rev_advance(&mr, (uint32_t)__builtin_popcount(sig));
        }
    }
    return 0;
}

int oracle_ht_decode_block(const uint8_t *data, uint32_t lcup, uint32_t lref, uint32_t width,
                           uint32_t height, uint32_t missing_msbs, uint32_t num_passes,
                           int stripe_causal, int phase_limit, uint32_t *out, uint32_t stride)
{
    oracle_build_tables();
    for (uint32_t y = 0; y < height; ++y)
        memset(out + (size_t)y * stride, 0, width * sizeof(uint32_t));
    if (width == 0 || height == 0)
        return ORACLE_OK;
    /* decode_impl, ht_block_decode.rs:1563-1599 */
    if (num_passes > 1 && lref == 0)
        num_passes = 1;
    if (num_passes > 3 || missing_msbs >= 30)
        return ORACLE_E_BLOCK;
    if (missing_msbs == 29 && num_passes > 1)
        num_passes = 1;
    uint32_t p = 30 - missing_msbs;
    if (lcup < 2)
        return ORACLE_E_BLOCK;
    /* cleanup_segment_suffix_length, :1042-1053 */
    size_t scup = ((size_t)data[lcup - 1] << 4) + (data[lcup - 2] & 0x0F);
    if (scup < 2 || scup > lcup || scup > 4079)
        return ORACLE_E_BLOCK;

    size_t sstr = cleanup_stride(width);
    size_t quad_rows = (height + 1) / 2;
    size_t mstr = sigma_stride(width);
    size_t sigma_rows = (height + 3) / 4 + 1;
    uint16_t *sc = (uint16_t *)calloc(sstr * (quad_rows + 1), sizeof(uint16_t));
    uint32_t *v_n = (uint32_t *)calloc((width + 1) / 2 + 2, sizeof(uint32_t));
    uint16_t *sigma = (uint16_t *)calloc(sigma_rows * mstr + 1, sizeof(uint16_t));
    uint16_t *prs = (uint16_t *)calloc((width + 3) / 4 + 8, sizeof(uint16_t));
    int rc = ORACLE_OK;
    if (!sc || !v_n || !sigma || !prs) {
        rc = ORACLE_E_NOMEM;
        goto done;
    }
    decode_cleanup_symbols(data, lcup, scup, width, height, sstr, sc);
    if (decode_magsgn(data, lcup, scup, sc, out, missing_msbs, width, height, stride, sstr, v_n)) {
        rc = ORACLE_E_BLOCK;
        goto done;
    }
    if (phase_limit == 0)
        goto done;
    if (num_passes > 1) {
        build_sigma(sc, sigma, width, height, sstr, mstr);
        apply_sigprop(data + lcup, lref, sigma, out, width, height, stride, mstr, stripe_causal,
                      p, prs);
        if (phase_limit == 1)
            goto done;
        if (num_passes > 2) {
            if (apply_magref(data + lcup, lref, sigma, out, width, height, stride, mstr, p)) {
                rc = ORACLE_E_BLOCK;
                goto done;
            }
        }
    }
done:
    free(sc);
    free(v_n);
    free(sigma);
    free(prs);
    return rc;
}
