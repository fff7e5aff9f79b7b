/* CPU oracle for the HTJ2K tile-decode path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * CPU decoder (frames-sg/signinum, crates/signinum-j2k-native) used as the
 * parity checker by the test suite, by the smoke check and by the
 * CPU-baseline leg of the benchmark.  The product library never links,
 * loads or calls it.
 */
#ifndef SGN_ORACLE_H
#define SGN_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    ORACLE_OK = 0,
    ORACLE_E_ARG = -1,
    ORACLE_E_UNSUPPORTED = -2,
    ORACLE_E_MALFORMED = -3,
    ORACLE_E_BLOCK = -4,
    ORACLE_E_BUFFER = -5,
    ORACLE_E_NOMEM = -6
};

/* Pixel formats, numbered like the product's C header. */
enum {
    ORACLE_GRAY8 = 0,
    ORACLE_RGB8 = 1,
    ORACLE_RGBA8 = 2,
    ORACLE_GRAY16 = 3,
    ORACLE_RGB16 = 4
};

typedef struct {
    uint32_t width, height;
    uint32_t num_components;
    uint32_t bit_depth;
    uint32_t num_levels;
    uint32_t reversible;
    uint32_t mct;
    uint32_t high_throughput;
    uint32_t tiles_x, tiles_y;
    uint32_t cb_width, cb_height;
    uint32_t progression;
} oracle_info_t;

void oracle_build_tables(void);
void oracle_get_tables(uint16_t *vlc0, uint16_t *vlc1, uint16_t *uvlc0, uint16_t *uvlc1);

int oracle_inspect(const uint8_t *data, size_t len, oracle_info_t *info);

/* Output dimensions of a decode at 1/denominator resolution. */
int oracle_scaled_dims(const uint8_t *data, size_t len, uint32_t denominator,
                       uint32_t *out_w, uint32_t *out_h);

/* Full decode into an interleaved buffer.  denominator is 1, 2, 4 or 8. */
int oracle_decode(const uint8_t *data, size_t len, int fmt, uint32_t denominator,
                  uint8_t *out, size_t out_len, size_t stride);

/* Decode to per-component planes before rounding (sign shift and MCT
 * applied).  use_f64 selects the textbook double-precision restatement of
 * the 9/7 and ICT steps instead of the reference's f32 arithmetic. */
int oracle_decode_planes(const uint8_t *data, size_t len, uint32_t denominator, int use_f64,
                         double *planes, size_t plane_len);

/* Tile batch like the reference's decode_tiles_into: `workers` threads,
 * contiguous chunks of ceil(n / workers), one decoder state per worker.
 * Returns ORACLE_OK or the first failing status, with its index. */
int oracle_decode_batch(const uint8_t *const *datas, const size_t *lens, size_t n, int fmt,
                        uint32_t denominator, uint8_t *const *outs, const size_t *out_lens,
                        const size_t *strides, int workers, size_t *first_fail_index);

/* One HT code-block: cleanup segment of lcup bytes followed by lref bytes of
 * refinement data.  Writes the u32 sign-magnitude samples (bit 31 sign,
 * magnitude aligned below bit 30) with row stride `stride`.  phase_limit 0
 * stops after cleanup, 1 after SigProp, 2 runs all passes. */
int oracle_ht_decode_block(const uint8_t *data, uint32_t lcup, uint32_t lref, uint32_t width,
                           uint32_t height, uint32_t missing_msbs, uint32_t num_passes,
                           int stripe_causal, int phase_limit, uint32_t *out, uint32_t stride);

int32_t oracle_coefficient_to_i32(uint32_t value, uint32_t k_max);

/* Code-block style bits of SPcod/SPcoc (Table A.19). */
enum {
    ORACLE_CB_BYPASS = 0x01,
    ORACLE_CB_RESET = 0x02,
    ORACLE_CB_TERMALL = 0x04,
    ORACLE_CB_CAUSAL = 0x08,
    ORACLE_CB_PTERM = 0x10,
    ORACLE_CB_SEGSYM = 0x20,
    ORACLE_CB_HT = 0x40
};

/* One codeword segment of a classic code-block: bytes [off, off + len) of
 * the block's combined data carry passes [start, end); arith 0 marks a raw
 * (bypass) segment. */
typedef struct {
    uint32_t off, len;
    uint8_t start, end, arith, pad;
} oracle_seg_t;

/* Segment index of coding pass p under selective bypass (Table D.9). */
uint32_t oracle_bypass_segment(uint32_t p);

/* One classic (EBCOT) code-block: `bitplanes` = the sub-band's bit-planes
 * less the missing ones, `passes` coding passes; band 0 LL, 1 HL, 2 LH,
 * 3 HH; style the code-block style bits.  Without bypass or termination on
 * each pass the whole data is one arithmetic segment and segs is unused.
 * Writes signed integer coefficients (bit-plane 0 at bit 0). */
int oracle_classic_decode_block(const uint8_t *data, size_t len, const oracle_seg_t *segs,
                                size_t nseg, uint32_t w, uint32_t h, uint32_t bitplanes,
                                uint32_t passes, int band, int style, int32_t *out, size_t stride);

/* One 2D synthesis level over [x0,x1) x [y0,y1) from four f32 bands
 * (idwt.rs:535-1143); out holds the level row-major. */
int oracle_idwt_level(int reversible, uint32_t x0, uint32_t y0, uint32_t x1, uint32_t y1,
                      const float *ll, const float *hl, const float *lh, const float *hh,
                      float *out);

/* Inverse RCT / ICT on f32 planes followed by the addends (mct.rs:71-156). */
int oracle_inverse_mct(int reversible, float *p0, float *p1, float *p2, size_t n, float a0,
                       float a1, float a2);

typedef struct {
    uint32_t tile, comp, res, band;
    uint32_t x0, y0, x1, y1;
    int64_t offset; /* within the tile's concatenated tile-part bodies */
    uint32_t lcup, lref, zblk, passes;
} oracle_cb_t;

/* Code-blocks of a full-resolution decode in geometry order. */
int oracle_code_blocks(const uint8_t *data, size_t len, oracle_cb_t *out, size_t cap,
                       size_t *count);

/* Test double: zero every refinement segment before block decoding. */
void oracle_set_zero_refinement(int enable);

/* Statistics of the last oracle_decode call on this thread: HT code-blocks
 * decoded, code-blocks with refinement data, maximum coding passes. */
void oracle_last_stats(uint32_t *blocks, uint32_t *refinement_blocks, uint32_t *max_passes);

/* Decode-cost model (mlp.c): 5-32-32-1 ReLU MLP, 1281 parameters. */
#define ORACLE_MLP_PARAMS 1281
double oracle_mlp_predict(const double *params, const float *x);
double oracle_mlp_loss(const double *params, const float *X, const float *Y, size_t n);
/* Mean loss; grad (1281 doubles) receives its gradient. */
double oracle_mlp_loss_grad(const double *params, const float *X, const float *Y, size_t n,
                            double *grad);

#ifdef __cplusplus
}
#endif
#endif
