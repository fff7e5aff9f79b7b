/* signinum-j2k-hip: MI355X (gfx950) HTJ2K tile decoder, C ABI.
 *
 * This is the drop-in boundary for the HTJ2K tile-decode path of
 * frames-sg/signinum.  Every entry point names the reference interface it
 * replaces.  Signatures carry plain pointers and sizes only.  All functions
 * return SGN_OK (0) or a negative status; sgn_last_error_message() returns
 * a thread-local description of the last failure.
 *
 * Threading: a session owns one HIP stream and its device buffers and must
 * be used from one host thread at a time; separate sessions may be used
 * concurrently (crates/signinum-j2k-cuda/src/session.rs:11-49 has the same
 * rule for CudaSession).
 */
#ifndef SIGNINUM_J2K_HIP_H
#define SIGNINUM_J2K_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes (mirrors J2kError, crates/signinum-j2k/src/error.rs:5-74,
 * and the adapter Error enum, crates/signinum-j2k-cuda/src/error.rs:6-41) */
#define SGN_OK 0
#define SGN_E_ARG (-1)         /* invalid argument / buffer / stride */
#define SGN_E_UNSUPPORTED (-2) /* J2kError::Unsupported */
#define SGN_E_MALFORMED (-3)   /* codestream parse failure */
#define SGN_E_BLOCK (-4)       /* DecodingError::CodeBlockDecodeFailure */
#define SGN_E_BUFFER (-5)      /* output buffer too small */
#define SGN_E_NOMEM (-6)
#define SGN_E_DEVICE (-7)      /* HIP runtime failure / device unavailable */
#define SGN_E_REGION (-8)      /* J2kError::InvalidRegion: region outside the image */

/* PixelFormat (crates/signinum-core/src/pixel.rs:13-58). */
typedef enum {
    SGN_GRAY8 = 0,
    SGN_RGB8 = 1,
    SGN_RGBA8 = 2,
    SGN_GRAY16 = 3,
    SGN_RGB16 = 4
} sgn_pixel_format;

/* Downscale denominators (crates/signinum-core/src/scale.rs:3-30):
 * None = 1, Half = 2, Quarter = 4, Eighth = 8. */

/* Image information (crates/signinum-j2k/src/view.rs:102-112, Info). */
typedef struct {
    uint32_t width, height;
    uint32_t num_components;
    uint32_t bit_depth;
    uint32_t num_levels;      /* decomposition levels (minimum over components) */
    uint32_t reversible;      /* 1: 5/3, 0: 9/7 */
    uint32_t mct;             /* multiple-component transform signalled */
    uint32_t high_throughput; /* HT block coder (Part 15) */
    uint32_t tiles_x, tiles_y;
    uint32_t cb_width, cb_height;
    uint32_t progression;
    /* Colorspace (crates/signinum-core/src/types.rs:9-20): 0 Grayscale,
     * 1 YCbCr, 2 Rgb, 3 Cmyk, 4 Ycck, 5 SRgb, 6 SGray, 7 IccTagged, 8 Rct,
     * 9 Ict; from a JP2 colr box, else inferred from the codestream
     * (crates/signinum-j2k/src/parse/mod.rs:41-49). */
    uint32_t colorspace;
    uint32_t container; /* 0: raw codestream, 1: JP2 file (CompressedPayloadKind) */
} sgn_j2k_info;

const char *sgn_last_error_message(void);

/* J2kDecoder::inspect (crates/signinum-j2k/src/view.rs:102). */
int sgn_j2k_inspect(const uint8_t *bytes, size_t len, sgn_j2k_info *info);

/* Output dimensions of decode_scaled_into at 1/denominator
 * (crates/signinum-j2k/src/decode.rs:12-33). */
int sgn_j2k_scaled_dims(const uint8_t *bytes, size_t len, uint32_t denominator,
                        uint32_t *width, uint32_t *height);

/* ---- containers ----------------------------------------------------------
 * Every entry point that takes codestream bytes also accepts a JP2 file
 * (crates/signinum-j2k/src/parse/boxes.rs:14-160) and decodes its jp2c
 * codestream; JP2 files whose samples change after the codestream decode
 * (palette, channel reordering, sYCC or CIELab) are refused with
 * SGN_E_UNSUPPORTED by the decode entry points (the native decoder's direct
 * device plane reuse rule, crates/signinum-j2k-native/src/lib.rs:1493-1509).
 *
 * DICOM encapsulated Pixel Data: extract_dicom_encapsulated_frames_with_limit
 * (crates/signinum-j2k-metal/src/dicom.rs:28-89).  Frames (at most `limit`)
 * are written back to back into `out`, frame k at frame_offsets[k] with
 * frame_lengths[k] bytes.  With out == NULL only *frame_count and
 * *total_bytes are set.  A Pixel Data element that is not encapsulated
 * gives SGN_E_UNSUPPORTED; malformed or truncated Pixel Data SGN_E_MALFORMED. */
int sgn_dicom_extract_frames(const uint8_t *bytes, size_t len, size_t limit, uint8_t *out,
                             size_t out_capacity, size_t *frame_offsets, size_t *frame_lengths,
                             size_t frames_capacity, size_t *frame_count, size_t *total_bytes);

/* ---- sessions (CudaSession, crates/signinum-j2k-cuda/src/session.rs:11-49) */
typedef struct sgn_hip_session sgn_hip_session;
int sgn_hip_session_create(int device, sgn_hip_session **out);
void sgn_hip_session_destroy(sgn_hip_session *session);
/* HIP stream handle of the session (a hipStream_t), for callers that order
 * their own work after a decode. */
void *sgn_hip_session_stream(sgn_hip_session *session);
int sgn_hip_session_synchronize(sgn_hip_session *session);
/* Kernel launches issued by the session so far
 * (CudaSurfaceStats::kernel_dispatches, surface.rs:18-27). */
uint64_t sgn_hip_session_dispatches(sgn_hip_session *session);

/* ---- device buffers and events (the role of CudaDeviceBuffer and the
 * launch/copy helpers, crates/signinum-cuda-runtime/src/lib.rs:325-376,
 * 778-950).  Copies and fills are ordered on the session stream; the
 * download/upload helpers return after the copy completed. */
int sgn_device_alloc(sgn_hip_session *session, size_t bytes, void **out);
void sgn_device_free(sgn_hip_session *session, void *ptr);
/* Fills bytes with `value` using a kernel (no memset node), asynchronous. */
int sgn_device_fill(sgn_hip_session *session, void *ptr, size_t bytes, uint8_t value);
/* Device-to-device copy on the session stream (asynchronous). */
int sgn_device_copy(sgn_hip_session *session, void *dst, const void *src, size_t bytes);
int sgn_device_download(sgn_hip_session *session, void *host, const void *dev, size_t bytes);
int sgn_device_upload(sgn_hip_session *session, void *dev, const void *host, size_t bytes);
/* Pitched download of `rows` rows of `row_bytes` each. */
int sgn_device_download_2d(sgn_hip_session *session, void *host, size_t host_stride,
                           const void *dev, size_t dev_pitch, size_t row_bytes, size_t rows);
/* Page-locked host memory (hipHostMalloc): codestreams placed in it go to
 * the device by DMA without a staging copy.  Free with sgn_host_free. */
int sgn_host_alloc(size_t bytes, void **out);
void sgn_host_free(void *ptr);
int sgn_hip_event_create(void **event);
void sgn_hip_event_destroy(void *event);
/* Records the event on the session stream. */
int sgn_hip_event_record(sgn_hip_session *session, void *event);
/* Milliseconds between two recorded events (waits for `end`). */
int sgn_hip_event_elapsed_ms(void *start, void *end, float *ms);

/* ---- single codestream decode ------------------------------------------
 * J2kDecoder::decode_into / decode_scaled_into (crates/signinum-j2k/src/
 * view.rs:168,267; adapter decoder.rs:156-200).  Decodes on the device and
 * copies the result to host memory `out` with row stride `stride`. */
int sgn_j2k_decode_into(sgn_hip_session *session, const uint8_t *bytes, size_t len, int fmt,
                        uint32_t denominator, uint8_t *out, size_t out_len, size_t stride);

/* ImageDecodeSubmit::submit_to_device / submit_scaled_to_device
 * (crates/signinum-j2k-cuda/src/decoder.rs:225-266): decodes into device
 * memory `dst` (pitch in bytes) on the session stream; asynchronous. */
int sgn_j2k_decode_device(sgn_hip_session *session, const uint8_t *bytes, size_t len, int fmt,
                          uint32_t denominator, void *dst, size_t dst_pitch, size_t dst_capacity);

/* ---- regions (J2kDecoder::decode_region_into / decode_region_scaled_into,
 * crates/signinum-j2k/src/view.rs:215-311; decode.rs:36-71) --------------
 * The region is in full-resolution image pixels; at 1/denominator the
 * output is its covering rectangle, x / d .. ceil((x + w) / d).  A region
 * outside the image gives SGN_E_REGION.  Only the code-blocks and synthesis
 * blocks the region's pixels depend on are decoded, and tiles outside it
 * are skipped. */
int sgn_j2k_region_dims(const uint8_t *bytes, size_t len, uint32_t denominator, uint32_t x,
                        uint32_t y, uint32_t w, uint32_t h, uint32_t *out_x, uint32_t *out_y,
                        uint32_t *out_w, uint32_t *out_h);
/* Decodes the region into host memory `out` (row stride in bytes). */
int sgn_j2k_decode_region_into(sgn_hip_session *session, const uint8_t *bytes, size_t len,
                               int fmt, uint32_t denominator, uint32_t x, uint32_t y, uint32_t w,
                               uint32_t h, uint8_t *out, size_t out_len, size_t stride);
/* Decodes the region into device memory `dst` (pitch in bytes) on the
 * session stream; asynchronous. */
int sgn_j2k_decode_region_device(sgn_hip_session *session, const uint8_t *bytes, size_t len,
                                 int fmt, uint32_t denominator, uint32_t x, uint32_t y,
                                 uint32_t w, uint32_t h, void *dst, size_t pitch,
                                 size_t capacity);

/* ---- tile batches --------------------------------------------------------
 * TileDecodeJob (crates/signinum-j2k/src/batch.rs:24-33) with a device
 * destination. */
typedef struct {
    const uint8_t *bytes;
    size_t len;
    void *dst;        /* device pointer */
    size_t pitch;     /* bytes per output row */
    size_t capacity;  /* bytes available at dst */
} sgn_tile_job;

/* Execution modes of a batch decode. */
#define SGN_MODE_PLAIN 0    /* one launch sequence per tile */
#define SGN_MODE_BATCHED 1  /* one launch per stage for the whole batch */
#define SGN_MODE_GRAPH 2    /* batched, captured once in a hipGraph and replayed */

/* decode_tiles_into / TileBatchDecodeManyDevice::decode_tiles_to_device
 * (crates/signinum-j2k/src/batch.rs:110-147; crates/signinum-j2k-cuda/src/
 * codec.rs:190-210).  On failure *first_fail holds the first failing tile
 * index (TileBatchError, batch.rs:47-67).  Synchronous.
 * SGN_MODE_BATCHED pipelines batches of more than 768 tiles in chunks
 * (SGN_PIPELINE_TILES in the environment sets the chunk): the host parses
 * chunk k+1 while chunk k's bytes travel and chunk k-1 decodes.  Codestreams
 * that lie back to back (gaps under 4 KB) in page-locked host memory (see
 * sgn_host_alloc) are copied to the device straight from it; others are
 * staged through the session's pinned buffers. */
int sgn_j2k_decode_tiles_device(sgn_hip_session *session, const sgn_tile_job *jobs, size_t n,
                                int fmt, uint32_t denominator, int mode, size_t *first_fail);

/* The same batch decode (SGN_MODE_BATCHED, one plan, no pipelining) with no
 * device or page-locked allocation: no counterpart in the reference; it is
 * for callers that cannot allocate (code under HIP graph capture, serving
 * loops and harnesses that forbid hipMalloc).  The plan (codestream bytes,
 * job tables, status words, work counters and coefficients) lives in a
 * file-scope device array of sgn_j2k_noalloc_scratch_bytes() bytes (64 MiB;
 * e.g. one 1024x1024 RGB tile or about fifty 256x256 RGB tiles), which each
 * GPU holds until the process ends; a batch that needs more is rejected
 * with SGN_E_ARG before any device work.  Codestreams are copied straight
 * from the caller's host memory.  Calls on one GPU share its scratch: this
 * process serialises them with a lock, and every counter the kernels need
 * at zero is cleared by a kernel at the start of the call.  Synchronous;
 * first_fail as above. */
int sgn_j2k_decode_tiles_device_noalloc(sgn_hip_session *session, const sgn_tile_job *jobs,
                                        size_t n, int fmt, uint32_t denominator,
                                        size_t *first_fail);
/* Capacity of the no-allocation scratch array in bytes. */
uint64_t sgn_j2k_noalloc_scratch_bytes(void);

/* ---- plans: parse once, decode many times -------------------------------
 * The flattened device plan plays the role of J2kDirectColorPlan
 * (crates/signinum-j2k-native/src/direct_plan.rs:12-205): code-block table,
 * IDWT steps and store/colour steps for every tile of a batch. */
typedef struct sgn_j2k_plan sgn_j2k_plan;

/* Parses all codestreams on the host (worker threads), builds the tables
 * and uploads codestream bytes and tables to device memory.  tile_status
 * (optional, n entries) receives each tile's parse status. */
int sgn_j2k_plan_create(sgn_hip_session *session, const sgn_tile_job *jobs, size_t n, int fmt,
                        uint32_t denominator, sgn_j2k_plan **out, int32_t *tile_status);
/* One request of a heterogeneous batch: per-tile pixel format, scale and
 * region (TileRegionScaledDecodeJob, crates/signinum-j2k/src/batch.rs:34-46;
 * the request record of the device plan). */
typedef struct {
    const uint8_t *bytes;
    size_t len;
    void *dst;            /* device pointer */
    size_t pitch;         /* bytes per output row */
    size_t capacity;      /* bytes available at dst */
    int32_t fmt;          /* sgn_pixel_format */
    uint32_t denominator; /* 1, 2, 4 or 8 */
    /* has_roi != 0: decode only the region [roi_x, roi_x + roi_w) x [roi_y,
     * roi_y + roi_h) of the full-resolution image; the output is the
     * region's covering rectangle at the requested scale
     * (Rect::scaled_covering), see sgn_j2k_region_dims */
    uint32_t has_roi;
    uint32_t roi_x, roi_y, roi_w, roi_h;
} sgn_tile_req;
/* decode_tiles_region_scaled_into (crates/signinum-j2k/src/batch.rs:
 * 151-195): every request's region (or whole tile) at its scale and format,
 * one batched plan; *first_fail as for sgn_j2k_decode_tiles_device.
 * Synchronous. */
int sgn_j2k_decode_tiles_region_device(sgn_hip_session *session, const sgn_tile_req *reqs,
                                       size_t n, size_t *first_fail);
/* As sgn_j2k_plan_create, with format, scale and region per tile. */
int sgn_j2k_plan_create_mixed(sgn_hip_session *session, const sgn_tile_req *reqs, size_t n,
                              sgn_j2k_plan **out, int32_t *tile_status);
void sgn_j2k_plan_destroy(sgn_j2k_plan *plan);
/* Re-targets tile i of the plan to a new device destination. */
int sgn_j2k_plan_set_output(sgn_j2k_plan *plan, size_t i, void *dst, size_t pitch,
                            size_t capacity);
/* Enqueues the decode of the whole plan on the session stream
 * (asynchronous); mode as above. */
int sgn_j2k_plan_decode(sgn_hip_session *session, sgn_j2k_plan *plan, int mode);

/* Enqueues one stage of the batched decode, for per-kernel timing:
 * SGN_STAGE_BLOCKS (HT block decode), SGN_STAGE_LEVELS (intermediate
 * synthesis levels), SGN_STAGE_FINAL (last level + colour + packing).
 * Running the three in order equals sgn_j2k_plan_decode(SGN_MODE_BATCHED). */
#define SGN_STAGE_BLOCKS 0
#define SGN_STAGE_LEVELS 1
#define SGN_STAGE_FINAL 2
int sgn_j2k_plan_decode_stage(sgn_hip_session *session, sgn_j2k_plan *plan, int stage);

/* ---- autotuning of the HT lane kernel ------------------------------------
 * No counterpart in the reference (its CPU decoder has no launch shapes):
 * the lane kernel comes in variants of work-group size, quad pairs decoded
 * between LDS refills, one-shot grid versus persistent waves pulling chunks
 * of 64, 128 or 256 code-blocks from a device work counter, and the
 * cooperative one-block-per-wave kernel. */
int sgn_ht_lane_variant_count(void);
const char *sgn_ht_lane_variant_name(int variant);
int sgn_j2k_plan_set_lane_variant(sgn_j2k_plan *plan, int variant);
int sgn_j2k_plan_get_lane_variant(const sgn_j2k_plan *plan);
/* Times every variant `reps` times on the plan's code-blocks (HIP events on
 * the session stream), keeps the fastest for later decodes of the plan and
 * writes it to *chosen; ms (optional, `count` entries) receives each
 * variant's mean time in milliseconds.  Synchronous. */
int sgn_j2k_plan_autotune(sgn_hip_session *session, sgn_j2k_plan *plan, int reps, float *ms,
                          size_t count, int *chosen);

/* ---- autotuning of the synthesis-level tile shape -------------------------
 * No counterpart in the reference.  The level stage's launches come in
 * shapes: 64x64 output blocks computed by four or by two waves per
 * work-group, each with or without the one-wave 32x32 kernel for launches
 * whose levels are all at most 32x32.  Every shape computes the same
 * coefficients. */
int sgn_dwt_shape_count(void);
const char *sgn_dwt_shape_name(int shape);
int sgn_j2k_plan_set_dwt_shape(sgn_j2k_plan *plan, int shape);
int sgn_j2k_plan_get_dwt_shape(const sgn_j2k_plan *plan);
/* Times the level stage `reps` times in every shape (HIP events on the
 * session stream; the code-block stage runs once first so the levels read
 * decoded bands), keeps the fastest and writes it to *chosen; ms (optional,
 * `count` entries) receives each shape's mean time.  Synchronous. */
int sgn_j2k_plan_autotune_dwt(sgn_hip_session *session, sgn_j2k_plan *plan, int reps, float *ms,
                              size_t count, int *chosen);

typedef struct {
    uint64_t tiles;
    uint64_t code_blocks;
    uint64_t coded_bytes;      /* codestream bytes uploaded */
    uint64_t output_bytes;     /* W*H*C*bytes-per-sample over all tiles */
    uint64_t output_pixels;
    uint64_t coefficient_bytes;/* device coefficient arena */
    uint32_t launches_batched; /* kernel launches per batched decode */
    uint32_t narrow_samples;   /* 1: code-block samples stored as int16 (every band has at
                                  most 15 magnitude bit-planes); 0: int32 */
    /* host time of the plan build (ms): per-tile parse on the worker pool,
     * merge into the batch tables, staging fill, copy submission (until the
     * copies completed, for a synchronous build) */
    float host_parse_ms, host_merge_ms, host_stage_ms, host_copy_ms;
    /* code-blocks of the retained resolutions that region decodes skip
     * (outside every region's dependencies, or in tiles outside it) */
    uint64_t pruned_code_blocks;
} sgn_j2k_plan_stats;
int sgn_j2k_plan_get_stats(const sgn_j2k_plan *plan, sgn_j2k_plan_stats *stats);
/* Output geometry of tile i: width, height, bytes per row (unpadded). */
int sgn_j2k_plan_tile_dims(const sgn_j2k_plan *plan, size_t i, uint32_t *width,
                           uint32_t *height, size_t *row_bytes);

/* ---- stage hooks (HtCodeBlockDecoder, crates/signinum-j2k-native/src/
 * lib.rs:832-909).  Host buffers in, host buffers out; for parity tests. */

/* HtCodeBlockDecodeJob (lib.rs:149-174). */
typedef struct {
    uint64_t data_offset;      /* cleanup bytes followed by refinement bytes */
    uint32_t cleanup_length;
    uint32_t refinement_length;
    uint32_t width, height;
    uint32_t output_offset;    /* element offset of the block's first sample */
    uint32_t output_stride;    /* in samples */
    uint8_t missing_bit_planes;
    uint8_t number_of_coding_passes;
    uint8_t num_bitplanes;
    uint8_t stripe_causal;
    float dequantization_step;
} sgn_ht_code_block_job;

/* The HT code-block jobs a full-resolution decode of `bytes` hands to the
 * hook (decode.rs:1572-1669 collects them per sub-band): data_offset is
 * relative to `bytes`, and each block gets its own packed output region
 * (output_offset = sum of preceding w*h, output_stride = width).  Blocks
 * that decode to zero are included with number_of_coding_passes = 0.  With
 * jobs == NULL only *count is written. */
int sgn_j2k_code_blocks(const uint8_t *bytes, size_t len, sgn_ht_code_block_job *jobs,
                        size_t capacity, size_t *count);

/* decode_code_block / decode_sub_band (lib.rs:864-909): decodes every job
 * on the device; writes f32 = coefficient * step into out (host). */
int sgn_ht_decode_code_blocks(sgn_hip_session *session, const uint8_t *data, size_t data_len,
                              const sgn_ht_code_block_job *jobs, size_t n, float *out,
                              size_t out_len);
/* Same, returning the raw u32 sign-magnitude samples of the HT passes
 * (ht_block_decode.rs:1563-1680) before conversion. */
int sgn_ht_decode_code_blocks_raw(sgn_hip_session *session, const uint8_t *data,
                                  size_t data_len, const sgn_ht_code_block_job *jobs, size_t n,
                                  uint32_t *out, size_t out_len);

/* HtCodeBlockBatchJob (lib.rs:188-197): a code-block at (output_x,
 * output_y) of its sub-band; code_block.data_offset is relative to `data`,
 * code_block.output_offset is ignored (the position is the block's place
 * in the band) and code_block.output_stride is the band width. */
typedef struct {
    uint32_t output_x, output_y;
    sgn_ht_code_block_job code_block;
} sgn_ht_code_block_batch_job;

/* decode_sub_band (lib.rs:864; HtSubBandDecodeJob lib.rs:200-209, called at
 * decode.rs:1641-1650): decodes every block of one width x height sub-band
 * on the device into `out` (row-major, out_len >= width * height) as
 * f32 = coefficient * step.  Samples outside the listed blocks are left as
 * the caller passed them. */
int sgn_ht_decode_sub_band(sgn_hip_session *session, const uint8_t *data, size_t data_len,
                           uint32_t width, uint32_t height,
                           const sgn_ht_code_block_batch_job *jobs, size_t n, float *out,
                           size_t out_len);

/* decode_single_decomposition_idwt (lib.rs:877; J2kSingleDecompositionIdwtJob
 * lib.rs:767-780): one 2D synthesis level over rect [x0,x1) x [y0,y1);
 * band arrays are row-major with the Annex B band sizes.  For 5/3 the f32
 * inputs are integers; each is truncated toward zero (NaN to 0, saturating
 * at the int32 range) before integer lifting, except on a rect one sample
 * wide or high at an odd origin, which halves that sample: there the 5/3
 * steps run in f32 on the inputs as given, as the reference computes them. */
int sgn_idwt_level(sgn_hip_session *session, int reversible, uint32_t x0, uint32_t y0,
                   uint32_t x1, uint32_t y1, const float *ll, const float *hl, const float *lh,
                   const float *hh, float *out);

/* decode_inverse_mct (lib.rs:890; J2kInverseMctJob lib.rs:785-800):
 * planes updated in place, addends applied after the transform. */
int sgn_inverse_mct(sgn_hip_session *session, int reversible, float *plane0, float *plane1,
                    float *plane2, size_t n, float addend0, float addend1, float addend2);

/* decode_store_component (lib.rs:899; J2kStoreComponentJob lib.rs:805-828,
 * called at decode.rs:1792-1810): on the device,
 * output[(output_y + y) * output_width + output_x + x] =
 *     input[(source_y + y) * input_width + source_x + x] + addend
 * for x < copy_width, y < copy_height; the rest of `output` is unchanged. */
int sgn_store_component(sgn_hip_session *session, const float *input, size_t input_len,
                        uint32_t input_width, uint32_t source_x, uint32_t source_y,
                        uint32_t copy_width, uint32_t copy_height, float *output,
                        size_t output_len, uint32_t output_width, uint32_t output_x,
                        uint32_t output_y, float addend);

/* ---- decode-cost model ---------------------------------------------------
 * No counterpart in the reference: a 5-32-32-1 ReLU MLP over code-block
 * header features (coded bytes, coding passes, Zblk, sub-band, sample
 * count; 5 floats per block) predicting decode cycles, used to order the
 * code-blocks of a plan so that the lanes of a wave carry similar work.  It
 * trains on cycles that the lane kernel itself records, with the library's
 * own HIP forward, backward and Adam kernels.  1281 float parameters:
 * W1[32][5] | b1[32] | W2[32][32] | b2[32] | W3[32] | b3. */
typedef struct sgn_cost_model sgn_cost_model;
int sgn_cost_model_create(sgn_hip_session *session, uint64_t seed, sgn_cost_model **out);
void sgn_cost_model_destroy(sgn_cost_model *model);
int sgn_cost_model_param_count(void);
int sgn_cost_model_get_params(const sgn_cost_model *model, float *out, size_t count);
int sgn_cost_model_set_params(sgn_cost_model *model, const float *params, size_t count);
/* Gradient of the mean loss 0.5 * (f(x) - y)^2 over n samples (X: 5 floats
 * per sample) into grad (1281 floats), computed on the device. */
int sgn_cost_model_gradient(sgn_hip_session *session, const sgn_cost_model *model,
                            const float *X, const float *Y, size_t n, float *grad);
/* `epochs` full-batch Adam steps on the device; *loss (optional) receives the
 * final mean loss. */
int sgn_cost_model_train(sgn_hip_session *session, sgn_cost_model *model, const float *X,
                         const float *Y, size_t n, uint32_t epochs, float lr, float *loss);
int sgn_cost_model_predict(sgn_hip_session *session, const sgn_cost_model *model,
                           const float *X, size_t n, float *out);
/* Decode-time samples: every `stride`-th non-empty code-block of the plan's
 * lane queue (at most max_samples) is decoded alone by one wave, which
 * records its shader-clock cycles.  features: 5 floats per sample. */
int sgn_j2k_plan_time_code_blocks(sgn_hip_session *session, sgn_j2k_plan *plan, size_t stride,
                                  size_t max_samples, float *features, float *cycles,
                                  size_t *count);
/* Order of the plan's lane code-blocks: 0 by the compressed-length
 * heuristic (coded bytes + samples / 4, the default), 1 input order, 2 by the
 * model's predicted cost (largest first), 3 by block shape (largest first,
 * input order within a shape), 4 by block shape and then by the tile's summed
 * heuristic cost (a tile's blocks stay together), 5 as 4 with the model's
 * predicted costs.  2 and 5 need `model`. */
enum {
    SGN_ORDER_LENGTH = 0,
    SGN_ORDER_INPUT = 1,
    SGN_ORDER_MODEL = 2,
    SGN_ORDER_GEOMETRY = 3,
    SGN_ORDER_GEOMETRY_COST = 4,
    SGN_ORDER_GEOMETRY_MODEL = 5
};
int sgn_j2k_plan_order_code_blocks(sgn_hip_session *session, sgn_j2k_plan *plan, int order,
                                   const sgn_cost_model *model);
/* Autotunes order and lane-kernel variant together: for each of the length,
 * input, geometry and geometry-cost orders runs sgn_j2k_plan_autotune, keeps
 * the fastest (order, variant) pair on the plan.  ms (optional, 4 entries, in
 * that order) receives each order's best time; *chosen the chosen order. */
int sgn_j2k_plan_autotune_order(sgn_hip_session *session, sgn_j2k_plan *plan, int reps,
                                float *ms, int *chosen);

/* ---- serving ---------------------------------------------------------------
 * No counterpart in the reference (its entry points decode per call): a
 * continuous-batching server.  Requests of 1..n tiles (any size, bit depth,
 * format, scale) are admitted as they arrive; a batch forms when the oldest
 * waiting request has waited window_us or max_batch_tiles tiles wait, and
 * takes waiting requests in arrival order while their device footprint fits
 * hbm_budget.  Up to max_inflight batches run at once, each on its own
 * stream, so new batches overlap running ones.  Outputs stay in device
 * memory until the request is released.  The caller drives the server with
 * sgn_server_pump(); one server is used from one thread at a time.  Device
 * blocks of finished batches and released requests are kept for reuse (up
 * to hbm_budget / 4 bytes of them; freed when the server is destroyed),
 * because freeing device memory waits for every batch in flight. */
typedef struct sgn_server sgn_server;
typedef struct {
    uint32_t window_us;    /* batching window (default 2000) */
    uint64_t hbm_budget;   /* device bytes for outputs and in-flight plans (default 64 GiB) */
    uint32_t max_batch_tiles;  /* default 4096 */
    uint32_t max_inflight;     /* concurrent batches / streams, 1..16 (default 4) */
} sgn_server_config;
typedef struct {
    const uint8_t *bytes;  /* copied at submission */
    size_t len;
    int32_t fmt;           /* sgn_pixel_format */
    uint32_t denominator;  /* 1, 2, 4 or 8 */
} sgn_serve_tile;
typedef struct {
    uint64_t id;
    int32_t status;        /* SGN_OK or the error of the first failing tile */
    uint32_t first_fail;   /* index of that tile within the request */
    uint32_t tiles;
    uint64_t pixels;       /* decoded output pixels */
    double latency_ms;     /* submission to completion observed by a pump */
    double queue_ms;       /* submission to batch launch */
} sgn_request_info;
typedef struct {
    uint64_t requests_submitted, requests_done, requests_failed;
    uint64_t batches, tiles;
    uint64_t pending, running_batches;
    uint64_t hbm_in_use;
} sgn_server_stats;
void sgn_server_default_config(sgn_server_config *config);
int sgn_server_create(int device, const sgn_server_config *config, sgn_server **out);
/* Waits for running batches, frees every output. */
void sgn_server_destroy(sgn_server *server);
/* Validates the tiles' main headers and queues the request. */
int sgn_server_submit(sgn_server *server, const sgn_serve_tile *tiles, size_t n, uint64_t *id);
/* Completes finished batches, then launches due batches on free streams
 * (flush != 0: launch regardless of the window).  Non-blocking. */
int sgn_server_pump(sgn_server *server, int flush, uint32_t *completed);
/* Moves up to cap finished-request records to out. */
int sgn_server_poll(sgn_server *server, sgn_request_info *out, size_t cap, size_t *n);
/* Device surface of tile `tile` of a finished request. */
int sgn_server_output(sgn_server *server, uint64_t id, size_t tile, void **dev, size_t *pitch,
                      uint32_t *width, uint32_t *height);
int sgn_server_release(sgn_server *server, uint64_t id);
int sgn_server_get_stats(const sgn_server *server, sgn_server_stats *out);

/* ---- test-input encoder (encode_j2k_lossless, crates/signinum-j2k/src/
 * encode.rs:166-182; J2kLosslessEncodeOptions encode.rs:71-94) */
typedef struct {
    uint32_t levels;          /* clamped to floor(log2(min tile side)) */
    uint32_t cb_width_log2;   /* code-block width exponent, 2..10 */
    uint32_t cb_height_log2;
    uint32_t reversible;      /* 1: 5/3 lossless; 0: 9/7 quantised */
    uint32_t use_mct;         /* RCT / ICT for 3+ components */
    uint32_t progression;     /* 0 LRCP .. 4 CPRL */
    uint32_t guard_bits;      /* 0: 1 (2 with RCT) reversible, 2 irreversible */
    float base_step;          /* irreversible base step relative to 2^P (0: 1/64) */
    uint32_t precinct_log2;   /* 0: maximal precincts; else PPx = PPy */
    uint32_t sop, eph;        /* SOP / EPH markers */
    uint32_t tile_width, tile_height; /* 0: one tile */
    /* image origin (XOsiz, YOsiz) and tile-grid origin (XTOsiz <= XOsiz,
     * YTOsiz <= YOsiz) on the reference grid; the image occupies
     * [x0, x0 + width) x [y0, y0 + height) */
    uint32_t image_x0, image_y0, tile_x0, tile_y0;
    uint32_t packed_headers; /* 0: packet headers in the packets, 1: PPM, 2: PPT */
    uint32_t block_coder;    /* 0: HT (Part 15), 1: classic EBCOT / MQ (Part 1) */
    uint32_t cb_style;       /* classic code-block style bits: 0x01 selective bypass,
                                0x02 context reset, 0x04 termination on each pass,
                                0x08 vertically causal context, 0x20 segmentation symbols */
    uint32_t ht_refine;      /* HT: code bit 0 in SigProp / MagRef passes after a cleanup
                                of the bits above; 1 only in code-blocks where SigProp
                                reaches every sample of magnitude 1 (lossless), 2 in every
                                code-block (the unreached samples are dropped) */
    uint32_t subsampling;    /* byte c: component c's XRsiz (low nibble) and YRsiz (high
                                nibble), 0 read as 1; the colour transform is skipped
                                unless the first three components share one */
} sgn_j2k_encode_options;

void sgn_j2k_encode_default_options(sgn_j2k_encode_options *opts);
/* pixels: interleaved u8 (bit_depth <= 8) or native-endian u16 samples.
 * On SGN_E_BUFFER *out_len holds the required size. */
int sgn_j2k_encode_ht(const void *pixels, uint32_t width, uint32_t height,
                      uint32_t num_components, uint32_t bit_depth, size_t row_stride,
                      const sgn_j2k_encode_options *opts, uint8_t *out, size_t out_cap,
                      size_t *out_len);

#ifdef __cplusplus
}
#endif
#endif
