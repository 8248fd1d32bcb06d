/*
 * mars_codec.h -- C ABI of the CPU side of the Mars encode path that sits
 * above the search provider: the quadtree walks, RD mode decision, `.mars`
 * v0 writer and rANS entropy coding.  These stay on the CPU (they are
 * serial per image and cheap next to the search); they exist in this
 * library so that a search table computed on the GPU produces the same
 * `.mars` bytes as the reference's own CPU encode.
 *
 * Reference interfaces replaced:
 *   mars_encode_plane      encode::encode_image_with_search
 *                          (crates/mars-codec/src/encode.rs:1054-1126)
 *                          followed by mars_format::write (mars_format.rs:232-297)
 *   mars_search_fn         trait SearchProvider::search / search_candidates
 *                          (encode.rs:557-579)
 *   mars_wrap_gray         color::wrap_gray_stream (color.rs:382-384, 778-790)
 *   mars_hip_encode_plane  encmars --method <hip> (encmars.rs:780-796): the
 *                          provider is the GPU search table of a context.
 *   mars_hip_encode_prepared  the same for one plane of a prepared batch.
 *   mars_hip_encode_plane_method  encode_image_with_search with
 *                          IndexedSearchProvider::build(.., method)
 *                          (crates/mars-search/src/lib.rs:468-500; the
 *                          reference's own tests do exactly this,
 *                          crates/mars-search/tests/production_search.rs:79-90).
 *   mars_ycbcr             mars_core::metrics::ycbcr (metrics.rs:377-417)
 *   mars_downsample_box    color::downsample_box (color.rs:60-78)
 *   mars_write_container   color.rs:778-790 write_container
 *   mars_hip_encode_rgb    color::encode_color_image_with_residual_quantisation
 *                          (color.rs:400-489, RGB arm, no colour regions)
 *                          with each plane encoded through
 *                          mars_hip_encode_plane_method.
 *   mars_read, mars_write  mars_format::read / write (mars_format.rs:232-388)
 *   mars_ifs_parse         ifs::parse (ifs.rs:183-278)
 *   mars_decode_iterative  ifs::decode_iterative (ifs.rs:281-309, 406-484)
 *
 * Status: 0 ok, < 0 error (MARS_HIP_E_* codes of mars_hip.h).
 */
#ifndef MARS_CODEC_H
#define MARS_CODEC_H

#include <stddef.h>
#include <stdint.h>

#include "mars_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* encode.rs:27-60 EncodeParams (lambda: has_lambda = 0 is None). */
typedef struct {
    uint32_t min_size;
    uint32_t max_size;
    uint32_t shift;
    uint32_t bits_alfa;
    uint32_t bits_beta;
    uint32_t zero_threshold;
    double max_alfa;
    double t_rms;
    int32_t has_lambda;
    int32_t _pad;
    double lambda;
} mars_encode_params;

/* encode.rs:76-101 EncodeOptions (lambda_regions empty; rd_candidates 1). */
typedef struct {
    uint8_t allowed_modes[4];
    uint8_t adaptive_density;
    uint8_t lambda_adaptive_qstep; /* ResidualQuantisation::LambdaAdaptive */
    uint8_t _pad[2];
    double fixed_qstep;            /* ResidualQuantisation::Fixed (default 8.0) */
    uint32_t rd_candidates;        /* P5c ranked candidates for modes 2/3: 1 (default) .. 32 */
    uint32_t _pad2;
} mars_encode_options;

/* encode.rs:1010-1021 EncodeCounters + 1494-1500 ModeStats. */
typedef struct {
    uint64_t search_evals;
    uint64_t warmup_evals;
    uint64_t leaf_modes[4];
    uint64_t split_decisions;
    uint64_t leaf_decisions;
    uint64_t n_leaves;
} mars_encode_stats;

/* Most ranked candidates a provider is asked for (rd_candidates). */
#define MARS_MAX_RANKED 32
/* The RD warm-up's threshold walk runs at this t_rms (encode.rs:1139-1169). */
#define MARS_WARMUP_T_RMS 8.0

/*
 * SearchProvider::search_candidates (encode.rs:557-579): up to `max` of the
 * best candidates for range (row, col, size) at domain stride `shift`,
 * ascending by rms (ties in the provider's encounter order), into out[0..max);
 * *n_out = how many (0 when there is none).  max == 1 is SearchProvider::search.
 * A provider that cannot rank returns its winner only (the trait default).
 * Returns 0 or an error.
 */
typedef int (*mars_search_fn)(void *user, uint32_t row, uint32_t col, uint32_t size,
                              uint32_t shift, uint32_t max, mars_hip_candidate *out,
                              uint32_t *n_out, uint64_t *evals);

/*
 * encode_image_with_search + mars_format::write.  `search` serves the final
 * walk, `warmup` the RD rate warm-up (the reference hard-codes the
 * exhaustive search there, encode.rs:1153; any provider returning the same
 * winners keeps the bytes).  *out is malloc'd; free it with mars_free.
 */
int mars_encode_plane(const uint8_t *px, uint32_t w, uint32_t h, const mars_encode_params *p,
                      const mars_encode_options *o, mars_search_fn search, void *search_user,
                      mars_search_fn warmup, void *warmup_user, uint8_t **out, size_t *out_len,
                      mars_encode_stats *stats);

/* color::wrap_gray_stream: the MARC container around one grey plane stream. */
int mars_wrap_gray(const uint8_t *plane, size_t len, uint8_t **out, size_t *out_len);

void mars_free(void *p);

/* Colour container modes (color.rs:369-375, MODE_*). */
#define MARS_COLOR_GRAY 0
#define MARS_COLOR_RGB_444 1
#define MARS_COLOR_RGB_420 2

/* metrics.rs:377-417: n interleaved RGB pixels -> BT.601 full-range Y, Cb, Cr
 * (binary64, rounded half away from zero, clamped to 0..255). */
int mars_ycbcr(const uint8_t *rgb, size_t n, uint8_t *y, uint8_t *cb, uint8_t *cr);

/* color.rs:60-78: 2x2 box average with +2 rounding into ceil(w/2) x ceil(h/2);
 * odd sides repeat the last row / column. */
int mars_downsample_box(const uint8_t *px, uint32_t w, uint32_t h, uint8_t *out);

/* color.rs:778-790: "MARC", version 0, mode, stream count, then each stream
 * as a u32 little-endian length and its bytes.  *out is malloc'd. */
int mars_write_container(uint8_t mode, const uint8_t *const *streams, const size_t *lens,
                         uint32_t n_streams, uint8_t **out, size_t *out_len);

/*
 * The encode of mars_encode_plane with the GPU search table of `ctx` as the
 * provider for both the warm-up and the final walk.  Prepares ctx for the
 * plane (strides shift, and 2*shift when adaptive density is on).
 */
int mars_hip_encode_plane(mars_hip_ctx *ctx, const uint8_t *px, uint32_t w, uint32_t h,
                          const mars_encode_params *p, const mars_encode_options *o,
                          uint8_t **out, size_t *out_len, mars_encode_stats *stats);

/*
 * Encode plane `plane` of the batch prepared in ctx (mars_hip_prepare_batch)
 * from its search table: the walks of mars_encode_plane served by
 * mars_hip_search_plane.  The search fields of *p must equal the prepared
 * params, and the strides the walk requests must have been prepared.
 * Read-only on the context: several planes may be encoded concurrently.
 */
int mars_hip_encode_prepared(mars_hip_ctx *ctx, uint32_t plane, const mars_encode_params *p,
                             const mars_encode_options *o, uint8_t **out, size_t *out_len,
                             mars_encode_stats *stats);

/*
 * mars_hip_encode_plane with the provider built for `method`
 * (MARS_HIP_METHOD_*): the final walk is served by that method's table, the
 * RD warm-up by the exhaustive one, as in the reference.
 */
int mars_hip_encode_plane_method(mars_hip_ctx *ctx, const uint8_t *px, uint32_t w, uint32_t h,
                                 const mars_encode_params *p, const mars_encode_options *o,
                                 uint32_t method, uint8_t **out, size_t *out_len,
                                 mars_encode_stats *stats);

/* ---- reading streams back: the `.mars` v0 reader, the legacy `.ifs` parser
 * and the iterative decoder (all CPU) ---- */

/* Typed errors of the readers, mars_format.rs:141-181 MarsFormatError and
 * ifs.rs:130-156 IfsError (one code per variant; the field values the Rust
 * variants carry are not reported). */
#define MARS_FMT_E_BAD_MAGIC (-101)
#define MARS_FMT_E_INVALID_QSTEP (-102)
#define MARS_FMT_E_UNSUPPORTED_VERSION (-103)
#define MARS_FMT_E_TRUNCATED (-104)
#define MARS_FMT_E_DEGENERATE_HEADER (-105)
#define MARS_FMT_E_SECTION_OUT_OF_BOUNDS (-106)
#define MARS_FMT_E_MISSING_TREE (-107)
#define MARS_FMT_E_INVALID_TREE_PAYLOAD (-108) /* lib.rs:292-297 InvalidCompressedData */
#define MARS_FMT_E_DOMAIN_OUT_OF_BOUNDS (-109)
#define MARS_FMT_E_TRAILING_BITS (-110)         /* .ifs: 8 or more bits after the tree */
#define MARS_FMT_E_NONZERO_PADDING (-111)       /* .ifs */
/* A context alphabet above 2^20 (bits_alfa or bits_beta above 20): the
 * format allows up to 24, but every decoded symbol scans its alphabet; the
 * reference panics at 24. */
#define MARS_FMT_E_UNSUPPORTED (-112)

/* mars_format.rs:59-64 Header: the ifs.rs Header geometry plus the residual
 * step as binary32 bits (quant.rs:18-64 ResidualQstep). */
typedef struct {
    uint32_t min_size, max_size, shift, bits_alfa, bits_beta, int_max_alfa;
    uint32_t width, height;
    uint32_t residual_qstep_bits;
    uint32_t _pad;
} mars_header;

/* ifs.rs:100-117 Leaf.  A mode-3 leaf's size*size residual levels are
 * residual[residual_off .. residual_off + residual_len) of the array that
 * comes with the leaves (residual_len = 0 for the other modes). */
typedef struct {
    uint32_t row, col, size, mode;
    uint32_t qalfa, qbeta, isometry, dom_row, dom_col;
    int32_t qgx, qgy;
    uint32_t residual_off, residual_len;
} mars_leaf;

/*
 * mars_format::read (mars_format.rs:302-388): decode a `.mars` v0 plane
 * stream into its header and leaves (traversal order: NW, SW, NE, SE).
 * Arbitrary input returns 0 or a MARS_FMT_E_* code, never crashes, and does
 * work bounded by the header (at most 2^24 leaf positions).  *leaves and
 * *residual are malloc'd (free with mars_free; null when empty).
 */
int mars_read(const uint8_t *data, size_t len, mars_header *hdr, mars_leaf **leaves,
              size_t *n_leaves, int32_t **residual, size_t *n_residual);

/* mars_format::write (mars_format.rs:232-297) of a header and leaves. */
int mars_write(const mars_header *hdr, const mars_leaf *leaves, size_t n_leaves,
               const int32_t *residual, size_t n_residual, uint8_t **out, size_t *out_len);

/* ifs::parse (ifs.rs:183-278): a legacy Mars 1 `.ifs` bitstream. */
int mars_ifs_parse(const uint8_t *data, size_t len, mars_header *hdr, mars_leaf **leaves,
                   size_t *n_leaves);

/*
 * ifs::decode_iterative (ifs.rs:281-309, decode_leaf 406-484): `iterations`
 * fixed-point iterations from a flat 128 image into out (width * height
 * bytes).  Leaves are checked first (inside the plane, domain inside the
 * plane, mode 0..3, isometry 0..7, residual length size^2 for mode 3);
 * MARS_HIP_E_ARG otherwise.
 */
int mars_decode_iterative(const mars_header *hdr, const mars_leaf *leaves, size_t n_leaves,
                          const int32_t *residual, size_t n_residual, uint32_t iterations,
                          uint8_t *out);

/* color.rs:338-345 ColorEncodeStats, with each plane's encode counters. */
typedef struct {
    uint64_t y_bytes, cb_bytes, cr_bytes;
    mars_encode_stats y, cb, cr;
} mars_color_stats;

/*
 * Encode a w x h interleaved RGB image: BT.601 YCbCr, chroma box-downsampled
 * when subsampling is MARS_COLOR_RGB_420 (MARS_COLOR_RGB_444 keeps it), Y
 * encoded with *y_params and Cb, Cr with *chroma_params, every plane with
 * `method` as its provider, and the three streams wrapped in the MARC
 * container of that mode.  *out is malloc'd.
 */
int mars_hip_encode_rgb(mars_hip_ctx *ctx, const uint8_t *rgb, uint32_t w, uint32_t h,
                        uint32_t subsampling, const mars_encode_params *y_params,
                        const mars_encode_params *chroma_params, const mars_encode_options *o,
                        uint32_t method, uint8_t **out, size_t *out_len, mars_color_stats *stats);

#ifdef __cplusplus
}
#endif
#endif
