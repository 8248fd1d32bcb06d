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
