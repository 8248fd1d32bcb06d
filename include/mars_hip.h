/*
 * mars_hip.h -- C ABI of the MI355X (gfx950) range x domain x isometry search
 * backend for the Mars fractal codec.
 *
 * The reference plugs search backends in through the Rust trait
 *   trait SearchProvider: Sync {
 *       fn search(&self, &SearchRequest) -> SearchOutcome;
 *       fn search_candidates(&self, &SearchRequest, max) -> RankedCandidates }
 * (crates/mars-codec/src/encode.rs:557-579), built once per plane by the
 * factory passed to encode_image_with_search (encode.rs:1054-1076).  This
 * library is what a thin FFI provider crate binds: `mars_hip_prepare` is the
 * factory body (it owns the plane's pool and runs the whole-plane search table
 * on the GPU), `mars_hip_search` is SearchProvider::search (a read-only,
 * thread-safe table lookup) of the prepared method -- the exhaustive search,
 * or IndexedSearchProvider with the Fisher retriever
 * (crates/mars-search/src/lib.rs:468-552).  The binding is one extern block
 * over these entry points, each taking plain pointers and sizes.
 *
 * Conventions: status 0 = ok, < 0 = error (mars_hip_last_error() says why).
 * Nothing unwinds across the ABI.  No CPU search fallback exists: a caller
 * that needs one must own it.
 */
#ifndef MARS_HIP_H
#define MARS_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MARS_HIP_OK 0
#define MARS_HIP_E_ARG (-1)         /* bad argument / unsupported geometry */
#define MARS_HIP_E_DEVICE (-2)      /* HIP runtime error */
#define MARS_HIP_E_NOT_PREPARED (-3)
#define MARS_HIP_E_NOT_IN_TABLE (-4) /* (row, col, size, shift) was not prepared */
#define MARS_HIP_E_UNSUPPORTED (-5)

/* Search kernels (mars_hip_set_kernel). */
#define MARS_HIP_KERNEL_MFMA_I8 0 /* exact MFMA correlation (binary16 at sizes 4, 8; split INT8
                                     at 2, 16, 32) + certified epilogue (default) */
#define MARS_HIP_KERNEL_FP32 1    /* plain FP32-FMA correlation + binary32 fit (baseline) */

/* Candidate retrieval methods (crates/mars-search/src/lib.rs:357-410 MethodName). */
#define MARS_HIP_METHOD_EXHAUSTIVE 0 /* every position x isometry (exhaustive.rs:15-33) */
#define MARS_HIP_METHOD_FISHER 1     /* Fisher classified buckets (fisher.rs:29-78) */

typedef struct mars_hip_ctx mars_hip_ctx;

/* The EncodeParams fields the search reads (encode.rs:27-60), plus the
 * retrieval method IndexedSearchProvider::build takes (lib.rs:478-484).
 * With MARS_HIP_METHOD_FISHER the context holds both the Fisher tables
 * (the provider) and the exhaustive ones (the RD rate warm-up, which the
 * reference keeps exhaustive for every method, encode.rs:1050, 1139-1169). */
typedef struct {
    uint32_t min_size;
    uint32_t max_size;
    uint32_t shift;
    uint32_t bits_alfa;
    uint32_t bits_beta;
    uint32_t method; /* MARS_HIP_METHOD_* (0: exhaustive) */
    double max_alfa;
} mars_hip_params;

/* encode.rs:224-244 Candidate with its RawMoments. */
typedef struct {
    uint32_t dom_row;
    uint32_t dom_col;
    uint32_t isometry;
    uint32_t qalfa;
    uint32_t qbeta;
    uint32_t _pad;
    double rms;
    int64_t s0, s1_x4, s2_x16, t0, t1_x4, t2;
} mars_hip_candidate;

/* Per-run device timings (milliseconds, HIP events on the context stream). */
typedef struct {
    double ms_step;     /* whole step: pool build + range prep + search + merge */
    double ms_pool;     /* preparation: pool build, range operands and sums, threshold seed */
    double ms_search;   /* search kernels, all sizes */
    double ms_by_size[6]; /* search kernel per size 2,4,8,16,32,64 (index log2(size)-1) */
    uint64_t evals;     /* range x position x isometry candidates per step (walk-only
                           tables: of the nodes the walk visits) */
    uint64_t survivors; /* candidates past the integer bound (MFMA kernel) */
    uint64_t survivors_by_size[6];  /* the same per size 2..64 */
    uint64_t exact_fits_by_size[6]; /* candidates that took the verbatim binary64 fit */
    double ms_fisher;      /* Fisher classification, buckets and bucket scans */
    uint64_t fisher_evals; /* Fisher candidates fitted per step */
    double ms_pool_build;  /* the domain pool alone: contraction, split planes, summed-area
                              tables, per-position window sums (within ms_pool) */
} mars_hip_timing;

/* One plane of a batch: w x h u8 pixels, row-major. */
typedef struct {
    const uint8_t *px;
    uint32_t w, h;
} mars_hip_plane;

/* Search chunking shapes chosen by mars_hip_autotune, per range size
 * 2, 4, 8, 16, 32, 64 (index log2(size) - 1). */
typedef struct {
    uint32_t target_blocks[6]; /* blocks per search launch the chunking aims at */
    uint32_t min_tiles[6];     /* least position tiles per block */
    double ms[6];              /* search time per size with the chosen shape */
} mars_hip_tuning;

/* Open a context on HIP device `device` (0-based).  Creates its own stream. */
int mars_hip_open(int device, mars_hip_ctx **out);
void mars_hip_close(mars_hip_ctx *ctx);

/*
 * Factory body (encode.rs:1054-1076 make_search + the provider's own index):
 * copy the w x h u8 plane to the device, build the domain pool, and run the
 * search for every range node a quadtree walk with these params can request
 * -- every aligned non-forced node from max_size down to min_size, plus the
 * border nodes down to size 2 (encode.rs:1330-1362, 1865-1919) -- once per
 * stride in shifts[0..n_shifts).  Single-threaded; copies its inputs.
 * Supported sizes: 2..32.  MARS_HIP_E_UNSUPPORTED for a plane whose window
 * copies exceed 32-bit offsets or a (size, stride) table of more than
 * 2^27 - 32 domain positions (32-bit byte offsets of its records).
 */
int mars_hip_prepare(mars_hip_ctx *ctx, const uint8_t *px, uint32_t w, uint32_t h,
                     const mars_hip_params *p, const uint32_t *shifts, uint32_t n_shifts);

/*
 * The factory body for a batch of planes (the same params and strides for
 * all): one pool per plane and one search table per (plane, size, stride),
 * built and searched by one step whose search launches each cover one range
 * size over every plane of the batch.  encode_image_with_search calls its
 * factory once per plane (encode.rs:1054-1076); a batch front-loads those
 * calls for many images.  Plane i is looked up with
 * mars_hip_search_plane(ctx, i, ...).  Copies its inputs.
 */
int mars_hip_prepare_batch(mars_hip_ctx *ctx, const mars_hip_plane *planes, uint32_t n_planes,
                           const mars_hip_params *p, const uint32_t *shifts, uint32_t n_shifts);

/*
 * SearchProvider::search (encode.rs:557-602): the winner for the range block
 * (row, col, size) at domain stride `shift`; with k > 1,
 * ExhaustiveSearch::search_candidates (encode.rs:598-601, search_top_k
 * 695-788): the best k ascending by rms, ties in enumeration order, into
 * out[0..k) -- k must not exceed what the context was prepared to rank
 * (mars_hip_set_ranked).  *n_out = how many (0 when no domain position is
 * legal); *evals is the number of fitted candidates the reference counts
 * (encode.rs:768).  Thread-safe, read-only.
 */
int mars_hip_search(const mars_hip_ctx *ctx, uint32_t row, uint32_t col, uint32_t size,
                    uint32_t shift, uint32_t k, mars_hip_candidate *out, uint32_t *n_out,
                    uint64_t *evals);

/* SearchProvider::search of plane `plane` of a prepared batch (plane 0 is
 * what mars_hip_search serves). */
int mars_hip_search_plane(const mars_hip_ctx *ctx, uint32_t plane, uint32_t row, uint32_t col,
                          uint32_t size, uint32_t shift, uint32_t k, mars_hip_candidate *out,
                          uint32_t *n_out, uint64_t *evals);

/*
 * The winner of one retrieval method for (plane, row, col, size, shift):
 * method MARS_HIP_METHOD_EXHAUSTIVE is ExhaustiveSearch::search
 * (encode.rs:582-602), MARS_HIP_METHOD_FISHER is IndexedSearchProvider::search
 * with the Fisher retriever (lib.rs:515-552: search_block_fitted over the
 * range's bucket; *evals = the bucket's size, *n_out = 0 for an empty pool).
 * Fisher needs a context prepared with that method.
 */
int mars_hip_search_method(const mars_hip_ctx *ctx, uint32_t plane, uint32_t method, uint32_t row,
                           uint32_t col, uint32_t size, uint32_t shift, uint32_t k,
                           mars_hip_candidate *out, uint32_t *n_out, uint64_t *evals);

/*
 * How many candidates per range the next prepare keeps, k = 1..32
 * (search_top_k's k; default 1).  Ranked tables (k > 1) serve
 * mars_hip_search with up to k and need the MFMA kernel.
 */
int mars_hip_set_ranked(mars_hip_ctx *ctx, uint32_t k);

/*
 * Exhaustive tables for a threshold walk only.  With t_rms > 0, the next
 * MARS_HIP_METHOD_FISHER prepare (k = 1) computes the exhaustive winners of
 * just the nodes a threshold walk at t_rms visits at the base stride -- the
 * RD warm-up's walk (encode.rs:1139-1169, t_rms 8) asks nothing else when the
 * final walk uses the Fisher retriever: every node of the largest size, and
 * the children of each searched node whose best rms exceeds t_rms and whose
 * size exceeds min_size (nodes split without a search pass theirs on).  The
 * step then runs the sizes one after another, largest first, each activated
 * on the device from the previous size's winners.  An exhaustive lookup of a
 * node the walk does not visit returns MARS_HIP_E_NOT_IN_TABLE; Fisher
 * lookups are unaffected.  t_rms <= 0 (the default) keeps full tables.  No
 * counterpart in the reference (whose warm-up searches on the CPU node by
 * node); mars_hip_encode_plane_method and mars_hip_encode_rgb set it for
 * Fisher encodes with a lambda.
 */
int mars_hip_set_walk_tables(mars_hip_ctx *ctx, double t_rms);

/* Device scratch of mars_hip_search_grid_noalloc, per GPU, for the life of
 * the process. */
#define MARS_HIP_NOALLOC_SCRATCH_BYTES ((size_t)64 << 20)

/*
 * The fixed-grid search of mars_hip_prepare (min_size == max_size) without
 * any allocation, for callers that cannot allocate device memory: code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc.
 * No counterpart in the reference (whose searches run on the CPU).
 *
 * Input: d_px, a w x h u8 plane in device memory.  Output: the winner of
 * every size x size range block of the grid inside the plane (row r, column
 * c of the grid at index r * (w / size) + c; the border blocks a quadtree
 * walk would split are not part of the grid) in device memory: d_out[i] the
 * reference Candidate (8-byte aligned), d_found[i] = 1, or 0 with a zeroed
 * record when no domain position is legal.  Same results as
 * mars_hip_prepare with min_size == max_size and p->shift as the one stride,
 * by the exact MFMA search (whatever mars_hip_set_kernel chose).
 *
 * Enqueues its work on `stream` (a hipStream_t; null: the context's stream)
 * and returns; the results are there once the stream reaches that point, and
 * d_px must stay valid until then.  Every buffer of the step lives in a
 * static device array of MARS_HIP_NOALLOC_SCRATCH_BYTES per GPU; a plane
 * whose step needs more is refused with MARS_HIP_E_UNSUPPORTED before
 * anything runs (a 1024 x 1024 plane fits at sizes 4..32 with stride >= 2
 * and at size 2 with stride >= 4; a 2048 x 2048 plane does not).
 * Every call on a GPU shares that array and starts by clearing the part it
 * uses: calls that overlap on one GPU (two streams, two threads, two
 * contexts of one process) race and must be serialised by the caller.
 * Calls no hipMalloc/hipFree and allocates no host memory on success.
 */
int mars_hip_search_grid_noalloc(mars_hip_ctx *ctx, const uint8_t *d_px, uint32_t w, uint32_t h,
                                 const mars_hip_params *p, mars_hip_candidate *d_out,
                                 uint8_t *d_found, void *stream);

/* Last error message of this context ("" if none). */
const char *mars_hip_last_error(const mars_hip_ctx *ctx);

/* ---- measurement and diagnostics (no reference counterpart) ---- */

/* Select the search kernel used by later prepare/run calls. */
int mars_hip_set_kernel(mars_hip_ctx *ctx, int kernel);

/*
 * Re-run the prepared step `iters` times on device-resident data (no host
 * copies inside), replayed from one captured HIP graph when use_graph != 0,
 * then refresh the lookup table.  Fills *timing with per-step averages.
 */
int mars_hip_run(mars_hip_ctx *ctx, int iters, int use_graph, mars_hip_timing *timing);

/*
 * Pick the search chunking shape per range size by timing the prepared step
 * over a fixed space of shapes (reps timed steps each, median); results do
 * not depend on the shape.  Keeps the best shapes for later runs, refreshes
 * the lookup table and reports the choice in *out (may be null).
 */
int mars_hip_autotune(mars_hip_ctx *ctx, uint32_t reps, mars_hip_tuning *out);

/* Number of prepared (plane, size, shift) tables and the nodes of table i. */
int mars_hip_table_count(const mars_hip_ctx *ctx, uint32_t *n_tables);
/* The batch plane table i belongs to. */
int mars_hip_table_plane(const mars_hip_ctx *ctx, uint32_t i, uint32_t *plane);
int mars_hip_table_info(const mars_hip_ctx *ctx, uint32_t i, uint32_t *size, uint32_t *shift,
                        uint32_t *n_ranges, uint32_t *n_positions);
/* Copy table i's winners (n_ranges entries, found[j] = 0 for no candidate). */
int mars_hip_table_read(const mars_hip_ctx *ctx, uint32_t i, uint32_t *rows, uint32_t *cols,
                        mars_hip_candidate *out, uint8_t *found);

/* Table i's ranked lists: *k per range, out[r*k + j] (n_ranges * k entries),
 * counts[r] filled entries of range r. */
int mars_hip_table_read_ranked(const mars_hip_ctx *ctx, uint32_t i, uint32_t *k,
                               mars_hip_candidate *out, uint8_t *counts);

/* Table i's winners of one method and each range's evals (evals may be null). */
int mars_hip_table_read_method(const mars_hip_ctx *ctx, uint32_t i, uint32_t method,
                               mars_hip_candidate *out, uint8_t *found, uint64_t *evals);

/*
 * The search kernel's own score of each winner of table i (n_ranges
 * entries): the MFMA search's binary64 rms (equal to the table's rms), or
 * the FP32 baseline's binary32 rms of the candidate it picked, before the
 * binary64 re-score that fills the table.  For precision reports.
 */
int mars_hip_table_scores(mars_hip_ctx *ctx, uint32_t i, double *score);

/*
 * Device self-test of the binary64 fit: evaluates the search's own fit
 * (encode.rs:247-289) on the GPU for n moment sets (6 int64 each, in
 * RawMoments order) and writes qalfa, qbeta, rms.
 */
int mars_hip_fit_f64(mars_hip_ctx *ctx, const int64_t *moments, uint32_t n, double max_alfa,
                     uint32_t bits_alfa, uint32_t bits_beta, uint32_t *qalfa, uint32_t *qbeta,
                     double *rms);

/*
 * Run one plain device-to-device copy kernel (k_copy_probe) over `bytes`
 * (a multiple of 16): the known-traffic reference against which hardware
 * HBM counters of the search kernels are scaled.  *ms = its duration.
 */
int mars_hip_copy_probe(mars_hip_ctx *ctx, size_t bytes, double *ms);

#ifdef __cplusplus
}
#endif
#endif
