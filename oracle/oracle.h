/*
 * CPU oracle for the Mars range x domain x isometry search.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/, the
 * smoke() entry point and bench.py's cpu_baseline leg may load it.  The
 * product (libmars_hip.so) never links, loads or falls back to it.
 *
 * It restates, operation for operation, the reference Rust CPU search
 * (daleobrien/mars, crates/mars-codec/src/encode.rs and
 * crates/mars-simd/src/moments.rs scalar references).  Every function cites
 * the reference lines it follows.  Build flags: -O2 -ffp-contract=off, no
 * fast-math, so every binary64 expression rounds exactly as Rust's does
 * (Rust never contracts a*b+c into an FMA).
 *
 * Parity pins (what fixes each function independently of this restatement)
 * are listed in tests/test_oracle.py.
 */
#ifndef MARS_ORACLE_H
#define MARS_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* encode.rs:224-232 RawMoments: s1/s2/t1 stored as 4x/16x/4x their value. */
typedef struct {
    int64_t s0, s1_x4, s2_x16, t0, t1_x4, t2;
} or_moments;

/* encode.rs:234-244 Candidate. */
typedef struct {
    uint32_t dom_row, dom_col;
    uint32_t isometry, qalfa, qbeta, _pad;
    double rms;
    or_moments m;
} or_candidate;

/* isometry.rs:7-19 */
void or_iso_map(uint32_t k, uint32_t u, uint32_t v, uint32_t size, uint32_t *i, uint32_t *j);

/* encode.rs:473-501 Contracted::build: out is (w/2)*(h/2), stride w/2. */
void or_contract(const uint8_t *px, uint32_t w, uint32_t h, int32_t *out);

/* moments.rs:21-44 domain_sums_scalar */
void or_domain_sums(const int32_t *plane, size_t stride, size_t dr, size_t dc, size_t size,
                    int64_t *s1, int64_t *s2);

/* encode.rs:833-842 permute_range */
void or_permute_range(const uint8_t *range, uint32_t k, uint32_t size, uint8_t *out);

/* moments.rs:157-176 dot_u8_i32_window_scalar (== cross_term_permuted) */
int64_t or_cross_term_permuted(const int32_t *plane, size_t stride, size_t dr, size_t dc,
                               size_t size, const uint8_t *range_k);

/* encode.rs:247-289 fit_f64 */
void or_fit_f64(const or_moments *m, double max_alfa, uint32_t bits_alfa, uint32_t bits_beta,
                uint32_t *qalfa, uint32_t *qbeta, double *rms);

/* encode.rs:295-337 fit_f32 */
void or_fit_f32(const or_moments *m, float max_alfa, uint32_t bits_alfa, uint32_t bits_beta,
                uint32_t *qalfa, uint32_t *qbeta, float *rms);

/* encode.rs:446-448 quantise_f64 */
uint32_t or_quantise_f64(double x, uint32_t max);

/* encode.rs:456-460 best_beta */
uint32_t or_best_beta(int64_t range_sum, int64_t pixel_count, uint32_t bits_beta);

/*
 * encode.rs:695-788 search_top_k (with retain_top_k, 676-685).
 * px: w*h plane; contracted: or_contract(px).  Writes up to k candidates
 * (ascending rms, encounter order on ties) to out, their count to *n_out and
 * the number of fitted candidates to *evals.
 */
void or_search_top_k(const uint8_t *px, uint32_t w, uint32_t h, const int32_t *contracted,
                     uint32_t row, uint32_t col, uint32_t size, uint32_t shift, double max_alfa,
                     uint32_t bits_alfa, uint32_t bits_beta, uint32_t k, or_candidate *out,
                     uint32_t *n_out, uint64_t *evals);

/*
 * Top-1 search over a list of range blocks of one size, split over `threads`
 * POSIX threads by contiguous range chunks (order-preserving: out[i] is
 * range i's winner whatever the thread count).  found[i] = 0 when no domain
 * position is legal.  Returns total evals.
 */
uint64_t or_search_table(const uint8_t *px, uint32_t w, uint32_t h, uint32_t size,
                         uint32_t shift, double max_alfa, uint32_t bits_alfa,
                         uint32_t bits_beta, const uint32_t *rows, const uint32_t *cols,
                         uint32_t n, int threads, or_candidate *out, uint8_t *found);
/* The same with the best k per range (search_top_k): out holds n*k records,
 * found[i] = how many of range i's k slots are filled. */
uint64_t or_search_table_k(const uint8_t *px, uint32_t w, uint32_t h, uint32_t size,
                           uint32_t shift, double max_alfa, uint32_t bits_alfa,
                           uint32_t bits_beta, const uint32_t *rows, const uint32_t *cols,
                           uint32_t n, int threads, uint32_t k, or_candidate *out,
                           uint8_t *found);

/*
 * The exhaustive search as a search-provider callback (the reference's
 * ExhaustiveSearch, encode.rs:582-602), for driving the encoder chain in
 * tests.  `out` receives up to `max` or_candidate records (the same layout
 * as the product's candidate record), ascending.
 */
typedef struct {
    const uint8_t *px;
    uint32_t w, h;
    int32_t *contracted;
    double max_alfa;
    uint32_t bits_alfa, bits_beta;
} or_provider;

or_provider *or_provider_new(const uint8_t *px, uint32_t w, uint32_t h, double max_alfa,
                             uint32_t bits_alfa, uint32_t bits_beta);
void or_provider_free(or_provider *p);
int or_search_cb(void *user, uint32_t row, uint32_t col, uint32_t size, uint32_t shift,
                 uint32_t max, void *out, uint32_t *n_out, uint64_t *evals);

/* ---- Fisher classified search and colour plumbing (fisher.c) ---- */

#define OR_FISHER_BUCKETS 72 /* 3 shape classes x 24 variance orders */
#define OR_FISHER_SLOTS 16

typedef struct {
    uint32_t dom_row, dom_col, dom_iso;
} or_fisher_entry;

/* fisher.rs:23-27 buckets, stored flat: bucket b = clas*24 + var_class holds
 * entries[start[b] .. start[b]+count[b]] in domain-position order; an empty
 * bucket aliases the first non-empty one (own_count[b] is its own size). */
typedef struct {
    uint32_t size, shift, P;
    uint32_t start[OR_FISHER_BUCKETS], count[OR_FISHER_BUCKETS], own_count[OR_FISHER_BUCKETS];
    or_fisher_entry *entries;
} or_fisher_index;

typedef struct {
    or_provider *base;
    or_fisher_index *ix[OR_FISHER_SLOTS];
} or_fisher_provider;

void or_fisher_tables(int32_t *ordering /* [24][6] */, uint8_t *mapping /* [8][8] */);
uint32_t or_fisher_match(const int32_t order[4]);
void or_fisher_newclass(const double *b, uint32_t size, uint32_t *iso, uint32_t *clas);
uint32_t or_fisher_variance_class(const double *b, uint32_t size);
void or_fisher_flips(const double *b, uint32_t size, uint32_t iso, double *out);
void or_fisher_classify(const double *b, uint32_t size, uint32_t *iso, uint32_t *clas,
                        uint32_t *var_class);
or_fisher_index *or_fisher_index_build(const int32_t *contracted, uint32_t w, uint32_t h,
                                       uint32_t size, uint32_t shift);
void or_fisher_index_free(or_fisher_index *ix);
uint32_t or_fisher_range_class(const uint8_t *px, uint32_t w, uint32_t row, uint32_t col,
                               uint32_t size, uint32_t *isom);
void or_fisher_search(const uint8_t *px, uint32_t w, uint32_t h, const int32_t *contracted,
                      const or_fisher_index *ix, uint32_t row, uint32_t col, double max_alfa,
                      uint32_t bits_alfa, uint32_t bits_beta, or_candidate *out, uint32_t *n_out,
                      uint64_t *evals);
/* Fisher winners of a list of ranges of one size (threaded like or_search_table);
 * evals_each (optional) receives each range's candidate count. */
uint64_t or_fisher_table(const uint8_t *px, uint32_t w, uint32_t h, uint32_t size,
                         uint32_t shift, double max_alfa, uint32_t bits_alfa, uint32_t bits_beta,
                         const uint32_t *rows, const uint32_t *cols, uint32_t n, int threads,
                         or_candidate *out, uint8_t *found, uint64_t *evals_each);
or_fisher_provider *or_fisher_provider_new(const uint8_t *px, uint32_t w, uint32_t h,
                                           double max_alfa, uint32_t bits_alfa,
                                           uint32_t bits_beta);
void or_fisher_provider_free(or_fisher_provider *p);
int or_fisher_cb(void *user, uint32_t row, uint32_t col, uint32_t size, uint32_t shift,
                 uint32_t max, void *out, uint32_t *n_out, uint64_t *evals);

/* metrics.rs:377-417: interleaved RGB (n pixels) -> Y, Cb, Cr planes. */
void or_ycbcr(const uint8_t *rgb, size_t n, uint8_t *y, uint8_t *cb, uint8_t *cr);
/* color.rs:60-78: out is ceil(w/2) x ceil(h/2). */
void or_downsample_box(const uint8_t *px, uint32_t w, uint32_t h, uint8_t *out);

#ifdef __cplusplus
}
#endif
#endif
