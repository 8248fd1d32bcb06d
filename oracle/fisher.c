/*
 * CPU oracle: restatement of the reference Fisher classified search and of
 * the colour plumbing around it (BT.601 YCbCr, 4:2:0 box downsampling).
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Reference: daleobrien/mars
 *   crates/mars-search/src/classify.rs   (quadrant sums, ordering sort,
 *                                         variance_2, newclass,
 *                                         variance_class, flips)
 *   crates/mars-search/src/tables.rs     (ORDERING, MAPPING, match_order)
 *   crates/mars-search/src/fisher.rs     (index, candidates)
 *   crates/mars-search/src/lib.rs        (domain_positions, domain_block,
 *                                         search_block_fitted)
 *   crates/mars-core/src/metrics.rs      (ycbcr)
 *   crates/mars-codec/src/color.rs       (downsample_box)
 */
#include "oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* tables.rs:13-38: descending quadrant order (a,b,c,d) -> (isometry, class). */
static const int ORDERING[24][6] = {
    {3, 2, 1, 0, 3, 0}, {2, 3, 1, 0, 5, 1}, {2, 1, 3, 0, 5, 2}, {3, 1, 2, 0, 7, 0},
    {1, 3, 2, 0, 1, 1}, {1, 2, 3, 0, 1, 2}, {3, 2, 0, 1, 3, 1}, {2, 3, 0, 1, 5, 0},
    {2, 1, 0, 3, 2, 2}, {3, 1, 0, 2, 7, 1}, {1, 3, 0, 2, 1, 0}, {1, 2, 0, 3, 4, 2},
    {3, 0, 2, 1, 3, 2}, {2, 0, 3, 1, 2, 0}, {2, 0, 1, 3, 2, 1}, {3, 0, 1, 2, 7, 2},
    {1, 0, 2, 3, 4, 1}, {1, 0, 3, 2, 4, 0}, {0, 3, 2, 1, 6, 2}, {0, 2, 3, 1, 6, 1},
    {0, 2, 1, 3, 6, 0}, {0, 3, 1, 2, 0, 2}, {0, 1, 3, 2, 0, 1}, {0, 1, 2, 3, 0, 0},
};

/* tables.rs:42-51: isometry applied to a candidate = MAPPING[range iso][domain iso]. */
static const uint8_t MAPPING[8][8] = {
    {0, 1, 2, 3, 4, 5, 6, 7}, {2, 0, 3, 1, 7, 6, 4, 5}, {1, 3, 0, 2, 6, 7, 5, 4},
    {3, 2, 1, 0, 5, 4, 7, 6}, {4, 7, 6, 5, 0, 3, 2, 1}, {5, 6, 7, 4, 3, 0, 1, 2},
    {6, 4, 5, 7, 1, 2, 0, 3}, {7, 5, 4, 6, 2, 1, 3, 0},
};

void or_fisher_tables(int32_t *ordering, uint8_t *mapping)
{
    memcpy(ordering, ORDERING, sizeof ORDERING);
    memcpy(mapping, MAPPING, sizeof MAPPING);
}

/* tables.rs:57-63 match_order */
uint32_t or_fisher_match(const int32_t order[4])
{
    for (uint32_t j = 0; j < 24; j++)
        if (ORDERING[j][0] == order[0] && ORDERING[j][1] == order[1] &&
            ORDERING[j][2] == order[2] && ORDERING[j][3] == order[3])
            return j;
    abort(); /* order is a permutation of 0..3 by construction */
}

/* classify.rs:63-75: three-pass descending bubble sort carrying the order. */
static void sort_desc(double a[4], int32_t order[4])
{
    for (int i = 0; i < 4; i++) order[i] = i;
    for (int i = 2; i >= 0; i--)
        for (int j = 0; j <= i; j++)
            if (a[j] < a[j + 1]) {
                double t = a[j];
                a[j] = a[j + 1];
                a[j + 1] = t;
                int32_t o = order[j];
                order[j] = order[j + 1];
                order[j + 1] = o;
            }
}

/* classify.rs:79-92 variance_2 over the h x h sub-block at (atx, aty). */
static double variance_2(const double *b, uint32_t size, uint32_t atx, uint32_t aty, uint32_t h)
{
    double sum = 0.0, sum2 = 0.0;
    for (uint32_t i = atx; i < atx + h; i++)
        for (uint32_t j = aty; j < aty + h; j++) {
            double v = b[i * size + j];
            sum += v;
            sum2 += v * v;
        }
    double n = (double)(h * h);
    return sum2 / n - (sum / n) * (sum / n);
}

/* classify.rs:97-103 newclass (quadrant sums classify.rs:42-56). */
void or_fisher_newclass(const double *b, uint32_t size, uint32_t *iso, uint32_t *clas)
{
    uint32_t h = size / 2;
    double a[4] = {0.0, 0.0, 0.0, 0.0};
    for (uint32_t i = 0; i < h; i++)
        for (uint32_t j = 0; j < h; j++) {
            a[0] += b[i * size + j];
            a[1] += b[i * size + j + h];
            a[2] += b[(i + h) * size + j];
            a[3] += b[(i + h) * size + j + h];
        }
    int32_t order[4];
    sort_desc(a, order);
    uint32_t idx = or_fisher_match(order);
    *iso = (uint32_t)ORDERING[idx][4];
    *clas = (uint32_t)ORDERING[idx][5];
}

/* classify.rs:108-119 variance_class */
uint32_t or_fisher_variance_class(const double *b, uint32_t size)
{
    uint32_t h = size / 2;
    double a[4] = {variance_2(b, size, 0, 0, h), variance_2(b, size, 0, h, h),
                   variance_2(b, size, h, 0, h), variance_2(b, size, h, h, h)};
    int32_t order[4];
    sort_desc(a, order);
    return or_fisher_match(order);
}

/* classify.rs:140-173 flips: out[i][j] = b[source(i, j)]. */
void or_fisher_flips(const double *b, uint32_t size, uint32_t iso, double *out)
{
    for (uint32_t i = 0; i < size; i++)
        for (uint32_t j = 0; j < size; j++) {
            uint32_t si, sj;
            switch (iso) {
            case 0: si = i; sj = j; break;
            case 1: si = j; sj = size - i - 1; break;
            case 2: si = size - j - 1; sj = i; break;
            case 3: si = size - i - 1; sj = size - j - 1; break;
            case 4: si = i; sj = size - j - 1; break;
            case 5: si = size - i - 1; sj = j; break;
            case 6: si = j; sj = i; break;
            default: si = size - j - 1; sj = size - i - 1; break;
            }
            out[i * size + j] = b[si * size + sj];
        }
}

/* fisher.rs:35-38 and 59-61: newclass, flips by its isometry, variance_class. */
void or_fisher_classify(const double *b, uint32_t size, uint32_t *iso, uint32_t *clas,
                        uint32_t *var_class)
{
    double *f = (double *)malloc(sizeof(double) * size * size);
    or_fisher_newclass(b, size, iso, clas);
    or_fisher_flips(b, size, *iso, f);
    *var_class = or_fisher_variance_class(f, size);
    free(f);
}

/*
 * fisher.rs:29-56 Fisher::index over DomainPool::domain_positions
 * (lib.rs:60-73, row-major over the stride grid) and domain_block
 * (lib.rs:78-88, the contracted sums as f64).  Bucket b = clas*24 + var_class;
 * an empty bucket is replaced by the first non-empty one in (clas,
 * var_class) row-major order (fisher.rs:42-54).
 */
or_fisher_index *or_fisher_index_build(const int32_t *contracted, uint32_t w, uint32_t h,
                                       uint32_t size, uint32_t shift)
{
    or_fisher_index *ix = (or_fisher_index *)calloc(1, sizeof *ix);
    ix->size = size;
    ix->shift = shift;
    if (shift == 0) shift = 1;
    uint32_t P = 0, nr = 0, nc = 0;
    if (h >= 2 * size && w >= 2 * size) {
        nr = (h - 2 * size) / shift + 1;
        nc = (w - 2 * size) / shift + 1;
        P = nr * nc;
    }
    ix->P = P;
    uint8_t *cls = (uint8_t *)malloc(P ? P : 1), *isos = (uint8_t *)malloc(P ? P : 1);
    double *blk = (double *)malloc(sizeof(double) * size * size);
    size_t cstride = w / 2;
    uint32_t counts[OR_FISHER_BUCKETS] = {0};
    for (uint32_t p = 0; p < P; p++) {
        uint32_t dom_row = (p / nc) * shift, dom_col = (p % nc) * shift;
        size_t dr = dom_row / 2, dc = dom_col / 2;
        for (uint32_t u = 0; u < size; u++)
            for (uint32_t v = 0; v < size; v++)
                blk[u * size + v] = (double)contracted[(dr + u) * cstride + dc + v];
        uint32_t iso, clas, vc;
        or_fisher_classify(blk, size, &iso, &clas, &vc);
        cls[p] = (uint8_t)(clas * 24 + vc);
        isos[p] = (uint8_t)iso;
        counts[cls[p]]++;
    }
    ix->entries = (or_fisher_entry *)malloc(sizeof(or_fisher_entry) * (P ? P : 1));
    uint32_t off = 0;
    for (int b = 0; b < OR_FISHER_BUCKETS; b++) {
        ix->start[b] = off;
        ix->count[b] = counts[b];
        off += counts[b];
    }
    uint32_t fill[OR_FISHER_BUCKETS];
    memcpy(fill, ix->start, sizeof fill);
    for (uint32_t p = 0; p < P; p++) {
        or_fisher_entry *e = &ix->entries[fill[cls[p]]++];
        e->dom_row = (p / nc) * shift;
        e->dom_col = (p % nc) * shift;
        e->dom_iso = isos[p];
    }
    int first = -1;
    for (int b = 0; b < OR_FISHER_BUCKETS && first < 0; b++)
        if (ix->count[b]) first = b;
    for (int b = 0; b < OR_FISHER_BUCKETS; b++) {
        ix->own_count[b] = ix->count[b];
        if (!ix->count[b] && first >= 0) {
            ix->start[b] = ix->start[first];
            ix->count[b] = ix->count[first];
        }
    }
    free(cls);
    free(isos);
    free(blk);
    return ix;
}

void or_fisher_index_free(or_fisher_index *ix)
{
    if (!ix) return;
    free(ix->entries);
    free(ix);
}

/* fisher.rs:58-78 candidates: the range block (lib.rs:115-118) classified. */
uint32_t or_fisher_range_class(const uint8_t *px, uint32_t w, uint32_t row, uint32_t col,
                               uint32_t size, uint32_t *isom)
{
    double *blk = (double *)malloc(sizeof(double) * size * size);
    for (uint32_t i = 0; i < size; i++)
        for (uint32_t j = 0; j < size; j++)
            blk[i * size + j] = (double)px[(size_t)(row + i) * w + col + j];
    uint32_t iso, clas, vc;
    or_fisher_classify(blk, size, &iso, &clas, &vc);
    free(blk);
    *isom = iso;
    return clas * 24 + vc;
}

/*
 * lib.rs:211-297 search_block_fitted over Fisher's candidate list: every
 * candidate fitted with fit_f64 in bucket order, first strict minimum of
 * rms kept, one eval per candidate; no candidates -> no fit, zero evals.
 */
void or_fisher_search(const uint8_t *px, uint32_t w, uint32_t h, const int32_t *contracted,
                      const or_fisher_index *ix, uint32_t row, uint32_t col, double max_alfa,
                      uint32_t bits_alfa, uint32_t bits_beta, or_candidate *out, uint32_t *n_out,
                      uint64_t *evals_out)
{
    (void)h;
    const uint32_t size = ix->size, n2 = size * size;
    size_t cstride = w / 2;
    uint8_t *range = (uint8_t *)malloc(n2), *by_iso = (uint8_t *)malloc((size_t)8 * n2);
    int64_t t0 = 0, t2 = 0;
    for (uint32_t i = 0; i < size; i++)
        for (uint32_t j = 0; j < size; j++) {
            uint8_t r = px[(size_t)(row + i) * w + col + j];
            range[i * size + j] = r;
            t0 += r;
            t2 += (int64_t)r * r;
        }
    for (uint32_t k = 0; k < 8; k++) or_permute_range(range, k, size, by_iso + k * n2);
    uint32_t isom;
    uint32_t bucket = or_fisher_range_class(px, w, row, col, size, &isom);
    const or_fisher_entry *e = ix->entries + ix->start[bucket];
    const uint32_t n = ix->count[bucket];
    *n_out = 0;
    for (uint32_t c = 0; c < n; c++) {
        uint32_t k = MAPPING[isom][e[c].dom_iso];
        size_t dr = e[c].dom_row / 2, dc = e[c].dom_col / 2;
        or_candidate cand;
        memset(&cand, 0, sizeof cand);
        cand.m.s0 = (int64_t)n2;
        or_domain_sums(contracted, cstride, dr, dc, size, &cand.m.s1_x4, &cand.m.s2_x16);
        cand.m.t0 = t0;
        cand.m.t1_x4 = or_cross_term_permuted(contracted, cstride, dr, dc, size, by_iso + k * n2);
        cand.m.t2 = t2;
        or_fit_f64(&cand.m, max_alfa, bits_alfa, bits_beta, &cand.qalfa, &cand.qbeta, &cand.rms);
        cand.dom_row = e[c].dom_row;
        cand.dom_col = e[c].dom_col;
        cand.isometry = k;
        if (*n_out == 0 || cand.rms < out->rms) {
            *out = cand;
            *n_out = 1;
        }
    }
    *evals_out = n;
    free(range);
    free(by_iso);
}

typedef struct {
    const uint8_t *px;
    const int32_t *contracted;
    const or_fisher_index *ix;
    uint32_t w, h, bits_alfa, bits_beta;
    double max_alfa;
    const uint32_t *rows, *cols;
    uint32_t lo, hi;
    or_candidate *out;
    uint8_t *found;
    uint64_t evals;
} fisher_job;

static void *fisher_worker(void *arg)
{
    fisher_job *j = (fisher_job *)arg;
    j->evals = 0;
    for (uint32_t i = j->lo; i < j->hi; i++) {
        uint32_t n;
        uint64_t e;
        or_fisher_search(j->px, j->w, j->h, j->contracted, j->ix, j->rows[i], j->cols[i],
                         j->max_alfa, j->bits_alfa, j->bits_beta, &j->out[i], &n, &e);
        j->found[i] = (uint8_t)(n > 0);
        j->evals += e;
    }
    return NULL;
}

uint64_t or_fisher_table(const uint8_t *px, uint32_t w, uint32_t h, uint32_t size,
                         uint32_t shift, double max_alfa, uint32_t bits_alfa, uint32_t bits_beta,
                         const uint32_t *rows, const uint32_t *cols, uint32_t n, int threads,
                         or_candidate *out, uint8_t *found, uint64_t *evals_each)
{
    int32_t *contracted = (int32_t *)malloc(sizeof(int32_t) * ((size_t)(w / 2) * (h / 2) + 1));
    or_contract(px, w, h, contracted);
    or_fisher_index *ix = or_fisher_index_build(contracted, w, h, size, shift);
    if (threads < 1) threads = 1;
    if ((uint32_t)threads > n && n > 0) threads = (int)n;
    fisher_job *jobs = (fisher_job *)calloc((size_t)threads, sizeof(fisher_job));
    pthread_t *tids = (pthread_t *)calloc((size_t)threads, sizeof(pthread_t));
    for (int t = 0; t < threads; t++) {
        fisher_job *j = &jobs[t];
        j->px = px;
        j->contracted = contracted;
        j->ix = ix;
        j->w = w;
        j->h = h;
        j->bits_alfa = bits_alfa;
        j->bits_beta = bits_beta;
        j->max_alfa = max_alfa;
        j->rows = rows;
        j->cols = cols;
        j->lo = (uint32_t)((uint64_t)n * t / threads);
        j->hi = (uint32_t)((uint64_t)n * (t + 1) / threads);
        j->out = out;
        j->found = found;
    }
    for (int t = 1; t < threads; t++) pthread_create(&tids[t], NULL, fisher_worker, &jobs[t]);
    fisher_worker(&jobs[0]);
    uint64_t evals = jobs[0].evals;
    for (int t = 1; t < threads; t++) {
        pthread_join(tids[t], NULL);
        evals += jobs[t].evals;
    }
    if (evals_each) {
        for (uint32_t i = 0; i < n; i++) {
            uint32_t isom;
            evals_each[i] = ix->count[or_fisher_range_class(px, w, rows[i], cols[i], size, &isom)];
        }
    }
    free(jobs);
    free(tids);
    or_fisher_index_free(ix);
    free(contracted);
    return evals;
}

/* Fisher as a search-provider callback: one index per (size, stride), built on first use. */
or_fisher_provider *or_fisher_provider_new(const uint8_t *px, uint32_t w, uint32_t h,
                                           double max_alfa, uint32_t bits_alfa,
                                           uint32_t bits_beta)
{
    or_fisher_provider *p = (or_fisher_provider *)calloc(1, sizeof *p);
    p->base = or_provider_new(px, w, h, max_alfa, bits_alfa, bits_beta);
    return p;
}

void or_fisher_provider_free(or_fisher_provider *p)
{
    if (!p) return;
    for (int i = 0; i < OR_FISHER_SLOTS; i++) or_fisher_index_free(p->ix[i]);
    or_provider_free(p->base);
    free(p);
}

/* IndexedSearchProvider keeps the trait's default search_candidates
 * (encode.rs:572-578): the winner alone, whatever `max` is. */
int or_fisher_cb(void *user, uint32_t row, uint32_t col, uint32_t size, uint32_t shift,
                 uint32_t max, void *out, uint32_t *n_out, uint64_t *evals)
{
    (void)max;
    or_fisher_provider *p = (or_fisher_provider *)user;
    const or_provider *b = p->base;
    or_fisher_index *ix = NULL;
    for (int i = 0; i < OR_FISHER_SLOTS; i++) {
        if (p->ix[i] && p->ix[i]->size == size && p->ix[i]->shift == shift) {
            ix = p->ix[i];
            break;
        }
        if (!p->ix[i]) {
            ix = p->ix[i] = or_fisher_index_build(b->contracted, b->w, b->h, size, shift);
            break;
        }
    }
    if (!ix) return -1;
    or_fisher_search(b->px, b->w, b->h, b->contracted, ix, row, col, b->max_alfa, b->bits_alfa,
                     b->bits_beta, (or_candidate *)out, n_out, evals);
    return 0;
}

/* metrics.rs:377-417 ycbcr: BT.601 full range, f64, round half away from zero, clamp. */
static uint8_t to_u8(double x)
{
    x = round(x);
    if (x < 0.0) x = 0.0;
    if (x > 255.0) x = 255.0;
    return (uint8_t)x;
}

void or_ycbcr(const uint8_t *rgb, size_t n, uint8_t *y, uint8_t *cb, uint8_t *cr)
{
    for (size_t i = 0; i < n; i++) {
        double r = rgb[3 * i], g = rgb[3 * i + 1], b = rgb[3 * i + 2];
        y[i] = to_u8(0.299 * r + 0.587 * g + 0.114 * b);
        cb[i] = to_u8(128.0 - 0.168736 * r - 0.331264 * g + 0.5 * b);
        cr[i] = to_u8(128.0 + 0.5 * r - 0.418688 * g - 0.081312 * b);
    }
}

/* color.rs:60-78 downsample_box: 2x2 sums + 2, /4; the last row/column repeats on odd sides. */
void or_downsample_box(const uint8_t *px, uint32_t w, uint32_t h, uint8_t *out)
{
    uint32_t ow = (w + 1) / 2, oh = (h + 1) / 2;
    for (uint32_t oy = 0; oy < oh; oy++)
        for (uint32_t ox = 0; ox < ow; ox++) {
            uint32_t x0 = 2 * ox, y0 = 2 * oy;
            uint32_t x1 = x0 + 1 < w ? x0 + 1 : w - 1, y1 = y0 + 1 < h ? y0 + 1 : h - 1;
            uint32_t s = (uint32_t)px[(size_t)y0 * w + x0] + px[(size_t)y0 * w + x1] +
                         px[(size_t)y1 * w + x0] + px[(size_t)y1 * w + x1];
            out[(size_t)oy * ow + ox] = (uint8_t)((s + 2) / 4);
        }
}
