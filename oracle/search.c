/*
 * CPU oracle: restatement of the reference exhaustive search.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Reference: daleobrien/mars crates/mars-codec/src/encode.rs,
 *            crates/mars-codec/src/isometry.rs,
 *            crates/mars-simd/src/moments.rs (scalar references).
 */
#include "oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* isometry.rs:7-19 */
void or_iso_map(uint32_t k, uint32_t u, uint32_t v, uint32_t size, uint32_t *i, uint32_t *j)
{
    switch (k) {
    case 0: *i = u; *j = v; break;                       /* identity */
    case 1: *i = size - 1 - v; *j = u; break;            /* left rotate 90 */
    case 2: *i = v; *j = size - 1 - u; break;            /* right rotate 90 */
    case 3: *i = size - 1 - u; *j = size - 1 - v; break; /* rotate 180 */
    case 4: *i = u; *j = size - 1 - v; break;            /* vertical mirror */
    case 5: *i = size - 1 - u; *j = v; break;            /* horizontal mirror */
    case 6: *i = v; *j = u; break;                       /* first diagonal */
    default: *i = size - 1 - v; *j = size - 1 - u; break; /* second diagonal */
    }
}

/* encode.rs:473-501: D(i,j) = sum of the 2x2 block at (2i,2j); odd last row/col dropped. */
void or_contract(const uint8_t *px, uint32_t w, uint32_t h, int32_t *out)
{
    uint32_t stride = w / 2, rows = h / 2;
    for (uint32_t i = 0; i < rows; i++) {
        uint32_t r0 = 2 * i;
        for (uint32_t j = 0; j < stride; j++) {
            uint32_t c0 = 2 * j;
            out[(size_t)i * stride + j] = (int32_t)px[(size_t)r0 * w + c0] +
                                          (int32_t)px[(size_t)r0 * w + c0 + 1] +
                                          (int32_t)px[(size_t)(r0 + 1) * w + c0] +
                                          (int32_t)px[(size_t)(r0 + 1) * w + c0 + 1];
        }
    }
}

/* moments.rs:21-44 */
void or_domain_sums(const int32_t *plane, size_t stride, size_t dr, size_t dc, size_t size,
                    int64_t *s1, int64_t *s2)
{
    int64_t a = 0, b = 0;
    for (size_t u = 0; u < size; u++) {
        const int32_t *row = plane + (dr + u) * stride + dc;
        for (size_t v = 0; v < size; v++) {
            int64_t d = row[v];
            a += d;
            b += d * d;
        }
    }
    *s1 = a;
    *s2 = b;
}

/* encode.rs:833-842: range_k[u*size+v] = range[map(k,u,v)] */
void or_permute_range(const uint8_t *range, uint32_t k, uint32_t size, uint8_t *out)
{
    for (uint32_t u = 0; u < size; u++)
        for (uint32_t v = 0; v < size; v++) {
            uint32_t i, j;
            or_iso_map(k, u, v, size, &i, &j);
            out[u * size + v] = range[i * size + j];
        }
}

/* moments.rs:157-176 */
int64_t or_cross_term_permuted(const int32_t *plane, size_t stride, size_t dr, size_t dc,
                               size_t size, const uint8_t *range_k)
{
    int64_t acc = 0;
    for (size_t u = 0; u < size; u++) {
        const int32_t *row = plane + (dr + u) * stride + dc;
        for (size_t v = 0; v < size; v++)
            acc += (int64_t)range_k[u * size + v] * (int64_t)row[v];
    }
    return acc;
}

/* encode.rs:446-448: (0.5 + x).trunc().clamp(0, max) as u32 */
uint32_t or_quantise_f64(double x, uint32_t max)
{
    double t = trunc(0.5 + x);
    if (t < 0.0) t = 0.0;
    if (t > (double)max) t = (double)max;
    return (uint32_t)t;
}

static uint32_t quantise_f32(float x, uint32_t max)
{
    float t = truncf(0.5f + x);
    if (t < 0.0f) t = 0.0f;
    if (t > (float)max) t = (float)max;
    return (uint32_t)t;
}

/* encode.rs:456-460 */
uint32_t or_best_beta(int64_t range_sum, int64_t pixel_count, uint32_t bits_beta)
{
    double mean = (double)range_sum / (double)pixel_count;
    uint32_t max_qbeta = (1u << bits_beta) - 1;
    return or_quantise_f64(mean / 255.0 * (double)max_qbeta, max_qbeta);
}

/* encode.rs:247-289, in the same operation order. */
void or_fit_f64(const or_moments *m, double max_alfa, uint32_t bits_alfa, uint32_t bits_beta,
                uint32_t *qalfa_out, uint32_t *qbeta_out, double *rms_out)
{
    double s0 = (double)m->s0;
    double s1 = (double)m->s1_x4 / 4.0;
    double s2 = (double)m->s2_x16 / 16.0;
    double t0 = (double)m->t0;
    double t1 = (double)m->t1_x4 / 4.0;
    double t2 = (double)m->t2;

    double det = s0 * s2 - s1 * s1;
    double alfa = det == 0.0 ? 0.0 : (s0 * t1 - s1 * t0) / det;
    if (alfa < 0.0) alfa = 0.0;

    uint32_t max_qalfa = (1u << bits_alfa) - 1;
    uint32_t qalfa = or_quantise_f64(alfa / max_alfa * (double)(1u << bits_alfa), max_qalfa);
    double alfa2 = (double)qalfa / (double)(1u << bits_alfa) * max_alfa;

    double beta = (t0 - alfa2 * s1) / s0;
    if (alfa2 > 0.0) beta += alfa2 * 255.0;
    uint32_t max_qbeta = (1u << bits_beta) - 1;
    uint32_t qbeta =
        or_quantise_f64(beta / ((1.0 + fabs(alfa2)) * 255.0) * (double)max_qbeta, max_qbeta);
    double beta2 = (double)qbeta / (double)max_qbeta * ((1.0 + fabs(alfa2)) * 255.0);
    if (alfa2 > 0.0) beta2 -= alfa2 * 255.0;

    double sum = t2 - 2.0 * alfa2 * t1 - 2.0 * beta2 * t0 + alfa2 * alfa2 * s2 +
                 2.0 * alfa2 * beta2 * s1 + s0 * beta2 * beta2;
    double q = sum / s0;
    /* f64::max(q, 0.0): q is never NaN here (det != 0 guards the division). */
    *rms_out = sqrt(q > 0.0 ? q : 0.0);
    *qalfa_out = qalfa;
    *qbeta_out = qbeta;
}

/* encode.rs:295-337, binary32 twin. */
void or_fit_f32(const or_moments *m, float max_alfa, uint32_t bits_alfa, uint32_t bits_beta,
                uint32_t *qalfa_out, uint32_t *qbeta_out, float *rms_out)
{
    float s0 = (float)m->s0;
    float s1 = (float)m->s1_x4 / 4.0f;
    float s2 = (float)m->s2_x16 / 16.0f;
    float t0 = (float)m->t0;
    float t1 = (float)m->t1_x4 / 4.0f;
    float t2 = (float)m->t2;

    float det = s0 * s2 - s1 * s1;
    float alfa = det == 0.0f ? 0.0f : (s0 * t1 - s1 * t0) / det;
    if (alfa < 0.0f) alfa = 0.0f;

    uint32_t max_qalfa = (1u << bits_alfa) - 1;
    uint32_t qalfa = quantise_f32(alfa / max_alfa * (float)(1u << bits_alfa), max_qalfa);
    float alfa2 = (float)qalfa / (float)(1u << bits_alfa) * max_alfa;

    float beta = (t0 - alfa2 * s1) / s0;
    if (alfa2 > 0.0f) beta += alfa2 * 255.0f;
    uint32_t max_qbeta = (1u << bits_beta) - 1;
    uint32_t qbeta =
        quantise_f32(beta / ((1.0f + fabsf(alfa2)) * 255.0f) * (float)max_qbeta, max_qbeta);
    float beta2 = (float)qbeta / (float)max_qbeta * ((1.0f + fabsf(alfa2)) * 255.0f);
    if (alfa2 > 0.0f) beta2 -= alfa2 * 255.0f;

    float sum = t2 - 2.0f * alfa2 * t1 - 2.0f * beta2 * t0 + alfa2 * alfa2 * s2 +
                2.0f * alfa2 * beta2 * s1 + s0 * beta2 * beta2;
    float q = sum / s0;
    *rms_out = sqrtf(q > 0.0f ? q : 0.0f);
    *qalfa_out = qalfa;
    *qbeta_out = qbeta;
}

/* encode.rs:676-685: strict `<` insertion, at most k kept. */
static void retain_top_k(or_candidate *top, uint32_t *n, const or_candidate *c, uint32_t k)
{
    uint32_t pos = *n;
    for (uint32_t i = 0; i < *n; i++)
        if (c->rms < top[i].rms) {
            pos = i;
            break;
        }
    if (pos < k) {
        uint32_t last = *n < k ? *n : k - 1;
        for (uint32_t i = last; i > pos; i--) top[i] = top[i - 1];
        top[pos] = *c;
        if (*n < k) (*n)++;
    }
}

/* encode.rs:695-788 */
void or_search_top_k(const uint8_t *px, uint32_t w, uint32_t h, const int32_t *contracted,
                     uint32_t row, uint32_t col, uint32_t size, uint32_t shift, double max_alfa,
                     uint32_t bits_alfa, uint32_t bits_beta, uint32_t k, or_candidate *out,
                     uint32_t *n_out, uint64_t *evals_out)
{
    *n_out = 0;
    *evals_out = 0;
    if (k == 0) k = 1;
    size_t cstride = w / 2;
    uint32_t n2 = size * size;
    uint8_t *range = (uint8_t *)malloc(n2);
    uint8_t *by_iso = (uint8_t *)malloc((size_t)8 * n2);
    int64_t t0 = 0, t2 = 0;
    for (uint32_t i = 0; i < size; i++)
        for (uint32_t j = 0; j < size; j++) {
            uint8_t r = px[(size_t)(row + i) * w + col + j];
            range[i * size + j] = r;
            t0 += r;
            t2 += (int64_t)r * r;
        }
    int64_t s0 = (int64_t)size * size;
    for (uint32_t kk = 0; kk < 8; kk++) or_permute_range(range, kk, size, by_iso + kk * n2);

    if (h >= 2 * size && w >= 2 * size) {
        uint32_t max_dom_row = h - 2 * size, max_dom_col = w - 2 * size;
        uint64_t evals = 0;
        for (uint32_t dom_row = 0; dom_row <= max_dom_row; dom_row += shift) {
            for (uint32_t dom_col = 0; dom_col <= max_dom_col; dom_col += shift) {
                size_t dr = dom_row / 2, dc = dom_col / 2;
                int64_t s1, s2;
                or_domain_sums(contracted, cstride, dr, dc, size, &s1, &s2);
                for (uint32_t kk = 0; kk < 8; kk++) {
                    or_candidate c;
                    memset(&c, 0, sizeof c);
                    c.m.s0 = s0;
                    c.m.s1_x4 = s1;
                    c.m.s2_x16 = s2;
                    c.m.t0 = t0;
                    c.m.t1_x4 =
                        or_cross_term_permuted(contracted, cstride, dr, dc, size, by_iso + kk * n2);
                    c.m.t2 = t2;
                    or_fit_f64(&c.m, max_alfa, bits_alfa, bits_beta, &c.qalfa, &c.qbeta, &c.rms);
                    c.dom_row = dom_row;
                    c.dom_col = dom_col;
                    c.isometry = kk;
                    evals++;
                    retain_top_k(out, n_out, &c, k);
                }
            }
        }
        *evals_out = evals;
    }
    free(range);
    free(by_iso);
}

typedef struct {
    const uint8_t *px;
    const int32_t *contracted;
    uint32_t w, h, size, shift, bits_alfa, bits_beta, k;
    double max_alfa;
    const uint32_t *rows, *cols;
    uint32_t lo, hi;
    or_candidate *out; /* k per range */
    uint8_t *found;    /* candidates kept per range */
    uint64_t evals;
} table_job;

static void *table_worker(void *arg)
{
    table_job *j = (table_job *)arg;
    j->evals = 0;
    for (uint32_t i = j->lo; i < j->hi; i++) {
        uint32_t n;
        uint64_t e;
        or_search_top_k(j->px, j->w, j->h, j->contracted, j->rows[i], j->cols[i], j->size, j->shift,
                        j->max_alfa, j->bits_alfa, j->bits_beta, j->k, &j->out[(size_t)i * j->k],
                        &n, &e);
        j->found[i] = (uint8_t)n;
        j->evals += e;
    }
    return NULL;
}

uint64_t or_search_table(const uint8_t *px, uint32_t w, uint32_t h, uint32_t size,
                         uint32_t shift, double max_alfa, uint32_t bits_alfa,
                         uint32_t bits_beta, const uint32_t *rows, const uint32_t *cols,
                         uint32_t n, int threads, or_candidate *out, uint8_t *found)
{
    return or_search_table_k(px, w, h, size, shift, max_alfa, bits_alfa, bits_beta, rows, cols,
                             n, threads, 1, out, found);
}

uint64_t or_search_table_k(const uint8_t *px, uint32_t w, uint32_t h, uint32_t size,
                           uint32_t shift, double max_alfa, uint32_t bits_alfa,
                           uint32_t bits_beta, const uint32_t *rows, const uint32_t *cols,
                           uint32_t n, int threads, uint32_t k, or_candidate *out,
                           uint8_t *found)
{
    if (k == 0) k = 1;
    int32_t *contracted = (int32_t *)malloc(sizeof(int32_t) * ((size_t)(w / 2) * (h / 2) + 1));
    or_contract(px, w, h, contracted);
    if (threads < 1) threads = 1;
    if ((uint32_t)threads > n && n > 0) threads = (int)n;
    table_job *jobs = (table_job *)calloc((size_t)threads, sizeof(table_job));
    pthread_t *tids = (pthread_t *)calloc((size_t)threads, sizeof(pthread_t));
    for (int t = 0; t < threads; t++) {
        table_job *j = &jobs[t];
        j->px = px;
        j->contracted = contracted;
        j->w = w;
        j->h = h;
        j->size = size;
        j->shift = shift;
        j->bits_alfa = bits_alfa;
        j->bits_beta = bits_beta;
        j->k = k;
        j->max_alfa = max_alfa;
        j->rows = rows;
        j->cols = cols;
        j->lo = (uint32_t)((uint64_t)n * t / threads);
        j->hi = (uint32_t)((uint64_t)n * (t + 1) / threads);
        j->out = out;
        j->found = found;
    }
    for (int t = 1; t < threads; t++) pthread_create(&tids[t], NULL, table_worker, &jobs[t]);
    table_worker(&jobs[0]);
    uint64_t evals = jobs[0].evals;
    for (int t = 1; t < threads; t++) {
        pthread_join(tids[t], NULL);
        evals += jobs[t].evals;
    }
    free(jobs);
    free(tids);
    free(contracted);
    return evals;
}

or_provider *or_provider_new(const uint8_t *px, uint32_t w, uint32_t h, double max_alfa,
                             uint32_t bits_alfa, uint32_t bits_beta)
{
    or_provider *p = (or_provider *)calloc(1, sizeof *p);
    uint8_t *copy = (uint8_t *)malloc((size_t)w * h);
    memcpy(copy, px, (size_t)w * h);
    p->px = copy;
    p->w = w;
    p->h = h;
    p->contracted = (int32_t *)malloc(sizeof(int32_t) * ((size_t)(w / 2) * (h / 2) + 1));
    or_contract(copy, w, h, p->contracted);
    p->max_alfa = max_alfa;
    p->bits_alfa = bits_alfa;
    p->bits_beta = bits_beta;
    return p;
}

void or_provider_free(or_provider *p)
{
    if (!p) return;
    free((void *)p->px);
    free(p->contracted);
    free(p);
}

/* ExhaustiveSearch::search_candidates (encode.rs:598-601): the best `max`. */
int or_search_cb(void *user, uint32_t row, uint32_t col, uint32_t size, uint32_t shift,
                 uint32_t max, void *out, uint32_t *n_out, uint64_t *evals)
{
    const or_provider *p = (const or_provider *)user;
    or_search_top_k(p->px, p->w, p->h, p->contracted, row, col, size, shift, p->max_alfa,
                    p->bits_alfa, p->bits_beta, max ? max : 1, (or_candidate *)out, n_out, evals);
    return 0;
}
