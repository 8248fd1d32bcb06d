// Fisher classified search (crates/mars-search/src/fisher.rs) on the GPU.
//
// Per (plane, range size, stride) table, after the pool (contracted plane,
// summed-area tables), k_possums (window sums S1, S2), k_ranges (each range
// under the 8 isometries) and k_rstats (t0, t2):
//   k_fisher_domains  every domain position classified from its quadrant
//                     sums and sums of squares, read off the summed-area
//                     tables (fisher.rs:33-38);
//   k_fisher_ranges   every range block classified the same way from its
//                     pixels (fisher.rs:58-61);
//   k_fisher_buckets  one block per table: the 72 bucket sizes, their
//                     offsets, the positions grouped by bucket, and every
//                     empty bucket aliased to the first non-empty one in
//                     (class, variance class) order (fisher.rs:42-54);
//   k_fisher_search   one wave per range: every candidate of the range's
//                     bucket with isometry MAPPING[range iso][domain iso]
//                     (fisher.rs:62-75), its exact integer t1 and the
//                     verbatim binary64 fit, the first strict minimum of rms
//                     in bucket order kept (lib.rs:254-286).
//
// Bucket order is domain-position order in the reference, and a position
// appears at most once per bucket, so "first strict minimum in bucket
// order" is the minimum of (rms, position index): the grouped lists need no
// particular order inside a bucket.  Evals = the bucket's size
// (lib.rs:270, one per candidate).

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "device.h"
#include "fisher_class.h"
#include "kernels.h"

#pragma clang fp contract(off)

namespace mars {

using fisher::kBuckets;

// Work item (table, block of 256 positions), one thread per position.
__global__ __launch_bounds__(256) void k_fisher_domains(const TableDesc *__restrict__ tabs,
                                                        const Work2 *__restrict__ work)
{
    const Work2 wk = work[blockIdx.x];
    const TableG T = view(tabs[wk.x]);
    const uint32_t p = wk.y * 256 + threadIdx.x;
    if (p >= T.P) return;
    const uint32_t W1 = T.wc + 1, h = T.size / 2;
    const uint32_t pi = p / T.npc, pj = p - pi * T.npc;
    const uint32_t dr = (pi * T.shift) >> 1, dc = (pj * T.shift) >> 1;
    // The 3 x 3 summed-area corners of the four h x h quadrants.
    int64_t c1[3][3], c2[3][3];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
            const size_t idx = (size_t)(dr + i * h) * W1 + dc + j * h;
            c1[i][j] = T.sat1[idx];
            c2[i][j] = T.sat2[idx];
        }
    int64_t sum[4], sq[4];
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const int i = q >> 1, j = q & 1;
        sum[q] = c1[i + 1][j + 1] - c1[i][j + 1] - c1[i + 1][j] + c1[i][j];
        sq[q] = c2[i + 1][j + 1] - c2[i][j + 1] - c2[i + 1][j] + c2[i][j];
    }
    T.fcls[p] = (uint16_t)fisher::classify(sum, sq, h);
}

__device__ __forceinline__ int32_t wave_sum(int32_t x)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) x += __shfl_xor(x, off, 64);
    return x;
}

// Work item (table, block of 4 ranges), one wave per range: quadrant sums
// and sums of squares of the range's pixels, then its class.
__global__ __launch_bounds__(256) void k_fisher_ranges(const TableDesc *__restrict__ tabs,
                                                       const Work2 *__restrict__ work)
{
    const Work2 wk = work[blockIdx.x];
    const TableG T = view(tabs[wk.x]);
    const uint32_t r = wk.y * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (r >= T.nr) return;
    const uint32_t size = T.size, h = size / 2, K = size * size, lg = __builtin_ctz(size);
    const uint32_t row = T.rows[r], col = T.cols[r];
    int32_t s[4] = {0, 0, 0, 0}, q2[4] = {0, 0, 0, 0};
    for (uint32_t e = lane; e < K; e += 64) {
        const uint32_t u = e >> lg, v = e & (size - 1);
        const int32_t x = T.px[(size_t)(row + u) * T.w + col + v];
        const int q = (u >= h ? 2 : 0) + (v >= h ? 1 : 0);
#pragma unroll
        for (int j = 0; j < 4; j++) {
            s[j] += q == j ? x : 0;
            q2[j] += q == j ? x * x : 0;
        }
    }
    int64_t sum[4], sq[4];
#pragma unroll
    for (int j = 0; j < 4; j++) {
        sum[j] = wave_sum(s[j]);
        sq[j] = wave_sum(q2[j]);
    }
    if (lane == 0) T.frcls[r] = (uint16_t)fisher::classify(sum, sq, h);
}

// Work item (table, 0), one 1024-thread block per table: bucket sizes,
// offsets (empty buckets aliased to the first non-empty one) and the
// positions grouped by bucket, each entry p * 8 + the position's isometry.
__global__ __launch_bounds__(1024) void k_fisher_buckets(const TableDesc *__restrict__ tabs,
                                                         const Work2 *__restrict__ work)
{
    __shared__ uint32_t cnt[kBuckets], cur[kBuckets];
    const TableG T = view(tabs[work[blockIdx.x].x]);
    const uint32_t t = threadIdx.x, P = T.P;
    if (t < kBuckets) cnt[t] = 0;
    __syncthreads();
    for (uint32_t p = t; p < P; p += 1024) atomicAdd(&cnt[T.fcls[p] >> 3], 1u);
    __syncthreads();
    if (t == 0) {
        uint32_t off = 0, first = kBuckets;
        for (uint32_t b = 0; b < kBuckets; b++) {
            cur[b] = off;
            off += cnt[b];
            if (first == kBuckets && cnt[b]) first = b;
        }
        for (uint32_t b = 0; b < kBuckets; b++) {
            const uint32_t src = cnt[b] || first == kBuckets ? b : first;
            T.fbuck[b] = cur[src];
            T.fbuck[kBuckets + b] = cnt[src];
        }
    }
    __syncthreads();
    for (uint32_t p = t; p < P; p += 1024) {
        const uint32_t c = T.fcls[p];
        const uint32_t slot = atomicAdd(&cur[c >> 3], 1u);
        T.flist[slot] = p * 8 + (c & 7);
    }
}

// Work item (table, block of 4 ranges), one wave per range.  The wave
// takes 64 / L candidates at a time, L = min(size, 16) lanes each; lane j
// of a group correlates columns j, j + L, ... of every window row (a
// coalesced row segment per load), and the group sums its t1 by shuffles.
// Each candidate is fitted with the reference's binary64 fit and the
// minimum by (rms, position) kept.
__global__ __launch_bounds__(256) void k_fisher_search(const TableDesc *__restrict__ tabs,
                                                       const Work2 *__restrict__ work, Common cm)
{
    const Work2 wk = work[blockIdx.x];
    const TableG T = view(tabs[wk.x]);
    const uint32_t r = wk.y * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (r >= T.nr) return;
    const uint32_t size = T.size, K = size * size;
    const uint32_t L = size < 16 ? size : 16, per = 64 / L, g = lane / L, li = lane % L;
    const uint32_t code = T.frcls[r], b = code >> 3, isom = code & 7;
    const uint32_t start = T.fbuck[b], cnt = T.fbuck[kBuckets + b];
    const int64_t t0 = T.rstat[2 * r], t2 = T.rstat[2 * r + 1];
    const GLB int32_t *d = T.d;
    const uint32_t wc = T.wc;
    double best = __longlong_as_double(0x7ff0000000000000ll);
    uint32_t bp = 0xffffffffu, bk = 0;
    int32_t bt1 = 0;
    for (uint32_t base = 0; base < cnt; base += per) {
        const uint32_t c = base + g;
        const bool valid = c < cnt;
        uint32_t p = 0, k = 0;
        int32_t t1 = 0;
        if (valid) {
            const uint32_t e = T.flist[start + c];
            p = e >> 3;
            k = fisher::kMapping[isom][e & 7];
            const uint32_t pi = p / T.npc, pj = p - pi * T.npc;
            const uint32_t dr = (pi * T.shift) >> 1, dc = (pj * T.shift) >> 1;
            const GLB int8_t *rk = T.rng8 + ((size_t)r * 8 + k) * T.kmem;
            for (uint32_t u = 0; u < size; u++) {
                const GLB int32_t *drow = d + (size_t)(dr + u) * wc + dc;
                for (uint32_t v = li; v < size; v += L)
                    t1 += ((int32_t)rk[u * size + v] + 128) * drow[v];
            }
        }
        for (uint32_t off = L / 2; off >= 1; off >>= 1) t1 += __shfl_xor(t1, (int)off, 64);
        if (valid) {
            uint32_t qa, qb;
            double rms, sum;
            fit_f64((int64_t)K, T.posx[2 * p], T.posx[2 * p + 1], t0, t1, t2, cm.max_alfa,
                    cm.bits_alfa, cm.bits_beta, &qa, &qb, &rms, &sum);
            if (rms < best || (rms == best && p < bp)) {
                best = rms;
                bp = p;
                bk = k;
                bt1 = t1;
            }
        }
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        const double ob = __shfl_xor(best, off, 64);
        const uint32_t op = __shfl_xor(bp, off, 64), ok = __shfl_xor(bk, off, 64);
        const int32_t ot1 = __shfl_xor(bt1, off, 64);
        if (ob < best || (ob == best && op < bp)) {
            best = ob;
            bp = op;
            bk = ok;
            bt1 = ot1;
        }
    }
    if (lane != 0) return;
    T.fevals[r] = cnt;
    if (cnt == 0) {
        st_cand(T.fout, r, mars_hip_candidate{});
        T.ffound[r] = 0;
        return;
    }
    mars_hip_candidate w = {};
    const uint32_t pi = bp / T.npc, pj = bp - pi * T.npc;
    w.dom_row = pi * T.shift;
    w.dom_col = pj * T.shift;
    w.isometry = bk;
    w.s0 = (int64_t)K;
    w.s1_x4 = T.posx[2 * bp];
    w.s2_x16 = T.posx[2 * bp + 1];
    w.t0 = t0;
    w.t1_x4 = bt1;
    w.t2 = t2;
    uint32_t qa, qb;
    double rms, sum;
    fit_f64(w.s0, w.s1_x4, w.s2_x16, w.t0, w.t1_x4, w.t2, cm.max_alfa, cm.bits_alfa, cm.bits_beta,
            &qa, &qb, &rms, &sum);
    w.qalfa = qa;
    w.qbeta = qb;
    w.rms = rms;
    st_cand(T.fout, r, w);
    T.ffound[r] = 1;
}

void launch_fisher(const TableDesc *t, const Work2 *dom_work, uint32_t n_dom,
                   const Work2 *rng_work, uint32_t n_rng, const Work2 *tab_work, uint32_t n_tab,
                   const Common &c, hipStream_t st)
{
    if (n_dom) hipLaunchKernelGGL(k_fisher_domains, dim3(n_dom), dim3(256), 0, st, t, dom_work);
    if (n_rng) hipLaunchKernelGGL(k_fisher_ranges, dim3(n_rng), dim3(256), 0, st, t, rng_work);
    if (n_tab) hipLaunchKernelGGL(k_fisher_buckets, dim3(n_tab), dim3(1024), 0, st, t, tab_work);
    if (n_rng) hipLaunchKernelGGL(k_fisher_search, dim3(n_rng), dim3(256), 0, st, t, rng_work, c);
}

} // namespace mars
