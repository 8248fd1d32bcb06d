// Fisher classified search (crates/mars-search/src/fisher.rs) on the GPU.
//
// Per (plane, range size, stride) table, after the pool (contracted plane,
// summed-area tables), k_possums (window sums S1, S2) and k_ranges (each
// range under the 8 isometries, and its sums t0, t2):
//   k_fisher_domains  every domain position classified from its quadrant
//                     sums and sums of squares, read off the summed-area
//                     tables (fisher.rs:33-38);
//   k_fisher_ranges   every range block classified the same way from its
//                     pixels (fisher.rs:58-61);
//   k_fisher_buckets  one block per table: the 72 bucket sizes, their
//                     offsets, the positions grouped by bucket, and every
//                     empty bucket aliased to the first non-empty one in
//                     (class, variance class) order (fisher.rs:42-54);
//   k_fisher_search   one wave per range, one lane per candidate of the
//                     range's bucket, isometry MAPPING[range iso][domain iso]
//                     (fisher.rs:62-75): exact integer t1 by i8 dot products
//                     on the split planes, the verbatim binary64 fit (of
//                     the candidates a least-squares bound cannot rule out), the
//                     first strict minimum of rms in bucket order kept
//                     (lib.rs:254-286).
//
// Bucket order is domain-position order in the reference, and a position
// appears at most once per bucket, so "first strict minimum in bucket
// order" is the minimum of (rms, position index): the grouped lists need no
// particular order inside a bucket.  Evals = the bucket's size
// (lib.rs:269, one per candidate).

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
// positions grouped by bucket, each entry (the position's slot in the
// search-order records) * 8 + the position's isometry.
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
        T.flist[slot] = (T.iperm ? T.iperm[p] : p) * 8 + (c & 7);
    }
}

// One row of S bytes of a window (or of a permuted range) as S/4 packed
// dwords; the split planes' copies keep every row segment aligned to its
// load width (4, 8 or 16 bytes).
template <uint32_t S> struct RowW {
    int32_t w[S / 4];
};

template <uint32_t S> __device__ __forceinline__ RowW<S> load_row(const GLB int8_t *p)
{
    typedef int32_t v2 __attribute__((ext_vector_type(2)));
    typedef int32_t v4 __attribute__((ext_vector_type(4)));
    RowW<S> r;
    if constexpr (S == 4) {
        r.w[0] = *(const GLB int32_t *)p;
    } else if constexpr (S == 8) {
        const v2 x = *(const GLB v2 *)p;
        r.w[0] = x.x;
        r.w[1] = x.y;
    } else {
#pragma unroll
        for (uint32_t q = 0; q < S / 16; q++) {
            const v4 x = *(const GLB v4 *)(p + 16 * q);
            r.w[4 * q] = x.x;
            r.w[4 * q + 1] = x.y;
            r.w[4 * q + 2] = x.z;
            r.w[4 * q + 3] = x.w;
        }
    }
    return r;
}

// sum r' * hi and sum r' * lo over an S x S window of the split planes
// (D = 4 * (hi + 128) + lo, r' = r - 128): four i8 products per dot
// instruction.
template <uint32_t S>
__device__ __forceinline__ void window_dots(const GLB int8_t *hi, const GLB int8_t *lo,
                                            uint32_t pitch, const GLB int8_t *rk, int32_t &ah,
                                            int32_t &al)
{
    if constexpr (S == 2) {
#pragma unroll
        for (uint32_t u = 0; u < 2; u++)
#pragma unroll
            for (uint32_t v = 0; v < 2; v++) {
                const int32_t rv = rk[u * 2 + v];
                ah += rv * (int32_t)hi[u * pitch + v];
                al += rv * (int32_t)lo[u * pitch + v];
            }
    } else {
        constexpr uint32_t U = S <= 8 ? S : 4;
#pragma unroll U
        for (uint32_t u = 0; u < S; u++) {
            const RowW<S> h = load_row<S>(hi + u * pitch), l = load_row<S>(lo + u * pitch);
            const RowW<S> x = load_row<S>(rk + u * S);
#pragma unroll
            for (uint32_t q = 0; q < S / 4; q++) {
                ah = __builtin_amdgcn_sdot4(x.w[q], h.w[q], ah, false);
                al = __builtin_amdgcn_sdot4(x.w[q], l.w[q], al, false);
            }
        }
    }
}

// One range (one wave): lane j takes candidates j, j + 64, ... of the
// range's bucket.  Per candidate the search-order record gives the window
// offset, S1 and the position index; t1 is exact from the split planes:
//   t1 = 4 sum r'hi + sum r'lo + 512 sum r' + 128 S1;
// then the reference's binary64 fit, skipped when a certified lower bound of
// the candidate's SSE exceeds the wave's best (shared after every round of
// 64), and the minimum by (rms, position).
template <uint32_t S>
__device__ __forceinline__ void fisher_range(const TableG &T, uint32_t r, uint32_t lane,
                                             const Common &cm)
{
    constexpr uint32_t K = S * S;
    const uint32_t code = T.frcls[r], b = code >> 3, isom = code & 7;
    const uint32_t start = T.fbuck[b], cnt = T.fbuck[kBuckets + b];
    const int64_t t0 = T.rstat[2 * r], t2 = T.rstat[2 * r + 1];
    const int32_t off_r = 512 * (int32_t)(t0 - 128 * (int64_t)K); // 512 sum r'
    const GLB int8_t *rng = T.rng8 + (size_t)r * 8 * T.kmem;
    const double inf = __longlong_as_double(0x7ff0000000000000ll);
    double best = inf;
    uint32_t bp = 0xffffffffu, bk = 0;
    int32_t bt1 = 0, bs1 = 0, bs2 = 0;
    // The wave's best SSE so far, inflated so that a candidate whose exact
    // SSE exceeds it has a strictly larger binary64 rms (relative 1e-9 and
    // absolute 1e-6 n cover the rounding of the fit's sum, /s0 and sqrt).
    double tau = inf;
    const double Ad = (double)((int64_t)K * t2 - t0 * t0); // A = n t2 - t0^2, exact
    struct Cand {
        uint32_t p, k;
        int32_t t1, s1, s2;
    };
    // Sizes 4 and 8 with the binary16 plane: the range's 8 isometries as
    // binary16 r (exact) in LDS, so a candidate reads one load per window
    // row (not one per split plane) and its record (S2 in the last word):
    // the gathers' L2 requests are what bounds this kernel.
    constexpr bool F16 = S == 4 || S == 8;
    __shared__ __align__(16) _Float16 s_r[F16 ? 8 * K : 2];
    const bool f16 = F16 && T.wf != nullptr;
    if constexpr (F16) {
        if (f16) {
            for (uint32_t i = lane; i < 8 * K; i += 64)
                s_r[i] = (_Float16)(float)((int32_t)rng[(i / K) * T.kmem + i % K] + 128);
            __syncthreads();
        }
    }
    // Exact t1 of candidate c of the bucket: binary16 dot products
    // accumulated in binary32 (every partial sum an integer below
    // 64 * 255 * 1020 < 2^24), or i8 dot products on the split planes.  The
    // search-order record is {window offset, S1, p, -}, at sizes 4 and 8
    // {window offset, p, S1, S2}.
    auto gather = [&](uint32_t c) {
        const uint32_t e = T.flist[start + c];
        const uint32_t k = fisher::kMapping[isom][e & 7];
        if constexpr (F16) {
            if (f16) {
                typedef _Float16 h2 __attribute__((ext_vector_type(2)));
                const uint4 rec = ld_u4(T.posinfo, e >> 3);
                float acc = 0.0f;
                const GLB uint16_t *w = T.wf + rec.x;
                const h2 *rr = (const h2 *)(s_r + k * K);
#pragma unroll
                for (uint32_t u = 0; u < S; u++) {
                    if constexpr (S == 4) {
                        const u32x2 d = *(const GLB u32x2 *)(w + u * T.pitch);
                        const uint32_t d0 = d.x, d1 = d.y;
                        acc = __builtin_amdgcn_fdot2(rr[2 * u], __builtin_bit_cast(h2, d0), acc, false);
                        acc = __builtin_amdgcn_fdot2(rr[2 * u + 1], __builtin_bit_cast(h2, d1), acc, false);
                    } else {
                        const u32x4 d = *(const GLB u32x4 *)(w + u * T.pitch);
                        // A vector element goes through a scalar: bit_cast of the
                        // element lvalue itself reads element 0 with this hipcc.
#pragma unroll
                        for (uint32_t q = 0; q < 4; q++) {
                            const uint32_t dq = d[q];
                            acc = __builtin_amdgcn_fdot2(rr[4 * u + q], __builtin_bit_cast(h2, dq), acc, false);
                        }
                    }
                }
                return Cand{rec.y, k, (int32_t)acc, (int32_t)rec.z, (int32_t)rec.w};
            }
        }
        const uint3 rec = ld_u3(T.posinfo, e >> 3);
        int32_t ah = 0, al = 0;
        window_dots<S>(T.hi + rec.x, T.lo + rec.x, T.pitch, rng + (size_t)k * T.kmem, ah, al);
        const int32_t s1 = (int32_t)rec.y;
        return Cand{rec.z, k, 4 * ah + al + off_r + 128 * s1, s1, T.posx[2 * rec.z + 1]};
    };
    auto consider = [&](const Cand &x) {
        // Free least-squares bound with alfa2 >= 0 (encode.rs:261-267):
        // n SSE >= A - max(Ni, 0)^2 / Di, Ni = n t1 - S1 t0, Di = n S2 - S1^2
        // (x4 / x16 units).  The candidate is skipped when that exceeds
        // n tau: A Di - Ni+^2 > n tau Di, compared in binary64 with a margin
        // of 1e-12 A Di, far above the products' rounding (5.5e-16 A Di);
        // Ni^2 <= A Di, so a right side above A Di never skips.
        const int64_t Ni = (int64_t)K * x.t1 - (int64_t)x.s1 * t0;
        const int64_t Di = (int64_t)K * x.s2 - (int64_t)x.s1 * x.s1;
        if (tau < inf && Di > 0) {
            const double Dd = (double)Di, Np = Ni > 0 ? (double)Ni : 0.0;
            const double ad = Ad * Dd;
            if (ad - Np * Np > (double)K * tau * Dd + 1e-12 * ad) return;
        }
        uint32_t qa, qb;
        double rms, sum;
        fit_f64((int64_t)K, x.s1, x.s2, t0, x.t1, t2, cm.max_alfa, cm.bits_alfa, cm.bits_beta, &qa,
                &qb, &rms, &sum);
        if (rms < best || (rms == best && x.p < bp)) {
            best = rms;
            bp = x.p;
            bk = x.k;
            bt1 = x.t1;
            bs1 = x.s1;
            bs2 = x.s2;
            const double tb = (sum > 0.0 ? sum : 0.0) * (1.0 + 1e-9) + 1e-6 * (double)K;
            tau = tb < tau ? tb : tau;
        }
    };
    for (uint32_t base = 0; base < cnt; base += 64) {
        if (base + lane < cnt) consider(gather(base + lane));
        if (base + 64 < cnt) { // share the wave's best before the next round
#pragma unroll
            for (int off = 32; off >= 1; off >>= 1) {
                const double o = __shfl_xor(tau, off, 64);
                tau = o < tau ? o : tau;
            }
        }
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        const double ob = __shfl_xor(best, off, 64);
        const uint32_t op = __shfl_xor(bp, off, 64), ok = __shfl_xor(bk, off, 64);
        const int32_t ot1 = __shfl_xor(bt1, off, 64), os1 = __shfl_xor(bs1, off, 64),
                      os2 = __shfl_xor(bs2, off, 64);
        if (ob < best || (ob == best && op < bp)) {
            best = ob;
            bp = op;
            bk = ok;
            bt1 = ot1;
            bs1 = os1;
            bs2 = os2;
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
    w.s1_x4 = bs1;
    w.s2_x16 = bs2;
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

// Work item (table, range), one wave.
__global__ __launch_bounds__(64) void k_fisher_search(const TableDesc *__restrict__ tabs,
                                                      const Work2 *__restrict__ work, Common cm)
{
    const Work2 wk = work[blockIdx.x];
    const TableG T = view(tabs[wk.x]);
    const uint32_t r = wk.y, lane = threadIdx.x;
    switch (T.size) {
    case 2: fisher_range<2>(T, r, lane, cm); break;
    case 4: fisher_range<4>(T, r, lane, cm); break;
    case 8: fisher_range<8>(T, r, lane, cm); break;
    case 16: fisher_range<16>(T, r, lane, cm); break;
    case 32: fisher_range<32>(T, r, lane, cm); break;
    }
}

void launch_fisher(const TableDesc *t, const Work2 *dom_work, uint32_t n_dom,
                   const Work2 *rng_work, uint32_t n_rng, const Work2 *tab_work, uint32_t n_tab,
                   const Work2 *range_work, uint32_t n_range, const Common &c, hipStream_t st)
{
    if (n_dom) hipLaunchKernelGGL(k_fisher_domains, dim3(n_dom), dim3(256), 0, st, t, dom_work);
    if (n_rng) hipLaunchKernelGGL(k_fisher_ranges, dim3(n_rng), dim3(256), 0, st, t, rng_work);
    if (n_tab) hipLaunchKernelGGL(k_fisher_buckets, dim3(n_tab), dim3(1024), 0, st, t, tab_work);
    if (n_range)
        hipLaunchKernelGGL(k_fisher_search, dim3(n_range), dim3(64), 0, st, t, range_work, c);
}

} // namespace mars
