// Device code of the Mars search backend for MI355X (gfx950, CDNA4).
//
// Data flow for one plane (all device resident):
//   k_contract   u8 plane -> contracted D (i32, the 2x2 box sums of
//                encode.rs:473-501) + the split planes Dl = D % 4 and
//                Dh' = D/4 - 128 (i8) in 16 byte-shifted copies, so that any
//                4/8/16-byte window row segment is one aligned load, and D as
//                binary16 in the same layout when a table has size 4 or 8.
//   k_sat_rows, k_sat_cols  summed-area tables of D and D^2.
//   k_possums    per (size, shift) table: per domain position the window sums
//                S1 = sum D, S2 = sum D^2 (moments.rs:21-44) and
//                Di = s0*S2 - S1^2; the search-order record; at sizes 4 and
//                8 the position's factors of the stage-1 bound.
//   k_ranges     per table: each range block permuted under the 8 isometries
//                (encode.rs:833-842) as i8 r-128 (MFMA operand) and as f32
//                (FP32 baseline operand), plus t0 = sum r, t2 = sum r^2.
//   k_seed       a starting threshold per range from a few exact fits.
//   k_search_mfma<S>  range x position x isometry correlation as a GEMM
//                [8 ranges x 8 isometries] x [32 positions] per wave tile,
//                K = S*S, exact: v_mfma_f32_32x32x16_f16 on the centred range
//                and D at sizes 4 and 8 (with one more K-step giving the
//                stage-1 limits), v_mfma_i32_32x32x{16,32}_i8 on the split
//                planes at 2, 16 and 32; then the certified epilogue
//                described at the kernel.
//   k_search_fp32<S>  the plain baseline: FP32 FMA correlation + binary32
//                fit (encode.rs:295-337) for every candidate.
//   k_merge_exact, k_merge  per range: min over position chunks by (rms,
//                key), the winner's moments and (qalfa, qbeta) recomputed
//                exactly.
//
// Every parity-bearing floating-point expression is compiled with
// -ffp-contract=off (and `#pragma clang fp contract(off)` below) so the
// binary64 fit rounds exactly like the reference's Rust.

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "device.h"
#include "kernels.h"

#pragma clang fp contract(off)

namespace mars {

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));
typedef _Float16 v8h __attribute__((ext_vector_type(8)));
typedef float v16f __attribute__((ext_vector_type(16)));

// ---------------------------------------------------------------- pool

// Work item (plane, block): one thread per contracted sample.  Writes D
// (i32), D as f32, and the i8 planes lo = D % 4, hi = D / 4 - 128 into
// copies c = 0..15 at x = j - c (copy c holds the row shifted left by c
// bytes).  Block 0 of the launch also clears the step's counters.
__device__ __forceinline__ void contract_item(const PlaneDesc *__restrict__ planes, const Work2 wk)
{
    const PlaneG pl = view(planes[wk.x]);
    const uint32_t idx = wk.y * blockDim.x + threadIdx.x;
    if (idx >= pl.wc * pl.hc) return;
    const uint32_t i = idx / pl.wc, j = idx - i * pl.wc;
    const GLB uint8_t *p0 = pl.px + (size_t)(2 * i) * pl.w + 2 * j;
    int32_t v = (int32_t)p0[0] + (int32_t)p0[1] + (int32_t)p0[pl.w] + (int32_t)p0[pl.w + 1];
    pl.d[idx] = v;
    pl.df[idx] = (float)v;
    const int8_t l = (int8_t)(v & 3);           // Dl = D % 4
    const int8_t hh = (int8_t)((v >> 2) - 128); // Dh' = D / 4 - 128
    const uint16_t vf = f16_bits(v);
    GLB uint16_t *wf = pl.wf;
#pragma unroll
    for (uint32_t c = 0; c < 16; c++) {
        if (j >= c) {
            size_t o = c * (size_t)pl.copy_stride + (size_t)i * pl.pitch + (j - c);
            pl.lo[o] = l;
            pl.hi[o] = hh;
            if (wf && c < 8) wf[o] = vf; // sizes 4 and 8 read copies 0..7 only
        }
    }
}

// Summed-area tables of D and D^2: sat[(i)(wc+1) + j] = sum over rows < i,
// columns < j.  Pass 1, work item (plane, row i; i = hc is the zero row):
// inclusive row prefix sums by a block-wide scan.  Pass 2, work item (plane,
// block of 16 columns): running sums down the columns.
__global__ void k_sat_rows(const PlaneDesc *__restrict__ planes, const Work2 *__restrict__ work)
{
    __shared__ int64_t s1[256], s2[256];
    const Work2 wk = work[blockIdx.x];
    const PlaneG pl = view(planes[wk.x]);
    const uint32_t wc = pl.wc, hc = pl.hc;
    GLB int64_t *sat1 = pl.sat1, *sat2 = pl.sat2;
    const uint32_t i = wk.y, t = threadIdx.x, W1 = wc + 1;
    if (i == hc) { // the zero row 0
        for (uint32_t j = t; j < W1; j += 256) {
            sat1[j] = 0;
            sat2[j] = 0;
        }
        return;
    }
    GLB int64_t *o1 = sat1 + (size_t)(i + 1) * W1, *o2 = sat2 + (size_t)(i + 1) * W1;
    if (t == 0) {
        o1[0] = 0;
        o2[0] = 0;
    }
    int64_t c1 = 0, c2 = 0;
    for (uint32_t base = 0; base < wc; base += 256) {
        uint32_t j = base + t;
        int64_t x = j < wc ? pl.d[(size_t)i * wc + j] : 0;
        s1[t] = x;
        s2[t] = x * x;
        __syncthreads();
        for (uint32_t off = 1; off < 256; off <<= 1) {
            int64_t a1 = t >= off ? s1[t - off] : 0, a2 = t >= off ? s2[t - off] : 0;
            __syncthreads();
            s1[t] += a1;
            s2[t] += a2;
            __syncthreads();
        }
        if (j < wc) {
            o1[j + 1] = c1 + s1[t];
            o2[j + 1] = c2 + s2[t];
        }
        c1 += s1[255];
        c2 += s2[255];
        __syncthreads();
    }
}

// Column pass: a block covers 16 columns x 16 row segments; each thread
// scans its segment, the segment totals are combined in LDS, and a second
// sweep adds the preceding segments' total.  Both sweeps go CH rows at a
// time, all loads of a batch before its stores, so a thread has CH loads in
// flight instead of one.
__global__ void k_sat_cols(const PlaneDesc *__restrict__ planes, const Work2 *__restrict__ work)
{
    __shared__ int64_t tot1[16][16], tot2[16][16];
    const Work2 wk = work[blockIdx.x];
    const PlaneG pl = view(planes[wk.x]);
    const uint32_t wc = pl.wc, hc = pl.hc;
    GLB int64_t *sat1 = pl.sat1, *sat2 = pl.sat2;
    const uint32_t cl = threadIdx.x & 15, seg = threadIdx.x >> 4, W1 = wc + 1;
    const uint32_t j = wk.y * 16 + cl;
    const uint32_t L = (hc + 15) / 16, r0 = 1 + seg * L;
    uint32_t r1 = r0 + L;
    if (r1 > hc + 1) r1 = hc + 1;
    constexpr uint32_t CH = 8;
    int64_t a1 = 0, a2 = 0;
    if (j < W1)
        for (uint32_t i0 = r0; i0 < r1; i0 += CH) {
            int64_t x1[CH], x2[CH];
#pragma unroll
            for (uint32_t c = 0; c < CH; c++) {
                const size_t o = (size_t)(i0 + c) * W1 + j;
                x1[c] = i0 + c < r1 ? sat1[o] : 0;
                x2[c] = i0 + c < r1 ? sat2[o] : 0;
            }
#pragma unroll
            for (uint32_t c = 0; c < CH; c++) {
                a1 += x1[c];
                a2 += x2[c];
                if (i0 + c < r1) {
                    const size_t o = (size_t)(i0 + c) * W1 + j;
                    sat1[o] = a1;
                    sat2[o] = a2;
                }
            }
        }
    tot1[seg][cl] = a1;
    tot2[seg][cl] = a2;
    __syncthreads();
    int64_t o1 = 0, o2 = 0;
    for (uint32_t s2 = 0; s2 < seg; s2++) {
        o1 += tot1[s2][cl];
        o2 += tot2[s2][cl];
    }
    if (j < W1 && seg > 0)
        for (uint32_t i0 = r0; i0 < r1; i0 += CH) {
            int64_t x1[CH], x2[CH];
#pragma unroll
            for (uint32_t c = 0; c < CH; c++) {
                const size_t o = (size_t)(i0 + c) * W1 + j;
                x1[c] = i0 + c < r1 ? sat1[o] : 0;
                x2[c] = i0 + c < r1 ? sat2[o] : 0;
            }
#pragma unroll
            for (uint32_t c = 0; c < CH; c++)
                if (i0 + c < r1) {
                    const size_t o = (size_t)(i0 + c) * W1 + j;
                    sat1[o] = x1[c] + o1;
                    sat2[o] = x2[c] + o2;
                }
        }
}

// Work item (table, block of 256 domain positions): window sums from the
// SATs.  Per position p: {S1, S2} and {Di, 1/Di} (binary64); the search's
// hot record {window byte offset, S1, p, sqrt(Di) as f32} goes to slot
// iperm[p] of the search order (identity when iperm is null), and slots
// P..P_pad-1 get a padding record (p = ~0).
__global__ void k_possums(const TableDesc *__restrict__ tabs, const Work2 *__restrict__ work)
{
    const Work2 wk = work[blockIdx.x];
    const TableG T = view(tabs[wk.x]);
    const uint32_t p = wk.y * blockDim.x + threadIdx.x;
    if (p >= T.P_pad) return;
    if (p >= T.P) { // padding slot of the last search tile: no position
        // S1 = -2^21: the MFMA search's stage-1 limit excludes the slot
        // (binary16 sizes: the 2^12 in k 9 of its bias record).
        T.posinfo[p] = T.posbias ? (u32x4){0u, 0xffffffffu, (uint32_t)-(1 << 21), 0u}
                                 : (u32x4){0u, (uint32_t)-(1 << 21), 0xffffffffu, 0u};
        if (T.posbias) {
            T.posbias[2 * p] = (u32x4){0u, 0u, 0u, 0u};
            T.posbias[2 * p + 1] = (u32x4){0x6c000000u, 0u, 0u, 0u}; // k 9 = 2^12
        }
        return;
    }
    const uint32_t W1 = T.wc + 1, size = T.size;
    const uint32_t pi = p / T.npc, pj = p - pi * T.npc;
    const uint32_t dr = (pi * T.shift) >> 1, dc = (pj * T.shift) >> 1;
    const size_t a = (size_t)dr * W1 + dc, b = (size_t)(dr + size) * W1 + dc;
    const int64_t s1 = T.sat1[b + size] - T.sat1[a + size] - T.sat1[b] + T.sat1[a];
    const int64_t s2 = T.sat2[b + size] - T.sat2[a + size] - T.sat2[b] + T.sat2[a];
    const int64_t n = (int64_t)size * size;
    const int64_t di = n * s2 - s1 * s1;
    // z: sqrt(Di) rounded down to binary32 (the stage-1 bound's factor).
    const float sq = __double2float_rd(sqrt((double)di));
    // x: byte offset of the window origin in the copy of the split planes
    // where its rows start aligned to the search's load width.
    const uint32_t cmask = size == 2 ? 0 : size == 4 ? 3 : size == 8 ? 7 : 15;
    const uint32_t cp = dc & cmask;
    const uint32_t boff = cp * T.copy_stride + dr * T.pitch + (dc - cp);
    const uint32_t q = T.iperm ? T.iperm[p] : p; // slot in search order
    // Sizes 2, 16, 32: {offset, S1, p, sqrt(Di)} for the search's limits.
    // Sizes 4 and 8 (limits from posbias): {offset, p, S1, S2}, the search
    // reading the first two words and the Fisher gather all four.
    T.posinfo[q] = T.posbias ? (u32x4){boff, p, (uint32_t)s1, (uint32_t)s2}
                             : (u32x4){boff, (uint32_t)s1, p, __float_as_uint(sq)};
    if (T.posbias) {
        // Binary16 sizes: the position's factors of the stage-1 bound (the
        // B operand of the bias K-step of k_search_mfma, 32 bytes).  Every value is
        // exact or rounded down: S1 = 256 S1h + S1l with S1h <= 63 (s = 4) or
        // 255 (s = 8), so S1h 256/n and S1l/n are binary16 integers times
        // powers of two; dh + dl <= sqrt(Di)/n, the (1 - 2^-20) covering the
        // binary64 square root's rounding before sq was rounded down.
        const uint32_t S1 = (uint32_t)s1, s1h = S1 >> 8, s1l = S1 & 255;
        const float inv_n = 1.0f / (float)n; // a power of two
        const float dn = sq * (1.0f - 0x1p-20f) * inv_n;
        const uint32_t dh = f16_rd_bits(dn);
        const uint32_t dl = f16_rd_bits(dn - (float)__builtin_bit_cast(_Float16, (uint16_t)dh));
        const uint32_t b0 = f16_bits_of((float)s1h * (256.0f * inv_n));
        const uint32_t b1 = f16_bits_of((float)s1l * inv_n);
        // k 0..7 (lane half 0), then k 8..15 (half 1); k 4..7 are the
        // constants {2^11, 1, 1, 1} of the intercept and margin slots
        T.posbias[2 * q] = (u32x4){b0 | (b1 << 16), dh | (dl << 16), 0x3c006800u, 0x3c003c00u};
        T.posbias[2 * q + 1] = (u32x4){dh, 0u, 0u, 0u};
    }
    T.posx[2 * p] = (int32_t)s1;
    T.posx[2 * p + 1] = (int32_t)s2;
    T.posd[2 * p] = (double)di;
    T.posd[2 * p + 1] = di > 0 ? 1.0 / (double)di : 0.0;
}

// Range operands, work item (table, block): one thread per 4 bytes of
// rng8, layout [range][iso][kmem] (i8 r-128, zero padded to kmem); with
// `fp32` also rngf, layout [range/4][k][iso*4 + range%4] (f32 r).  The block
// that starts a range also writes its sums t0 = sum r, t2 = sum r^2 and
// 3 * sum max(r - 128, 0) (zero for a padding range) and resets its pruning
// threshold: its pixels are in LDS here anyway.
__device__ __forceinline__ void ranges_item(const TableDesc *__restrict__ tabs, const Work2 wk,
                                            int fp32)
{
    // The block's 1024 output bytes cover whole ranges (sizes <= 8) or part
    // of one range; its ranges' pixels are staged in LDS first (row-wise,
    // coalesced), so each output byte is an LDS read rather than a gather.
    __shared__ uint8_t s_px[1024];
    const TableG T = view(tabs[wk.x]);
    const uint32_t lg_size = __builtin_ctz(T.size), lg_kmem = __builtin_ctz(T.kmem);
    const uint32_t size = T.size, K = size * size, kmem = T.kmem;
    const uint32_t e0 = wk.y * 1024, total = T.nr_pad * 8 * kmem;
    const uint32_t rb = e0 >> (lg_kmem + 3);
    const uint32_t per = (1024u >> (lg_kmem + 3)) ? (1024u >> (lg_kmem + 3)) : 1u;
    const uint32_t nrb = min(per, T.nr_pad - rb);
    for (uint32_t idx = threadIdx.x; idx < nrb * K; idx += 256) {
        const uint32_t rr = rb + (idx >> (2 * lg_size)), k = idx & (K - 1);
        s_px[idx] = rr < T.nr
                        ? T.px[(T.rows[rr] + (k >> lg_size)) * T.w + T.cols[rr] + (k & (size - 1))]
                        : (uint8_t)128;
    }
    __syncthreads();
    if ((e0 & (8 * kmem - 1)) == 0) { // the block starts its ranges (a wave per range)
        const uint32_t lane = threadIdx.x & 63;
        for (uint32_t i = threadIdx.x >> 6; i < nrb; i += 4) {
            const uint32_t rr = rb + i;
            int32_t s1 = 0, s2 = 0, sp = 0;
            if (rr < T.nr)
                for (uint32_t k = lane; k < K; k += 64) {
                    const int32_t px = s_px[i * K + k];
                    s1 += px;
                    s2 += px * px;
                    sp += px > 128 ? px - 128 : 0;
                }
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) {
                s1 += __shfl_xor(s1, off);
                s2 += __shfl_xor(s2, off);
                sp += __shfl_xor(sp, off);
            }
            if (lane == 0) {
                T.rstat[2 * rr] = s1;
                T.rstat[2 * rr + 1] = s2;
                T.rpos[rr] = 3 * sp; // largest sum of r'*Dl over Dl in [0, 3]
                T.gthr[rr] = 0x7f800000u; // +inf: no threshold yet
            }
        }
    }
    const uint32_t e = e0 + threadIdx.x * 4;
    if (e >= total) return;
    const uint32_t r = e >> (lg_kmem + 3), iso = (e >> lg_kmem) & 7, k0 = e & (kmem - 1);
    const bool valid = r < T.nr;
    const uint8_t *src = s_px + (r - rb) * K;
    GLB float *rngf = fp32 ? T.rngf : nullptr;
    uint32_t packed = 0;
#pragma unroll
    for (uint32_t b = 0; b < 4; b++) {
        const uint32_t k = k0 + b;
        if (valid && k < K) {
            const uint32_t u = k >> lg_size, v = k & (size - 1);
            uint32_t i, j;
            iso_map(iso, u, v, size, &i, &j);
            const uint32_t pix = src[i * size + j];
            packed |= ((pix - 128u) & 0xffu) << (8 * b);
            if (rngf) rngf[((size_t)(r >> 2) * K + k) * 32 + iso * 4 + (r & 3)] = (float)pix;
        } else if (rngf && k < K) {
            rngf[((size_t)(r >> 2) * K + k) * 32 + iso * 4 + (r & 3)] = 0.0f;
        }
    }
    *(uint32_t *)(T.rng8 + e) = packed;
}

__global__ void k_contract(const PlaneDesc *__restrict__ planes, const Work2 *__restrict__ work,
                           unsigned long long *__restrict__ counters, uint32_t n_counters)
{
    if (blockIdx.x == 0 && threadIdx.x < n_counters) glb(counters)[threadIdx.x] = 0;
    contract_item(planes, work[blockIdx.x]);
}

__global__ __launch_bounds__(256) void k_ranges(const TableDesc *__restrict__ tabs,
                                                const Work2 *__restrict__ work, int fp32)
{
    ranges_item(tabs, work[blockIdx.x], fp32);
}

// ---------------------------------------------------------------- MFMA search

template <int S> struct Geo {
    static constexpr int K = S * S;
    static constexpr bool SMALLK = S <= 4;                 // v_mfma_i32_32x32x16_i8
    static constexpr int KSTEP = SMALLK ? 16 : 32;         // K per instruction
    static constexpr int KMEM = K < 16 ? 16 : K;           // padded range row length
    static constexpr int NK = (KMEM + KSTEP - 1) / KSTEP;  // K steps
    static constexpr bool PRELOAD_A = NK <= 2;             // keep A in registers
    // Sizes 4 and 8 correlate in binary16 (v_mfma_f32_32x32x16_f16): r in
    // [0, 255] and D in [0, 1020] are exact, and t1 = sum r*D < 2^24 even at
    // s = 8 (64 * 255 * 1020 = 16 646 400), so the binary32 accumulation of
    // these non-negative integers is exact in any order.  One instruction per
    // K-step gives t1 itself: no hi/lo planes to combine.
    static constexpr bool F16 = S == 4 || S == 8;
    static constexpr int NKH = F16 ? K / 16 : 1;           // binary16 K-steps
};

// Binary16 B fragment of lane half h for K-step kk (16 values of k) of the
// window at element offset `boff` of the binary16 plane (the split planes'
// byte offset: the same copy layout with 2-byte elements).  k = row * S +
// column, lane half h holding k = 16 kk + 8 h .. + 7, as the A fragment.
template <int S>
__device__ __forceinline__ v8h load_bh(const GLB uint16_t *wf, uint32_t boff, uint32_t pitch, int kk,
                                       int h)
{
    const GLB char *base = (const GLB char *)wf;
    if constexpr (S == 4) {
        const uint32_t o = 2u * (boff + 2 * h * pitch); // rows 2h, 2h + 1
        const u32x2 r0 = *(const GLB u32x2 *)(base + o);
        const u32x2 r1 = *(const GLB u32x2 *)(base + (o + 2 * pitch));
        const u32x4 t = {r0.x, r0.y, r1.x, r1.y};
        return __builtin_bit_cast(v8h, t);
    } else {
        const uint32_t o = 2u * (boff + (2 * kk + h) * pitch); // row 2 kk + h
        return __builtin_bit_cast(v8h, *(const GLB u32x4 *)(base + o));
    }
}

// Binary16 A fragment: row (range, isometry) of lane & 31 as load_a, k = 16
// kk + 8 h .. + 7, built from the i8 operand (r - 128) as r - 128 + add
// (the search centres r: add = 128 - round(t0/n), |r - m| <= 255 exact):
// the fragments are loaded once per block, so the conversion costs nothing
// per tile and the range operand is stored once, in i8.
template <int S>
__device__ __forceinline__ v8h load_ah(const GLB int8_t *rng8, uint32_t g, int tt, int kk, int lane,
                                       int32_t add)
{
    constexpr int KMEM = Geo<S>::KMEM;
    const int m = lane & 31, h = lane >> 5;
    const int rl = ((m >> 2) & 1) * 4 + (m & 3);
    const int iso = 4 * tt + (m >> 3);
    const GLB int8_t *p = rng8 + ((size_t)(g * 8 + rl) * 8 + iso) * KMEM + kk * 16 + h * 8;
    const u32x2 w = *(const GLB u32x2 *)p;
    v8h a;
#pragma unroll
    for (int e = 0; e < 8; e++)
        a[e] = (_Float16)(float)((int32_t)(int8_t)(w[e >> 2] >> (8 * (e & 3))) + add);
    return a;
}

// B fragment (one split plane; sizes 2, 16 and 32) of lane half h for
// K-step kk of the window whose origin is byte `boff` of the plane's copy
// layout (at sizes >= 16 the copy in which the window's rows start 16-byte
// aligned; all copies of a plane span < 4 GiB).  The k order inside a fragment is the same one used for the A
// fragment, so the product does not depend on how the instruction orders K
// internally.
template <int S>
__device__ __forceinline__ void load_b(const GLB int8_t *plane, uint32_t boff, uint32_t pitch,
                                       int kk, int h, long *b8, v4i *b16)
{
    if constexpr (S == 2) {
        // K = 4 padded to 16: half 0 holds k 0..3 = (0,0) (0,1) (1,0) (1,1).
        uint32_t x = 0;
        if (h == 0) {
            const GLB uint8_t *p = (const GLB uint8_t *)plane + boff;
            x = (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[pitch] << 16) |
                ((uint32_t)p[pitch + 1] << 24);
        }
        *b8 = (long)x;
    } else {
        static_assert(S == 2, "sizes 4 and 8 search in binary16 (load_bh), sizes >= 16 by load_b16");
    }
}

// B fragment at sizes >= 16 (16 bytes of one split plane) of K-step kk:
// k = 32 kk + 16 h .. + 15, row 2 kk + h at S = 16, row kk and column 16 h
// at S = 32.  lane_off = the window's offset plus the lane's part (h pitch,
// or 16 h); the K-step's part, a multiple of the pitch, is the same for the
// whole wave, and the two add as 32-bit offsets from the plane: one add per
// load (a row index times the pitch had cost a v_mul_lo_u32 and 64-bit
// address arithmetic per load).
template <int S>
__device__ __forceinline__ v4i load_b16(const GLB int8_t *plane, uint32_t lane_off, uint32_t pitch,
                                        int kk)
{
    static_assert(S >= 16, "split-plane K-steps of 32 values");
    const uint32_t step = (S == 16 ? 2u * (uint32_t)kk : (uint32_t)kk) * pitch;
    return *(const GLB v4i *)(plane + (lane_off + step));
}

template <int S>
__device__ __forceinline__ void load_a(const GLB int8_t *rng8, uint32_t g, int tt, int kk, int lane,
                                       long *a8, v4i *a16)
{
    constexpr int KMEM = Geo<S>::KMEM;
    constexpr int KSTEP = Geo<S>::KSTEP;
    int m = lane & 31, h = lane >> 5;
    int rl = ((m >> 2) & 1) * 4 + (m & 3);
    int iso = 4 * tt + (m >> 3);
    const GLB int8_t *p = rng8 + ((size_t)(g * 8 + rl) * 8 + iso) * KMEM + kk * KSTEP + h * (KSTEP / 2);
    if constexpr (Geo<S>::SMALLK)
        *a8 = *(const GLB long *)p;
    else
        *a16 = *(const GLB v4i *)p;
}

struct BestRec {
    double rms;
    uint32_t key;
    int32_t t1; // exact t1_x4 of the record
};

// One entry of a ranked (top-k) list: the candidate, its exact t1 and its
// inflated threshold (tau_bits of its SSE).
struct TopRec {
    double rms;
    uint32_t key;
    int32_t t1;
    uint32_t tau;
    uint32_t pad;
};

// Inflated threshold of a candidate whose binary64 SSE (`sum` of
// encode.rs:283-285) is `sum`: anything whose exact SSE exceeds it has a
// strictly larger binary64 rms (relative 1e-9 plus absolute 1e-6*n cover the
// rounding of `sum`, of /s0 and of sqrt).  Rounded up to binary32.
__device__ __forceinline__ uint32_t tau_bits(double sum, int64_t n)
{
    double tau = (sum > 0.0 ? sum : 0.0) * (1.0 + 1e-9) + 1e-6 * (double)n;
    return __float_as_uint((float)(tau * (1.0 + 1e-7)));
}

// Walk-only tables, work item (table, block of 256 ranges): whether the
// threshold walk (encode.rs walk, the RD warm-up's) visits range r, from its
// parent's merged winner -- the parent splits when its best rms exceeds
// t_rms (no candidate: +inf) and its size exceeds min_size; a parent the walk
// splits without searching (larger than the largest range size, or across
// the border) always passes its children on.
__global__ __launch_bounds__(256) void k_activate(const TableDesc *__restrict__ tabs,
                                                  const Work2 *__restrict__ work, double t_rms,
                                                  uint32_t min_size, Work3 *__restrict__ wq,
                                                  uint32_t *__restrict__ wq_count)
{
    __shared__ uint8_t s_a[256];
    const Work2 wk = work[blockIdx.x];
    const TableG T = view(tabs[wk.x]);
    const uint32_t r = wk.y * 256 + threadIdx.x;
    uint8_t a = 0;
    if (r < T.nr) {
        const int32_t p = T.parent[r];
        if (p == -1) {
            a = 1;
        } else if (p >= 0) {
            const TableG U = view(tabs[T.parent_tab]);
            const GLB double *rp = (const GLB double *)((const GLB char *)(U.out + p) +
                                                        offsetof(mars_hip_candidate, rms));
            const double rms = U.found[p] ? *rp : __longlong_as_double(0x7ff0000000000000ll);
            a = U.active[p] && rms > t_rms && U.size > min_size;
        }
    }
    if (r < T.nr_pad) T.active[r] = a;
    s_a[threadIdx.x] = a;
    __syncthreads();
    // The block's 32 range groups of 8: each with a visited range gets all
    // its chunks appended to the search's work list.
    const uint32_t g = wk.y * 32 + threadIdx.x;
    if (threadIdx.x < 32 && g * 8 < T.nr_pad && T.P > 0) {
        bool any = false;
#pragma unroll
        for (int i = 0; i < 8; i++) any |= s_a[threadIdx.x * 8 + i] != 0;
        if (any) {
            const uint32_t nch = T.nchunks[0];
            const uint32_t base = atomicAdd(wq_count, nch);
            for (uint32_t ch = 0; ch < nch; ch++) wq[base + ch] = make_uint3(wk.x, ch, g);
        }
    }
}

__global__ void k_zero_u32(uint32_t *p, uint32_t n)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

// Correlation t1_x4 of one candidate from its range row rk (i8 r - 128,
// 16-byte aligned) and its window dw of the contracted plane (row stride wc):
// sizes <= 8 issue every load at once, larger sizes one window row at a time.
template <int S>
__device__ __forceinline__ int32_t seed_t1(const GLB int8_t *rk, const GLB int32_t *dw, uint32_t wc)
{
    constexpr int RS = S <= 8 ? S * S : S; // range bytes per load batch
    constexpr int UN = S <= 8 ? 1 : 2;
    int32_t t1 = 0;
#pragma unroll UN
    for (int u0 = 0; u0 < (S <= 8 ? 1 : S); u0++) {
        int32_t rv[RS], dv[RS];
        const GLB int8_t *rr = rk + u0 * RS;
        if constexpr (S == 2) {
            const uint32_t w = *(const GLB uint32_t *)rr;
#pragma unroll
            for (int k = 0; k < 4; k++) rv[k] = (int32_t)(int8_t)(w >> (8 * k));
        } else {
#pragma unroll
            for (int k = 0; k < RS; k += 16) {
                const v4i w = *(const GLB v4i *)(rr + k);
#pragma unroll
                for (int e = 0; e < 16; e++) rv[k + e] = (int32_t)(int8_t)((uint32_t)w[e >> 2] >> (8 * (e & 3)));
            }
        }
#pragma unroll
        for (int k = 0; k < RS; k++) {
            const int u = S <= 8 ? k / S : u0, v = S <= 8 ? k % S : k;
            dv[k] = dw[(size_t)u * wc + v];
        }
#pragma unroll
        for (int k = 0; k < RS; k++) t1 += (rv[k] + 128) * dv[k];
    }
    return t1;
}

// Threshold seed, work item (table, first of four ranges): wave w seeds
// range 4 y + w from 64 spread candidates (8 positions x 8 isometries), one
// per lane -- the lane's correlation with its loads in flight together,
// then its binary64 fit.  The smallest inflated SSE (ranked tables: the
// topk-th smallest over distinct candidates) becomes the range's starting
// threshold.  No block-wide barrier: a wave per range.
template <int S>
__device__ __forceinline__ void seed_wave(const TableG &T, const uint32_t r, const Common &cm,
                                          uint32_t *s_tau)
{
    constexpr int64_t n = (int64_t)S * S;
    const uint32_t lane = threadIdx.x & 63, P = T.P;
    const uint32_t j = lane >> 3, iso = lane & 7;
    const uint32_t p = P >= 8 ? (uint32_t)(((uint64_t)j * (P - 1)) / 7) : j % P;
    const uint32_t pi = p / T.npc, pj = p - pi * T.npc;
    const uint32_t dr = (pi * T.shift) >> 1, dc = (pj * T.shift) >> 1;
    const GLB int8_t *rk = T.rng8 + ((size_t)r * 8 + iso) * T.kmem; // 16-byte aligned rows
    int32_t t1;
    if constexpr (S >= 16) {
        // From the split planes, whose window rows start 16-byte aligned at
        // these sizes (the search order record's offset): 16 values per
        // load and per four dot4 per plane, where one i32 load per value of
        // the contracted plane made the s = 32 waves the seed's slowest.
        // t1 = 4 sum r' Dh' + sum r' Dl + 512 sum r' + 128 S1, exactly.
        const uint32_t boff = T.posinfo[T.iperm ? T.iperm[p] : p].x;
        int32_t ah = 0, al = 0;
        for (int u = 0; u < S; u++) {
#pragma unroll
            for (int c = 0; c < S; c += 16) {
                const v4i x = *(const GLB v4i *)(rk + u * S + c);
                const uint32_t o = boff + u * T.pitch + c;
                const v4i hv = *(const GLB v4i *)(T.hi + o);
                const v4i lv = *(const GLB v4i *)(T.lo + o);
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    ah = __builtin_amdgcn_sdot4(x[q], hv[q], ah, false);
                    al = __builtin_amdgcn_sdot4(x[q], lv[q], al, false);
                }
            }
        }
        const int32_t t0 = T.rstat[2 * r];
        t1 = 4 * ah + al + 512 * (t0 - (int32_t)(128 * n)) + 128 * T.posx[2 * p];
    } else {
        t1 = seed_t1<S>(rk, T.d + (size_t)dr * T.wc + dc, T.wc);
    }
    uint32_t qa, qb;
    double rms, sum;
    fit_f64(n, T.posx[2 * p], T.posx[2 * p + 1], T.rstat[2 * r], t1, T.rstat[2 * r + 1],
            cm.max_alfa, cm.bits_alfa, cm.bits_beta, &qa, &qb, &rms, &sum);
    const uint32_t tb = tau_bits(sum, n);
    if (cm.topk <= 1) {
        uint32_t m = tb;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) m = min(m, (uint32_t)__shfl_xor((int)m, off));
        if (lane == 0) T.gthr[r] = m;
        return;
    }
    // Ranked: only distinct samples count (positions repeat when P < 8).
    s_tau[lane] = (P >= 8 || j < P) ? tb : 0x7f800000u;
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    if (lane == 0) {
        // The topk-th smallest sampled threshold (+inf with fewer distinct
        // samples): topk real candidates reach it.
        uint32_t th = 0x7f800000u, prev = 0, taken = 0;
        for (uint32_t i = 0; i < cm.topk && i < 64; i++) {
            uint32_t m = 0x7f800000u, cnt = 0;
            for (uint32_t e = 0; e < 64; e++) {
                const uint32_t x = s_tau[e];
                if ((i == 0 || x > prev) && x < m) {
                    m = x;
                    cnt = 0;
                }
                if (x == m) cnt++;
            }
            if (m == 0x7f800000u) break;
            taken += cnt;
            prev = m;
            if (taken >= cm.topk) {
                th = m;
                break;
            }
        }
        T.gthr[r] = th;
    }
}

// Threshold seed of 8 ranges per wave (winner-only tables, sizes <= 8):
// lanes 8i..8i+7 take range r0 + i at the 8 spread positions; each lane
// correlates its window with the range's 8 isometries, fits the one with
// the largest Ni = n t1 - S1 t0 (the free least-squares favourite), and the
// range's threshold is the smallest inflated SSE of its 8 fits.  Any
// candidate's SSE bounds the best one, so the threshold stays valid; the
// fits, not the correlations, bound the seed, and a wave now does 8 ranges'.
template <int S>
__device__ __forceinline__ void seed8_wave(const TableG &T, const uint32_t r0, const Common &cm)
{
    constexpr int64_t n = (int64_t)S * S;
    constexpr int K = S * S;
    const uint32_t lane = threadIdx.x & 63, P = T.P;
    const uint32_t r = r0 + (lane >> 3), j = lane & 7;
    const bool live = r < T.nr && !(T.active && !T.active[r]);
    uint32_t tb = 0x7f800000u;
    if (live) {
        const uint32_t p = P >= 8 ? (uint32_t)(((uint64_t)j * (P - 1)) / 7) : j % P;
        const uint32_t pi = p / T.npc, pj = p - pi * T.npc;
        const uint32_t dr = (pi * T.shift) >> 1, dc = (pj * T.shift) >> 1;
        const GLB int32_t *dw = T.d + (size_t)dr * T.wc + dc;
        int32_t dv[K];
#pragma unroll
        for (int k = 0; k < K; k++) dv[k] = dw[(size_t)(k / S) * T.wc + k % S];
        const int32_t S1 = T.posx[2 * p];
        const int64_t t0 = T.rstat[2 * r];
        int32_t best = 0;
        int64_t best_ni = INT64_MIN;
        for (uint32_t iso = 0; iso < 8; iso++) {
            const GLB int8_t *rk = T.rng8 + ((size_t)r * 8 + iso) * T.kmem; // 16-byte aligned
            int32_t t1 = 0;
            if constexpr (S == 2) {
                const uint32_t w = *(const GLB uint32_t *)rk;
#pragma unroll
                for (int k = 0; k < 4; k++) t1 += ((int32_t)(int8_t)(w >> (8 * k)) + 128) * dv[k];
            } else {
#pragma unroll
                for (int k = 0; k < K; k += 16) {
                    const v4i w = *(const GLB v4i *)(rk + k);
#pragma unroll
                    for (int e = 0; e < 16; e++)
                        t1 += ((int32_t)(int8_t)((uint32_t)w[e >> 2] >> (8 * (e & 3))) + 128) * dv[k + e];
                }
            }
            const int64_t ni = n * (int64_t)t1 - (int64_t)S1 * t0;
            if (ni > best_ni) {
                best_ni = ni;
                best = t1;
            }
        }
        uint32_t qa, qb;
        double rms, sum;
        fit_f64(n, S1, T.posx[2 * p + 1], t0, best, T.rstat[2 * r + 1], cm.max_alfa, cm.bits_alfa,
                cm.bits_beta, &qa, &qb, &rms, &sum);
        tb = tau_bits(sum, n);
    }
#pragma unroll
    for (int off = 1; off < 8; off <<= 1) tb = min(tb, (uint32_t)__shfl_xor((int)tb, off));
    if (live && j == 0) T.gthr[r] = tb;
}

__global__ __launch_bounds__(256) void k_seed(const TableDesc *__restrict__ tabs,
                                              const Work2 *__restrict__ work, Common cm)
{
    __shared__ uint32_t s_tau[4][64];
    const Work2 wk = work[blockIdx.x];
    const TableG T = view(tabs[wk.x]);
    if (seed_ranges_per_block(T.size, cm.topk) == 32) {
        const uint32_t r0 = wk.y * 32 + (threadIdx.x >> 6) * 8;
        if (r0 >= T.nr) return; // the whole wave
        switch (T.size) {
        case 2: seed8_wave<2>(T, r0, cm); break;
        case 4: seed8_wave<4>(T, r0, cm); break;
        default: seed8_wave<8>(T, r0, cm); break;
        }
        return;
    }
    const uint32_t wave = threadIdx.x >> 6, r = wk.y * 4 + wave;
    if (r >= T.nr) return; // the whole wave: no barrier below is block-wide
    if (T.active && !T.active[r]) return; // walk-only: not visited this step
    switch (T.size) {
    case 2: seed_wave<2>(T, r, cm, s_tau[wave]); break;
    case 4: seed_wave<4>(T, r, cm, s_tau[wave]); break;
    case 8: seed_wave<8>(T, r, cm, s_tau[wave]); break;
    case 16: seed_wave<16>(T, r, cm, s_tau[wave]); break;
    default: seed_wave<32>(T, r, cm, s_tau[wave]); break;
    }
}

// Stage 2 (see k_search_mfma): true when the candidate may beat or tie the
// threshold, or sits on a quantisation boundary -- i.e. needs stage 3.
__device__ __forceinline__ bool stage2_keep(const Common &a, double nd, double inv_n,
                                            int32_t t1, int32_t S1, int64_t t0, double Di,
                                            double invDi, double A, float thr)
{
    const double t0d = (double)t0;
    const double Ni = fma(nd, (double)t1, -((double)S1 * t0d));
    double qa = 0.0;
    if (Ni > 0.0 && Di > 0.0) {
        // m is within one of floor(x + 0.5); settle it with exact compares.
        double m = floor(Ni * a.alfa_c * invDi + 0.5);
        if (m <= a.qa_max + 1.0) {
            const double tc = 2.0 * a.alfa_c * Ni;
            if (fma(-(2.0 * m + 1.0), Di, tc) >= 0.0) m += 1.0;
            else if (m >= 1.0 && fma(-(2.0 * m - 1.0), Di, tc) < 0.0) m -= 1.0;
        }
        qa = m > a.qa_max ? a.qa_max : m;
    }
    const double a2 = qa * a.alfa_unit;
    const double bs = (t0d - a2 * ((double)S1 * 0.25)) * inv_n;
    double beta = bs;
    if (a2 > 0.0) beta += a2 * 255.0;
    const double yp = beta * glb(a.yscale)[(int)qa] + 0.5;
    double mb = floor(yp);
    const double fr = yp - mb;
    const bool boundary = fr < 1e-9 || fr > 1.0 - 1e-9;
    mb = mb < 0.0 ? 0.0 : (mb > a.qb_max ? a.qb_max : mb);
    double b2 = mb / a.qb_max * ((1.0 + a2) * 255.0);
    if (a2 > 0.0) b2 -= a2 * 255.0;
    const double f = fma(a2 * a2 * 0.0625, Di, fma(-0.5 * a2, Ni, A));
    const double dd = bs - b2;
    const double sse = f * inv_n + nd * dd * dd;
    return boundary || sse <= (double)thr;
}

__device__ __forceinline__ int32_t vmax(int32_t a, int32_t b) { return max(a, b); }
// IEEE maximum (v_maximum3_f32): no NaN quieting of the MFMA results.
__device__ __forceinline__ float vmax(float a, float b) { return __builtin_elementwise_maximum(a, b); }

// Survivor queue: the wave drains full batches of 64 after every DRAIN
// values (one value per lane each), so DRAIN*64 + 63 entries always fit.
template <int S> struct Queue {
    static constexpr int DRAIN = S <= 8 ? 16 : 4;
    static constexpr int CAP = 64 + 64 * DRAIN;
};

// Certified exhaustive search for one range size S.
//
// Work item (table, position chunk, group of 8 ranges); block = 4 waves,
// each wave takes the chunk's 32-position tiles round robin.  A wave tile is
// 64 GEMM rows (8 ranges x 8 isometries, two 32-row MFMA tiles) x 32
// positions.  Output layout of the 32x32 MFMAs: item t of lane l is
// row 8*(t/4) + 4*(l/32) + t%4, column l%32; rows are arranged so that lane
// half h owns ranges 4h..4h+3 and item t of row tile tt is range t%4,
// isometry 4*tt + t/4.
//
// Correlation.  Sizes 4 and 8: binary16 operands r - m (m = round(t0/n))
// and D, binary32 accumulation of integers below 2^24: X0 = t1 - m*S1
// exactly (Geo<S>::F16).  Sizes 2, 16, 32: r' = r - 128 (i8),
// D = 4*(Dh' + 128) + Dl with Dh' = D/4 - 128 and Dl = D%4 (both i8):
//   t1_x4 = sum r*D = 4*hi + lo + 512*sum r' + 128*S1,
//   hi = sum r'Dh',  lo = sum r'Dl  (exact).
//
// Stage 1 (every candidate).  The least-squares residual with free slope
// and offset is a lower bound of the quantised SSE:
//   n*SSE >= A - Ni^2/Di,  A = n*t2 - t0^2, Di = n*S2 - S1^2,
//   Ni = n*t1 - S1*t0 (moments in the x4/x16 units of encode.rs:224-232).
// Given a threshold tau (an SSE some candidate of this range already
// reached, inflated so that anything above it has a strictly larger binary64
// rms), a candidate can only win if Ni >= sqrt((A - n tau) Di), i.e.
//   t1_x4 > floor((S1*t0 + sqrt((A - n tau) Di)) / n) - 1.
// Sizes <= 4 also use the cap on the slope: alfa2 <= qa_max * alfa_unit,
// so with u = alfa2/4 <= U, n*SSE >= A - 2u*Ni + u^2*Di and, where
// A - n tau > U^2 Di, the bound rises to the tangent (A - n tau + U^2 Di)/(2U)
// (about a fifth of the plain bound's survivors at s = 4 remain).
// Sizes 2, 16, 32: the limit is computed once per (range, position) from
// the lane's cached per-range factors and the position's sqrt(Di), rounded
// so that it never excludes more than the exact bound; then one integer
// compare per candidate.  Sizes 4 and 8: one more binary16 K-step per tile
// gives per (range, position) the plain bound and three tangents of the
// capped one (the construction is described where its A operand is built);
// one compare per (range, position) on the maximum over isometries.
// Candidates with Ni < 0 are
// covered too (their qalfa is 0, SSE >= A/n).  Survivors are appended to a
// per-wave LDS queue (wave ballot + mbcnt).
//
// Stage 2 (queue batches of 64, binary64 with exact operands): the
// candidate's quantised codes and SSE without any IEEE divide or square root.
// qalfa is decided exactly from integers (2c*Ni >= (2k-1)*Di, c the alfa
// scale); beta, alfa2 and the slope part of the SSE
//   n*SSE(alfa2, b*) = A - alfa2*Ni/2 + alfa2^2*Di/16
// are exact dyadic numbers; qbeta is the rounding of beta*(2^bb-1)/((1+a)255)
// unless that product lies within 1e-9 of a rounding boundary; then
//   SSE = n*SSE(alfa2, b*)/n + n*(b* - beta2)^2
// is within 1e-12 of the exact SSE.  A candidate whose SSE exceeds tau is
// dropped; the rest (new bests, near ties, boundary cases) take stage 3, the
// verbatim binary64 fit of encode.rs:247-289, and compete with the wave's
// best by (rms, key).  Winners tighten tau through LDS; blocks exchange tau
// through a global array every 16 tiles.  Tau starts from k_seed's sample.
// KCAP: 0 for the winner-only search, else the capacity of the ranked lists
// (>= c.topk; a smaller list keeps more blocks resident).
template <int S, int KCAP>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(S <= 4 ? 4 : S <= 8 ? 3 : 2))) void k_search_mfma(const TableDesc *__restrict__ tabs,
                                                     const Work3 *__restrict__ work, Common c)
{
    constexpr bool RANKED = KCAP > 0;
    using G = Geo<S>;
    constexpr int NK = G::NK;
    constexpr bool F16 = G::F16;
    constexpr int NKH = G::NKH;
    constexpr int64_t n = (int64_t)S * S;
    constexpr int LOG2S = S == 2 ? 1 : S == 4 ? 2 : S == 8 ? 3 : S == 16 ? 4 : 5;
    constexpr int LOG2N = 2 * LOG2S;
    constexpr bool INT_LIMIT = S <= 8; // every stage-1 quantity fits 32 bits
    // Sizes <= 4 also bound stage 1 by the contrast cap: most of their
    // Cauchy-Schwarz survivors sit on low-variance domains that only an
    // alfa above the largest quantised one could fit.
    constexpr bool CAP_LIMIT = S <= 4;
    // (wave through readfirstlane: the tile loop and its bounds tests are
    // then scalar)
    const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int h = lane >> 5;
    if (c.nwork && blockIdx.x >= *glb(c.nwork)) return; // past a compacted list's end
    const Work3 wk = work[blockIdx.x];
    const TableG T = view(tabs[wk.x]);
    const uint32_t chunk = wk.y, g = wk.z;
    const uint32_t nchunks = T.nchunks[0], tiles_per_chunk = T.tiles_per_chunk[0];
    const double nd = (double)n, inv_n = 1.0 / (double)n;
    if (T.active) { // walk-only table: a group none of whose ranges is visited is skipped
        bool any = false;
#pragma unroll
        for (int i = 0; i < 8; i++) any |= T.active[g * 8 + i] != 0; // nr_pad is a multiple of 8
        if (!any) return;
    }

    __shared__ __align__(16) uint32_t s_thr[8];
    // Bumped whenever one of s_thr drops, so a wave checks one word per tile
    // to know whether its stage-1 limits are stale.
    __shared__ uint32_t s_epoch;
    __shared__ unsigned long long s_count[2];
    __shared__ int64_t s_T0[8], s_T2[8];
    __shared__ double s_A[8];
    __shared__ BestRec s_wb[4][8];
    __shared__ uint2 s_q[4][Queue<S>::CAP];
    // RANKED: each wave's best c.topk candidates per range, ascending by
    // (rms, key), and how many are filled.
    __shared__ TopRec s_top[RANKED ? 4 : 1][RANKED ? 8 : 1][RANKED ? KCAP : 1];
    __shared__ uint32_t s_topn[4][8];
    // Winner-only: per wave and range, the least high word of a drain's
    // winning binary64 rms (a non-negative double orders as its bits).
    __shared__ uint32_t s_hmin[RANKED ? 1 : 4][8];
    // Range operand of the block (sizes >= 16): 64 rows of KMEM bytes,
    // padded rows so the lanes of a fragment read spread over the banks.
    constexpr int ASTRIDE = G::KMEM + 16;
    __shared__ __align__(16) int8_t s_rng[G::PRELOAD_A ? 16 : 64 * ASTRIDE];

    if (threadIdx.x < 2) s_count[threadIdx.x] = 0;
    if (threadIdx.x < 8) {
        uint32_t r = g * 8 + threadIdx.x;
        bool v = r < T.nr;
        s_thr[threadIdx.x] = v ? T.gthr[r] : 0x7f800000u;
        if (threadIdx.x == 0) s_epoch = 0;
        int64_t t0 = v ? T.rstat[2 * r] : 0, t2 = v ? T.rstat[2 * r + 1] : 0;
        s_T0[threadIdx.x] = t0;
        s_T2[threadIdx.x] = t2;
        s_A[threadIdx.x] = (double)(n * t2 - t0 * t0);
    }
    if (lane < 8) {
        s_wb[wave][lane].rms = __longlong_as_double(0x7ff0000000000000ll);
        s_wb[wave][lane].key = 0xffffffffu;
        s_wb[wave][lane].t1 = 0;
        s_topn[wave][lane] = 0;
    }

    // Range statistics of the lane's 4 ranges.
    int32_t T0[4];  // t0 (< 2^24)
    double A[4];    // A (sizes >= 16)
    float Alo[4];   // A rounded down to binary32 (sizes <= 8)
    int32_t rmin[4]; // INT_MIN for a real range, INT_MAX for padding
    // Sizes <= 8 fold the no-pruning and padding cases into the offset the
    // stage-1 limit subtracts.  There |v| < 2^23, |off| < 2^23, S1*128 <
    // 2^24, and X >> LOG2N < 2^25 for a real position (< 2^30 for any S1):
    // +2^27 while a range cannot prune puts its limit below every v; -2^28
    // for a padding range and, through S1 = -2^21 in the padding position
    // record, -2^28 for a padding position put the limit above every v; no
    // combination leaves int32.  Binary16 sizes compare t1 itself (0 <= t1
    // < 2^24, no offset): the same +2^27 / -2^28 range terms, and +2^28 added
    // for a padding position.
    // noff = -1 - (that offset), so a limit is one three-term add.
    int32_t noff[4];
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const uint32_t r = g * 8 + 4 * h + i;
        const bool rv = r < T.nr;
        const uint32_t rr = rv ? r : 0;
        const int64_t t0 = T.rstat[2 * rr], t2 = T.rstat[2 * rr + 1];
        T0[i] = (int32_t)t0;
        A[i] = (double)(n * t2 - t0 * t0);
        Alo[i] = (float)A[i] * (1.0f - 3e-7f); // <= A after both roundings
        rmin[i] = rv ? (int32_t)0x80000000 : 0x7fffffff;
        // no pruning until the first threshold
        noff[i] = -1 - ((F16 ? 0 : (int32_t)(512 * (t0 - 128 * n))) + (rv ? (1 << 27) : -(1 << 28)));
    }
    // t1 = 4*hi + lo + off + 128*S1 with off = 512 * (t0 - 128 n) (< 2^27 in
    // magnitude), recomputed from t0 where needed rather than kept live.
    auto off_of = [&](int i) { return (T0[i] - (int32_t)(128 * n)) * 512; };
    // Stage-1 cache per range: sqrt(A - n tau), and the threshold bits it
    // was computed from.  A - n tau <= 0 (no pruning) is sa < 0 for sizes
    // >= 16; sizes <= 8 fold it into noff and keep sa = 1 (positive, so the
    // limit needs no absolute value).
    float sa[4];
#pragma unroll
    for (int i = 0; i < 4; i++) sa[i] = INT_LIMIT ? 1.0f : -1.0f;
    // s_epoch the limits were computed at; ~0 (none yet, or this wave just
    // lowered a threshold itself) never equals the epoch, a count of drops.
    uint32_t seen = 0xffffffffu;
    uint32_t ep_pre = 0;
    // Binary16 sizes: the stage-1 limit rides on the matrix pipe.  The
    // range operand is centred, r - m with m = round(t0/n), so the K-steps
    // give the exact integer X0 = t1 - m S1 (|X0| <= 1020 |r - m|_1 < 2^24).
    // With a = A - n tau, x = sqrt(a), d = sqrt(Di), a candidate can win
    // only if n t1 - S1 t0 >= L(d) = x d + max(x - U d, 0)^2 / (2U) (the
    // slope cap U bounds alfa/4), i.e. X0 >= S1 e/n + L(d)/n, e = t0 - n m.
    // L is convex in d, so every tangent T_t(d) = (x - t) d + t (2x - t)/(2U)
    // (t >= 0) lies below it.  One more K-step per tile gives, per (range,
    // position) and per isometry row j of the range, a right-hand side
    //   nb_j = S1 e/n + ((x - t_j) d + t_j (2x - t_j)/(2U))/n - M
    // with t = 0 (the plain bound, L's line beyond d = x/U) and the tangents
    // at d = 0, x/(3U), 2x/(3U) (t = x, 2x/3, x/3); a candidate passes when
    // X0 exceeds the largest.  As a rank-10 product (A per row, B per
    // position; k 0..7 in lane half 0, k 8, 9 in half 1):
    //   A  {e, e, sh, sh, k1, k2, k3, -M | sl, 2^15, 0..}
    //   B  {S1h 256/n, S1l/n, dh, dl, 2^11, 1, 1, 1 | dh, 0 or 2^12, 0..}
    // S1 = 256 S1h + S1l and e exact binary16 integers (|e| <= n/2), sh + sl
    // <= x - t and dh + dl <= d/n, 2^11 k1 + k2 + k3 <= t (2x - t)/(2U n),
    // all rounded down, sl dl dropped: the exact nb_j never exceeds the
    // bound less M.  M = 2^-19 E with E the terms' largest possible sum of
    // magnitudes (tau = 0: 1020 (|e| + sqrt(A)) + A/(2U n) at most; 1020 >=
    // D and d/n) exceeds the binary32 rounding of the K-step (at most 9
    // additions of an ulp of E each, whatever the instruction's order).
    // While A - n tau <= 0 the 2^11 slot is -2^15 (nb < 2^15 - 2^26 < -2^24
    // <= X0: everything passes), and 2^15 for a padding range (nb > 2^26 -
    // 2^15 > X0); a padding position's 2^12 in k 9 adds 2^27 (nb > 2^26).
    const int brl = ((lane & 31) >> 2 & 1) * 4 + (lane & 3); // the A row's range
    const int bj = (lane & 31) >> 3;                          // its isometry row: the bound
    uint32_t bconst = 0;   // {e, e} (binary16) of a real range
    float balo = 0.0f;     // A of the row's range rounded down
    float be0 = 0.0f;      // 1020 |e|
    bool bvalid = false;   // the row's range is real
    int32_t acentre = 128; // r - 128 + acentre = r - m
    v8h abias = (v8h){};
    if constexpr (F16) {
        const uint32_t r = g * 8 + brl;
        bvalid = r < T.nr;
        const int64_t t0 = bvalid ? T.rstat[2 * r] : 0, t2 = bvalid ? T.rstat[2 * r + 1] : 0;
        const int64_t m = (t0 + n / 2) >> LOG2N, e = t0 - n * m;
        acentre = 128 - (int32_t)m;
        balo = (float)(double)(n * t2 - t0 * t0) * (1.0f - 3e-7f);
        be0 = 1020.0f * (float)(e < 0 ? -e : e);
        const uint32_t eh = f16_bits_of((float)e);
        bconst = eh | (eh << 16);
    }
    // Per lane: the row's tangent t = tf x gives slope (1 - tf) x = bsl x
    // and intercept tf (2 - tf) x^2 cap_k / n = (A - n tau) bkn (cap_k <=
    // 1/(2U)), both rounded down: 2^-19 covers the roundings of the
    // products, of A - n tau and of the square root.
    const float btf = bj == 0 ? 0.0f : bj == 1 ? 1.0f : bj == 2 ? 2.0f / 3.0f : 1.0f / 3.0f;
    const float bsl = (1.0f - btf) * (1.0f - 0x1p-19f);
    const float bkn = btf * (2.0f - btf) * c.cap_k * (1.0f / (float)n) * (1.0f - 0x1p-19f);
    // -M in the high half: M = 2^-19 (1020 (|e| + sqrt(A) bsl) + A bkn),
    // the terms' sum at tau = 0 (the largest), kept above that value through
    // the rounding toward zero by the 2^-8; at least 2^-14 (binary16 normal)
    uint32_t bmh = 0;
    if constexpr (F16) {
        const float sa = sqrtf(fmaxf(balo, 0.0f)) * (1.0f + 0x1p-10f);
        const float mm = fmaxf((be0 + 1020.0f * sa * bsl + fmaxf(balo, 0.0f) * (1.0f + 0x1p-10f) * bkn) * 0x1p-19f * (1.0f + 0x1p-8f), 0x1p-14f);
        bmh = ((uint32_t)__builtin_bit_cast(uint16_t, __builtin_amdgcn_cvt_pkrtz(mm, 0.0f)[0]) | 0x8000u) << 16;
    }
    // The A operand of the bias K-step for threshold bits `tb` of the row's
    // range; v_cvt_pkrtz_f16_f32 rounds two values toward zero, i.e. down
    // for these non-negative ones.  Rebuilt whenever a threshold of the
    // block drops, so it is kept short.
    auto make_abias = [&](uint32_t tb) -> v8h {
        const float amt = balo - (float)n * __uint_as_float(tb);
        const float x = __builtin_amdgcn_sqrtf(fmaxf(amt, 0.0f));
        const float sg = x * bsl;
        const float kn = fmaxf(amt, 0.0f) * bkn;
        const auto p1 = __builtin_bit_cast(uint32_t, __builtin_amdgcn_cvt_pkrtz(sg, kn * 0x1p-11f));
        const float sh = (float)__builtin_bit_cast(_Float16, (uint16_t)p1);
        const float k1 = (float)__builtin_bit_cast(_Float16, (uint16_t)(p1 >> 16)) * 2048.0f;
        const float r1 = kn - k1;
        const auto p2 = __builtin_bit_cast(uint32_t, __builtin_amdgcn_cvt_pkrtz(sg - sh, r1));
        const float r2 = r1 - (float)__builtin_bit_cast(_Float16, (uint16_t)(p2 >> 16));
        const auto p3 = __builtin_bit_cast(uint32_t, __builtin_amdgcn_cvt_pkrtz(r2, 0.0f)) | bmh;
        const uint32_t shb = p1 & 0xffffu;
        uint32_t w0 = bconst, w1 = shb | (shb << 16);
        uint32_t w2 = (p1 >> 16) | (p2 & 0xffff0000u);  // {k1, k2}
        uint32_t w3 = p3;                               // {k3, -M}
        uint32_t w4 = p2 & 0xffffu;
        if (!(amt > 0.0f) || !bvalid) {
            // no pruning yet: -2^15 in the 2^11 slot; a padding range: 2^15
            w0 = bvalid ? bconst : 0u;
            w1 = w3 = w4 = 0u;
            w2 = bvalid ? 0xf800u : 0x7800u;
        }
        // half 1: k 8 = sl, k 9 = 2^15 (times 2^12 for a padding position)
        const u32x4 w = h == 0 ? (u32x4){w0, w1, w2, w3} : (u32x4){w4 | 0x78000000u, 0u, 0u, 0u};
        return __builtin_bit_cast(v8h, w);
    };
    // Lower a threshold, noting the drop for every wave.
    auto lower = [&](uint32_t slot, uint32_t v) {
        if (atomicMin(&s_thr[slot], v) > v) atomicAdd(&s_epoch, 1u);
    };

    long a8[2][NK];
    v4i a16[2][NK];
    v8h ahf[2][NKH];
    if constexpr (F16) {
#pragma unroll
        for (int tt = 0; tt < 2; tt++)
#pragma unroll
            for (int kk = 0; kk < NKH; kk++) ahf[tt][kk] = load_ah<S>(T.rng8, g, tt, kk, lane, acentre);
    } else if constexpr (G::PRELOAD_A) {
#pragma unroll
        for (int tt = 0; tt < 2; tt++)
#pragma unroll
            for (int kk = 0; kk < NK; kk++) load_a<S>(T.rng8, g, tt, kk, lane, &a8[tt][kk], &a16[tt][kk]);
    } else {
        // The group's 8 ranges x 8 isometries are 64 contiguous rows of rng8;
        // every load of a thread is issued before the first LDS write, so
        // the staging costs one memory latency, not one per chunk.
        const GLB int8_t *src = T.rng8 + (size_t)g * 64 * G::KMEM;
        constexpr int CPR = G::KMEM / 16; // 16-byte chunks per row
        constexpr int PER = 64 * CPR / 256; // chunks per thread
        static_assert(64 * CPR % 256 == 0, "chunks divide over the block");
        v4i stg[PER];
#pragma unroll
        for (int j = 0; j < PER; j++) {
            const int cc = threadIdx.x + 256 * j;
            stg[j] = *(const GLB v4i *)(src + (size_t)cc * 16);
        }
#pragma unroll
        for (int j = 0; j < PER; j++) {
            const int cc = threadIdx.x + 256 * j, row = cc / CPR, col = cc - row * CPR;
            *(v4i *)(s_rng + row * ASTRIDE + col * 16) = stg[j];
        }
    }
    __syncthreads();
    // Row of the LDS range operand feeding lane `lane` in row tile tt.
    const int am = lane & 31;
    const int arow0 = (((am >> 2) & 1) * 4 + (am & 3)) * 8 + (am >> 3);

    // With c.sample > 1 (threshold pre-pass) the block covers every
    // sample-th tile of all of its group's positions (whatever its chunk):
    // tile index i below stands for tile i * sample.
    const uint32_t sample = S >= kPrepassMinSize ? c.sample : 1u;
    const bool whole = sample > 1;
    const uint32_t ntiles = (T.P_pad / 32 + sample - 1) / sample;
    const uint32_t tpc = whole ? ntiles : tiles_per_chunk;
    const uint32_t t_begin = whole ? 0 : chunk * tpc;
    uint32_t t_end = t_begin + tpc;
    if (t_end > ntiles) t_end = ntiles;
    unsigned long long qtotal = 0, stage3 = 0; // stage-1 survivors (wave), stage-3 fits (lane)
    uint32_t tcount = 0;
    uint32_t qn = 0; // wave-uniform queue length
    uint2 *q = s_q[wave];
    BestRec *wb = s_wb[wave];

    // Stage 2 + 3 over the queue entries [b0, b0 + cnt), cnt <= 64.
    auto process = [&](uint32_t b0, uint32_t cnt) {
        qtotal += cnt; // every queued survivor is drained exactly once
        const bool act = (uint32_t)lane < cnt;
        const uint2 e = q[b0 + (act ? lane : 0)];
        // The entry: {t1 | range << 28, position * 8 + isometry}; at the
        // binary16 sizes {X0 + 2^24 | range << 28, search slot * 8 +
        // isometry} (X0 = t1 - m S1, |X0| < 2^24), completed below from the
        // slot's record {offset, p, S1, S2} and the range's centre m.
        int32_t t1 = (int32_t)(e.x & 0x0fffffffu);
        const uint32_t slot = e.x >> 28;
        uint32_t key = e.y, p = key >> 3;
        bool keep = false, win = false;
        double rms = 0.0, sum = 0.0;
        if (act) {
            int32_t S1, S2;
            if constexpr (F16) {
                const uint4 rec = ld_u4(rec16(T.posinfo, p), 0);
                p = rec.y;
                key = p * 8 + (key & 7);
                S1 = (int32_t)rec.z;
                S2 = (int32_t)rec.w;
            } else {
                S1 = T.posx[2 * p];
                S2 = T.posx[2 * p + 1];
            }
            const double Di = T.posd[2 * p], invDi = T.posd[2 * p + 1];
            const int64_t t0 = s_T0[slot], t2 = s_T2[slot];
            if constexpr (F16) t1 += (int32_t)((t0 + n / 2) >> LOG2N) * S1 - (1 << 24);
            keep = !c.stage2 || stage2_keep(c, nd, inv_n, t1, S1, t0, Di, invDi, s_A[slot],
                                            __uint_as_float(s_thr[slot]));
            if (keep) {
                uint32_t qa, qb;
                fit_f64(n, S1, S2, t0, t1, t2, c.max_alfa, c.bits_alfa, c.bits_beta, &qa, &qb,
                        &rms, &sum);
                if constexpr (RANKED) {
                    const uint32_t kk = c.topk;
                    const TopRec &last = s_top[wave][slot][kk - 1];
                    win = s_topn[wave][slot] < kk || better(rms, key, last.rms, last.key);
                } else {
                    win = better(rms, key, wb[slot].rms, wb[slot].key);
                }
            }
        }
        stage3 += keep ? 1 : 0;
        unsigned long long m = __ballot(win);
        if (m) seen = 0xffffffffu; // this wave lowered a threshold: its own limits are stale now
        if constexpr (!RANKED) {
            // Several lanes of one range often beat the wave's best of
            // before the drain, and each would take a trip of the serial
            // loop below.  Only those whose rms has the least high word
            // among the range's winners can be its new best: the others
            // drop here (one LDS min and one read, in the wave's order).
            if (m) {
                if ((uint32_t)lane < 8) s_hmin[wave][lane] = 0xffffffffu;
                const uint32_t hw = (uint32_t)((uint64_t)__double_as_longlong(rms) >> 32);
                if (win) {
                    atomicMin(&s_hmin[wave][slot], hw);
                    win = lds_load(&s_hmin[wave][slot]) == hw;
                }
                m = __ballot(win);
            }
        }
        // Serialise the updates of the wave's best records.
        while (m) {
            const int l = __builtin_ctzll(m);
            m &= m - 1;
            const double r_l = __shfl(rms, l), s_l = __shfl(sum, l);
            const uint32_t k_l = __shfl(key, l), sl_l = __shfl(slot, l);
            const int32_t t_l = __shfl(t1, l);
            if constexpr (RANKED) {
                // retain_top_k (encode.rs:676-685) on the wave's list, one
                // entry per lane: the entries ahead of the new one are
                // counted by a ballot and those behind it move down one slot
                // in parallel.  Once the list holds c.topk entries its last
                // one bounds the range.
                TopRec *L = s_top[wave][sl_l];
                const uint32_t kk = c.topk;
                const uint32_t nn = lds_load(&s_topn[wave][sl_l]);
                TopRec e = {};
                bool ahead = false;
                if ((uint32_t)lane < nn) {
                    e = L[lane];
                    ahead = !better(r_l, k_l, e.rms, e.key);
                }
                const uint32_t pos = (uint32_t)__popcll(__ballot(ahead));
                if (pos < kk) {
                    if ((uint32_t)lane >= pos && (uint32_t)lane < nn && (uint32_t)lane + 1 < kk)
                        L[lane + 1] = e;
                    if (lane == 0) {
                        L[pos] = TopRec{r_l, k_l, t_l, tau_bits(s_l, n)};
                        const uint32_t nn2 = nn < kk ? nn + 1 : kk;
                        s_topn[wave][sl_l] = nn2;
                        if (nn2 == kk) lower(sl_l, L[kk - 1].tau);
                    }
                }
            } else if (lane == 0 && better(r_l, k_l, wb[sl_l].rms, wb[sl_l].key)) {
                wb[sl_l].rms = r_l;
                wb[sl_l].key = k_l;
                wb[sl_l].t1 = t_l;
                lower(sl_l, tau_bits(s_l, n));
            }
        }
    };

    // Software pipeline: the position record is loaded two tiles ahead and
    // (for sizes whose A fragments live in registers) the B fragments one
    // tile ahead, so neither load's latency sits in front of the MFMAs.
    struct Frag {
        uint4 pin; // {window byte offset, S1, position index or ~0, sqrt(Di) bits}; binary16: {offset}
        long l8[NK], h8[NK];
        v4i l16[NK], h16[NK];
        v8h wh[NKH]; // binary16 sizes
        v8h wb;      // binary16 sizes: the bias K-step's B operand (zero in lanes 32..63)
    };
    // The column (lane % 32) is recomputed at every use: kept live across the
    // tile loop it is the value the register allocator spills, and the
    // reload's wait would also wait for the prefetches issued before it.
    // Binary16 sizes read only the offset (their queue entries name the
    // search slot, not the position): a loaded word the tile never uses
    // would be dead at once, and the register allocator would reuse its
    // register while the load is in flight, waiting for the prefetch; and
    // with one word, the record of tile + 8 can load into the register of
    // tile + 4's once that one's fragment loads are issued, where a record
    // also read in its own tile would rotate through three registers in a
    // loop unrolled by two, a copy (and a wait for the load) per trip.
    // lc: lane_col(), computed once for the loads of one prefetch.
    auto load_pin = [&](uint32_t tile, uint32_t lc) {
        const GLB u32x4 *rec = rec16(T.posinfo, tile * sample * 32 + lc);
        if constexpr (F16) {
            return make_uint4(*(const GLB uint32_t *)rec, 0u, 0u, 0u); // the offset
        } else {
            return ld_u4(rec, 0);
        }
    };
    auto fetch_b = [&](const uint4 pin, const uint32_t tile, uint32_t lc, Frag &f) {
        f.pin = pin;
        if constexpr (F16) {
#pragma unroll
            for (int kk = 0; kk < NKH; kk++) f.wh[kk] = load_bh<S>(T.wf, pin.x, T.pitch, kk, h);
            // lane half h: record half h = k 8h..8h+7, one load straight into
            // the operand (a vector assembled from a load and constants
            // would wait for the load where it is assembled)
            f.wb = __builtin_bit_cast(v8h, *rec16(T.posbias, (tile * sample * 32 + lc) * 2 + h));
        } else if constexpr (G::PRELOAD_A) {
#pragma unroll
            for (int kk = 0; kk < NK; kk++) {
                load_b<S>(T.lo, pin.x, T.pitch, kk, h, &f.l8[kk], &f.l16[kk]);
                load_b<S>(T.hi, pin.x, T.pitch, kk, h, &f.h8[kk], &f.h16[kk]);
            }
        }
    };
    Frag nxt = {};
    uint4 pin_ahead = make_uint4(0, 0, 0, 0);
    const bool has_tiles = t_begin + wave < t_end;
    if (has_tiles) fetch_b(load_pin(t_begin + wave, lane_col()), t_begin + wave, lane_col(), nxt);
    if (t_begin + wave + 4 < t_end) pin_ahead = load_pin(t_begin + wave + 4, lane_col());
    // Sizes >= 16: window fragments in flight PD K-steps ahead, across tiles.
    constexpr int PD = NK < 8 ? NK : 8;
    v4i pbl[PD], pbh[PD];
    // The lane's part of its fragments' offsets (sizes >= 16), opaque to
    // the compiler so that it is not folded back into a row index times
    // the pitch per load.
    uint32_t hoff = S == 16 ? (uint32_t)h * T.pitch : 16u * (uint32_t)h;
    asm volatile("" : "+v"(hoff));
    if constexpr (!G::PRELOAD_A) {
#pragma unroll
        for (int d = 0; d < PD; d++) {
            pbl[d] = (v4i){};
            pbh[d] = (v4i){};
            if (has_tiles) {
                pbl[d] = load_b16<S>(T.lo, nxt.pin.x + hoff, T.pitch, d);
                pbh[d] = load_b16<S>(T.hi, nxt.pin.x + hoff, T.pitch, d);
            }
        }
    }
    const bool exchange = !whole && nchunks > 1;

    // One tile: `cur` holds its window fragments and position record, the
    // next tile's are fetched into `nxt`.  The loop below runs two tiles per
    // trip with the two fragment sets swapping roles, so no register copy
    // (and no wait for the prefetch) sits between the tiles.
    auto step = [&](const uint32_t tile, const Frag &cur, Frag &nxt) __attribute__((always_inline)) {
        // Exchange thresholds with the other blocks of this range group.
        if (exchange && (tcount & 15) == 15 && lane < 8) {
            uint32_t r = g * 8 + lane;
            if (r < T.nr) lower(lane, atomicMin((uint32_t *)&T.gthr[r], lds_load(&s_thr[lane])));
        }
        // Whether any threshold dropped since the limits were computed
        // (read before the MFMAs so the LDS latency overlaps them).
        // (Binary16 sizes read the epoch during the previous tile, hiding
        // the LDS latency; this wave's own drops reset `seen` at once.)
        const uint32_t ep = __builtin_amdgcn_readfirstlane(F16 ? ep_pre : lds_load(&s_epoch));
        // (cur: this tile; nxt: filled for tile + 4).  Sizes with register
        // A fragments issue these loads after the first MFMAs: every load in
        // flight at an MFMA is then one the MFMA needs, so the compiler's
        // in-order wait count cannot make it wait for the prefetch too.
        // Unconditional loads (a tile past the end reads the last tile
        // again, unused): the wait before the next tile's MFMAs then counts
        // only the loads it needs, where loads behind a branch made it wait
        // for every load in flight.
        auto prefetch = [&]() __attribute__((always_inline)) {
            const uint32_t lc = lane_col();
            fetch_b(pin_ahead, min(tile + 4, t_end - 1), lc, nxt);
            pin_ahead = load_pin(min(tile + 8, t_end - 1), lc);
        };
        if constexpr (!G::PRELOAD_A) prefetch();
        const uint32_t p = cur.pin.z; // position index, ~0 for padding
        const uint32_t boff = cur.pin.x;
        const int32_t S1p = (int32_t)cur.pin.y;
        const float sqDi = __uint_as_float(cur.pin.w);
        // INT_MIN, or INT_MAX for a padding slot
        const int32_t pmin = p < T.P ? (int32_t)0x80000000 : 0x7fffffff;

        // Per-range sqrt(A - n tau), recomputed only when a threshold dropped
        // since the last time (a scalar branch).  Thresholds only drop, so
        // one read a moment stale is a looser, still valid, limit.
        {
            if (ep != seen) {
                seen = ep;
                if constexpr (F16) abias = make_abias(lds_load(&s_thr[brl]));
#pragma unroll
                for (int i = 0; i < (F16 ? 0 : 4); i++) {
                    const float thr = __uint_as_float(lds_load(&s_thr[4 * h + i]));
                    float amt;
                    if constexpr (INT_LIMIT) {
                        amt = Alo[i] - (float)n * thr;
                    } else {
                        const double nthr = nd * (double)thr;
                        amt = nthr < A[i] ? (float)(A[i] - nthr) : 0.0f;
                    }
                    // sqrt within 1 ulp; (1 - 2e-5) keeps nt below
                    // sqrt((A - n tau) Di) after the product's rounding
                    // (sqDi is sqrt(Di) rounded down).
                    const bool prune = amt > 0.0f; // else A - n tau <= 0: no pruning
                    sa[i] = prune ? __builtin_amdgcn_sqrtf(amt) * (1.0f - 2e-5f)
                                  : (INT_LIMIT ? 1.0f : -1.0f);
                    if constexpr (INT_LIMIT)
                        noff[i] = -1 - ((F16 ? 0 : off_of(i)) + (rmin[i] == (int32_t)0x80000000
                                                         ? (prune ? 0 : (1 << 27))
                                                         : -(1 << 28)));
                }
            }
        }
        // Stage-1 limit per (range, position) on v = 4*hi + lo.  With
        // X = n t1 - S1 t0, a = A - n tau and u = alfa / 4 <= U (cap_u), a
        // candidate's SSE is at least (A - 2 u X + u^2 Di) / n, so it can
        // reach tau only if X >= sqrt(a Di) when a <= U^2 Di, else X >=
        // (a + U^2 Di) / (2U) = sqrt(a Di) + (sqrt(a) - U sqrt(Di))^2 / (2U).
        // nt is that bound from sa <= sqrt(a) and sqDi <= sqrt(Di); it rises
        // with both, and the (1 - 2e-5) in sa covers the rounding here.
        int32_t lim[4] = {0, 0, 0, 0};
        const int32_t o128 = (int32_t)((uint32_t)S1p << 7);
        // what the limit adds besides noff: -128 S1 (split planes), or 2^28
        // at a padding position (binary16)
        const int32_t padj = F16 ? (p < T.P ? 0 : (1 << 28)) : -o128;
        const float usq = CAP_LIMIT ? c.cap_u * sqDi : 0.0f;
#pragma unroll
        for (int i = 0; i < (F16 ? 0 : 4); i++) {
            const float sab = INT_LIMIT ? sa[i] : fabsf(sa[i]);
            float nt = sab * sqDi; // <= sqrt((A - n tau) Di) when pruning
            if constexpr (CAP_LIMIT) {
                const float d = fmaxf(sab - usq, 0.0f);
                nt = fmaf(d * d, c.cap_k, nt);
            }
            if constexpr (INT_LIMIT) {
                // S1*t0 + nt < 2^31 for sizes <= 8 (nt <= sqrt(A Di) + 8 A),
                // S1 and t0 < 2^24.
                const uint32_t X = __umul24((uint32_t)S1p, (uint32_t)T0[i]) + (uint32_t)nt;
                lim[i] = (int32_t)(X >> LOG2N) + noff[i] + padj;
            } else {
                const double X = ((double)S1p * (double)T0[i] + (double)nt) * inv_n;
                const int32_t L = (int32_t)floor(X) - 1 - (off_of(i) + o128);
                lim[i] = max(sa[i] < 0.0f ? (int32_t)0x80000000 : L, max(rmin[i], pmin));
            }
        }

        // Epilogue of one row tile.  Sizes >= 16: the queue is drained after
        // every DRAIN = 4 values per lane (size 16: one range's four
        // isometries, compared by their maximum; size 32: four ranges of one
        // isometry, a compare each); only a group some lane passes goes
        // value by value and appends (wave ballot + mbcnt).  Sizes <= 8: see
        // below.
        constexpr int DRAIN = Queue<S>::DRAIN;
        constexpr int GW = DRAIN < 8 ? DRAIN : 8; // values per compare group
        static_assert(16 % DRAIN == 0 && DRAIN % GW == 0, "groups tile a row tile's 16 values");
        static_assert(64 * DRAIN + 64 <= Queue<S>::CAP, "a drain interval fits the queue");
        // Sizes <= 8: one compare per (range, position).  The limit does not
        // depend on the isometry, so the row tile's 4 values of a range are
        // tested by their maximum (a v_max3 and a v_max); only a range some
        // lane passes is tested value by value.  Against a compare per value
        // this also drops most of the scalar mask ORs and per-value branches
        // (s = 4 launch 0.367 -> 0.330 ms at 512 x 512).
        // v: the row tile's values, lv: the range's limits in the same form
        // (int32 4 hi + lo and its limit; binary16: X0 = t1 - m S1 and
        // max_j nb_j, binary32); t1of(t, i) gives a survivor's queue word:
        // its exact t1, or X0 + 2^24 at the binary16 sizes; key0() the key
        // of the lane's column without the isometry (position or, binary16,
        // search slot times 8).
        auto epilogue16 = [&](const auto (&v)[16], const auto (&lv)[4], const int tt,
                              auto &&t1of, auto &&key0) {
            unsigned long long mr[4], any = 0;
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const auto mx = vmax(vmax(vmax(v[i], v[i + 4]), v[i + 8]), v[i + 12]);
                mr[i] = __ballot(mx > lv[i]);
                any |= mr[i];
            }
            if (any == 0) return;
            const uint32_t k0 = key0();
#pragma unroll
            for (int i = 0; i < 4; i++) {
                if (!mr[i]) continue;
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    const int t = 4 * j + i;
                    const bool sv = v[t] > lv[i];
                    const unsigned long long m = __ballot(sv);
                    if (m) {
                        const uint32_t pre = __builtin_amdgcn_mbcnt_hi(
                            (uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0));
                        if (sv) {
                            // (the word's low bits are below 2^28: + is |)
                            q[qn + pre] = make_uint2((uint32_t)t1of(t, i) + ((uint32_t)(4 * h + i) << 28),
                                                     k0 + 4 * tt + j);
                        }
                        qn += (uint32_t)__popcll(m);
                    }
                }
            }
            if (qn >= 64) {
                uint32_t b0 = 0;
                for (; qn - b0 >= 64; b0 += 64) process(b0, 64);
                const uint32_t rem = qn - b0;
                const uint2 tail = q[b0 + ((uint32_t)lane < rem ? lane : 0)];
                if ((uint32_t)lane < rem) q[lane] = tail;
                qn = rem;
            }
        };
        auto epilogue = [&](const v16i &ah, const v16i &al, const int tt) {
            if constexpr (DRAIN == 16) {
                int32_t v[16];
#pragma unroll
                for (int t = 0; t < 16; t++) v[t] = 4 * ah[t] + al[t];
                epilogue16(v, lim, tt, [&](int t, int i) { return v[t] + off_of(i) + o128; },
                           [&]() { return p * 8; });
                return;
            }
            // Size 16: drain interval grp = range grp of the lane half, its
            // four isometries' values tested by their maximum against the
            // range's limit (which does not depend on the isometry), one
            // compare and one scalar test; only a range some lane passes
            // goes value by value.  Size 32 keeps groups of four ranges of
            // one isometry, a compare per value: the range grouping raised
            // its register use to the limit and measured slower there
            // (0.122 -> 0.127 ms, where size 16 went 0.088 -> 0.085 ms).
#pragma unroll
            for (int grp = 0; grp < 16 / DRAIN; grp++) {
                if constexpr (S == 16) {
                    static_assert(DRAIN == 4, "one range (four isometries) per drain interval");
                    const int i = grp;
                    int32_t v[4];
#pragma unroll
                    for (int e = 0; e < 4; e++) v[e] = 4 * ah[i + 4 * e] + al[i + 4 * e];
                    const int32_t mx = max(max(v[0], v[1]), max(v[2], v[3]));
                    if (__ballot(mx > lim[i]) == 0) continue;
#pragma unroll
                    for (int e = 0; e < 4; e++) {
                        const bool sv = v[e] > lim[i];
                        const unsigned long long m = __ballot(sv);
                        if (m) {
                            const uint32_t pre = __builtin_amdgcn_mbcnt_hi(
                                (uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0));
                            if (sv) {
                                const uint32_t slot = 4 * h + i;
                                q[qn + pre] = make_uint2((uint32_t)(v[e] + off_of(i) + o128) | (slot << 28),
                                                         p * 8 + 4 * tt + e);
                            }
                            qn += (uint32_t)__popcll(m);
                        }
                    }
                } else {
#pragma unroll
                    for (int g0 = 0; g0 < DRAIN; g0 += GW) {
                        int32_t v[GW];
                        unsigned long long msk[GW], any = 0;
#pragma unroll
                        for (int e = 0; e < GW; e++) {
                            const int t = grp * DRAIN + g0 + e;
                            v[e] = 4 * ah[t] + al[t];
                            msk[e] = __ballot(v[e] > lim[t & 3]);
                            any |= msk[e];
                        }
                        if (any == 0) continue;
                        // The compare masks stay in scalar registers: a value
                        // without survivors costs one scalar test here.
#pragma unroll
                        for (int e = 0; e < GW; e++) {
                            const int t = grp * DRAIN + g0 + e, i = t & 3;
                            const unsigned long long m = msk[e];
                            if (m) {
                                const bool sv = v[e] > lim[i]; // the compare behind m
                                const uint32_t pre = __builtin_amdgcn_mbcnt_hi(
                                    (uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0));
                                if (sv) {
                                    const uint32_t slot = 4 * h + i;
                                    q[qn + pre] = make_uint2((uint32_t)(v[e] + off_of(i) + o128) | (slot << 28),
                                                             p * 8 + 4 * tt + (t >> 2));
                                }
                                qn += (uint32_t)__popcll(m);
                            }
                        }
                    }
                }
                // Drain full batches; keep the tail (< 64) for later.
                if (qn >= 64) {
                    uint32_t b0 = 0;
                    for (; qn - b0 >= 64; b0 += 64) process(b0, 64);
                    const uint32_t rem = qn - b0;
                    const uint2 tail = q[b0 + ((uint32_t)lane < rem ? lane : 0)];
                    if ((uint32_t)lane < rem) q[lane] = tail;
                    qn = rem;
                }
            }
        };

        if constexpr (F16) {
            // The bias K-step gives, per (range, position), the four
            // right-hand sides nb_j (items i + 4j of the lane's range i), the
            // same for both row tiles: one instruction per tile.  Then, row
            // tile by row tile, one binary16 instruction per K-step gives X0
            // = t1 - m S1 exactly, and a candidate passes stage 1 when X0 >
            // max_j nb_j, tested on the maximum over isometries.  (Issuing
            // row tile 1's chain before row tile 0's epilogue measured
            // slower: its accumulators stay live across the queue drain.)
            auto chain = [&](int tt) __attribute__((always_inline)) {
                v16f z = (v16f){};
#pragma unroll
                for (int kk = 0; kk < NKH; kk++)
                    z = __builtin_amdgcn_mfma_f32_32x32x16_f16(ahf[tt][kk], cur.wh[kk], z, 0, 0, 0);
                return z;
            };
            const v16f nbt = __builtin_amdgcn_mfma_f32_32x32x16_f16(abias, cur.wb, (v16f){}, 0, 0, 0);
            const v16f acc0 = chain(0);
            prefetch();
            // the epoch for the next tile, read now so its latency is hidden
            // (kept in a VGPR: a readfirstlane here would wait for the read)
            ep_pre = lds_load(&s_epoch);
            float nb[4];
#pragma unroll
            for (int i = 0; i < 4; i++) nb[i] = vmax(vmax(vmax(nbt[i], nbt[i + 4]), nbt[i + 8]), nbt[i + 12]);
#pragma unroll
            for (int tt = 0; tt < 2; tt++) {
                const v16f acc = tt == 0 ? acc0 : chain(1);
                float v[16];
#pragma unroll
                for (int t = 0; t < 16; t++) v[t] = acc[t];
                epilogue16(
                    v, nb, tt, [&](int t, int) { return (int32_t)v[t] + (1 << 24); },
                    [&]() { return (tile * sample * 32 + lane_col()) * 8; });
            }
        } else if constexpr (G::PRELOAD_A) {
            // Row tile by row tile: row tile 1's MFMAs run while row tile
            // 0's epilogue uses the VALU, and only one tile's accumulators
            // are live at a time.
#pragma unroll
            for (int tt = 0; tt < 2; tt++) {
                v16i ah = (v16i){}, al = (v16i){};
#pragma unroll
                for (int kk = 0; kk < NK; kk++) {
                    if constexpr (G::SMALLK) {
                        ah = __builtin_amdgcn_mfma_i32_32x32x16_i8(a8[tt][kk], cur.h8[kk], ah, 0, 0, 0);
                        al = __builtin_amdgcn_mfma_i32_32x32x16_i8(a8[tt][kk], cur.l8[kk], al, 0, 0, 0);
                    } else {
                        ah = __builtin_amdgcn_mfma_i32_32x32x32_i8(a16[tt][kk], cur.h16[kk], ah, 0, 0, 0);
                        al = __builtin_amdgcn_mfma_i32_32x32x32_i8(a16[tt][kk], cur.l16[kk], al, 0, 0, 0);
                    }
                }
                if (tt == 0) prefetch();
                epilogue(ah, al, tt);
            }
        } else {
            v16i acc_lo[2], acc_hi[2];
#pragma unroll
            for (int tt = 0; tt < 2; tt++) {
                acc_lo[tt] = (v16i){};
                acc_hi[tt] = (v16i){};
            }
            const bool has_next = tile + 4 < t_end;
            const uint32_t nboff = nxt.pin.x;
            // Runtime loop over blocks of PD K-steps (only the block is
            // unrolled, so the loads in flight stay at PD per plane).
#pragma unroll 1
            for (int k0 = 0; k0 < NK; k0 += PD) {
#pragma unroll
                for (int d = 0; d < PD; d++) {
                    const int kk = k0 + d;
                    const v4i a0 = *(const v4i *)(s_rng + arow0 * ASTRIDE + kk * 32 + h * 16);
                    const v4i a1 = *(const v4i *)(s_rng + (arow0 + 4) * ASTRIDE + kk * 32 + h * 16);
                    acc_hi[0] = __builtin_amdgcn_mfma_i32_32x32x32_i8(a0, pbh[d], acc_hi[0], 0, 0, 0);
                    acc_hi[1] = __builtin_amdgcn_mfma_i32_32x32x32_i8(a1, pbh[d], acc_hi[1], 0, 0, 0);
                    acc_lo[0] = __builtin_amdgcn_mfma_i32_32x32x32_i8(a0, pbl[d], acc_lo[0], 0, 0, 0);
                    acc_lo[1] = __builtin_amdgcn_mfma_i32_32x32x32_i8(a1, pbl[d], acc_lo[1], 0, 0, 0);
                    // Refill the slot with the fragment PD steps ahead: of
                    // this tile, of the next one, or (no next tile) a
                    // harmless re-read of this tile's first step.  The
                    // loads are unconditional, so the wait before each
                    // step's MFMAs counts only the loads it needs and the
                    // later ones stay in flight.
                    const bool here = kk + PD < NK;
                    const uint32_t sboff = (here || !has_next ? boff : nboff) + hoff;
                    const int skk = here ? kk + PD : has_next ? kk + PD - NK : 0;
                    pbl[d] = load_b16<S>(T.lo, sboff, T.pitch, skk);
                    pbh[d] = load_b16<S>(T.hi, sboff, T.pitch, skk);
                }
            }
            epilogue(acc_hi[0], acc_lo[0], 0);
            epilogue(acc_hi[1], acc_lo[1], 1);
        }
        tcount++;
    };
    Frag alt = {};
    for (uint32_t tile = t_begin + wave; tile < t_end; tile += 8) {
        step(tile, nxt, alt);
        if (tile + 4 >= t_end) break;
        step(tile + 4, alt, nxt);
    }
    if (qn > 0) process(0, qn);
    if (lane == 0) atomicAdd(&s_count[0], qtotal);
    atomicAdd(&s_count[1], stage3);
    __syncthreads();
    if (threadIdx.x == 0 && c.counters) {
        atomicAdd((unsigned long long *)&glb(c.counters)[LOG2S], s_count[0]);
        atomicAdd((unsigned long long *)&glb(c.counters)[8 + LOG2S], s_count[1]);
    }

    if (threadIdx.x < 8) {
        const uint32_t sl = threadIdx.x, r = g * 8 + sl;
        if (r < T.nr) {
            // The block's tightest threshold for later blocks and passes.
            atomicMin((uint32_t *)&T.gthr[r], lds_load(&s_thr[sl]));
            if constexpr (RANKED) {
                // The four waves' lists merged: the chunk's best c.topk,
                // padded with empty records (key ~0, rms +inf).
                if (sample == 1) {
                    const uint32_t kk = c.topk;
                    uint32_t hd[4] = {0, 0, 0, 0};
                    const size_t base = ((size_t)r * nchunks + chunk) * kk;
                    for (uint32_t i = 0; i < kk; i++) {
                        int bw = -1;
                        for (int wv = 0; wv < 4; wv++) {
                            if (hd[wv] >= s_topn[wv][sl]) continue;
                            const TopRec &e = s_top[wv][sl][hd[wv]];
                            if (bw < 0 || better(e.rms, e.key, s_top[bw][sl][hd[bw]].rms,
                                                 s_top[bw][sl][hd[bw]].key))
                                bw = wv;
                        }
                        if (bw < 0) {
                            st_partial(T.partial, base + i, __longlong_as_double(0x7ff0000000000000ll),
                                       0xffffffffu, 0);
                        } else {
                            const TopRec &e = s_top[bw][sl][hd[bw]++];
                            st_partial(T.partial, base + i, e.rms, e.key, e.t1);
                        }
                    }
                }
            } else {
                BestRec b = s_wb[0][sl];
                for (int wv = 1; wv < 4; wv++) {
                    const BestRec &e = s_wb[wv][sl];
                    if (better(e.rms, e.key, b.rms, b.key)) b = e;
                }
                // The pre-pass only tightens thresholds; the full pass writes
                // every partial.
                if (sample == 1)
                    st_partial(T.partial, (size_t)r * nchunks + chunk, b.rms, b.key, b.t1);
            }
        }
    }
}

// ---------------------------------------------------------------- FP32 baseline

// Plain FP32 search: a wave covers 4 ranges x 8 isometries (32 rows) and
// 64 positions (one per lane).  The range operand is wave-uniform (scalar
// loads), the window sample is per lane; 32 FMAs per k.  Every candidate
// then takes the binary32 fit of encode.rs:295-337 and competes by
// (rms_f32, key).  No certification: results may differ from the binary64
// reference wherever binary32 rounding moves a quantisation decision or an
// argmin near-tie (and t1 is inexact above 2^24).
template <int S>
__global__ __launch_bounds__(256) void k_search_fp32(const TableDesc *__restrict__ tabs,
                                                     const Work3 *__restrict__ work, Common c)
{
    constexpr int K = S * S;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const Work3 wk = work[blockIdx.x];
    const TableG T = view(tabs[wk.x]);
    const uint32_t chunk = wk.y, g = wk.z;
    __shared__ BestRec s_best[4][64][4];

    float T0[4], T2[4];
    bool rvalid[4];
#pragma unroll
    for (int i = 0; i < 4; i++) {
        uint32_t r = g * 4 + i;
        rvalid[i] = r < T.nr;
        uint32_t rr = rvalid[i] ? r : 0;
        T0[i] = (float)T.rstat[2 * rr];
        T2[i] = (float)T.rstat[2 * rr + 1];
    }
    float brms[4];
    uint32_t bkey[4];
#pragma unroll
    for (int i = 0; i < 4; i++) {
        brms[i] = __int_as_float(0x7f800000);
        bkey[i] = 0xffffffffu;
    }
    const float s0 = (float)K;
    const uint32_t ntiles = (T.P + 63) / 64;
    const uint32_t t_begin = chunk * T.tiles_per_chunk[1];
    uint32_t t_end = t_begin + T.tiles_per_chunk[1];
    if (t_end > ntiles) t_end = ntiles;
    const GLB float *Ar = T.rngf + (size_t)g * K * 32;

    for (uint32_t tile = t_begin + wave; tile < t_end; tile += 4) {
        const uint32_t p = tile * 64 + lane;
        const bool pv = p < T.P;
        const uint32_t pc = pv ? p : T.P - 1;
        const int32_t S1i = T.posx[2 * pc], S2i = T.posx[2 * pc + 1];
        const uint32_t pi = pc / T.npc, pj = pc - pi * T.npc;
        const uint32_t dr = (pi * T.shift) >> 1, dc = (pj * T.shift) >> 1;
        const GLB float *D = T.df + (size_t)dr * T.wc + dc;
        float acc[32];
#pragma unroll
        for (int r = 0; r < 32; r++) acc[r] = 0.0f;
        for (int u = 0; u < S; u++) {
#pragma unroll 8
            for (int v = 0; v < S; v++) {
                float d = D[(size_t)u * T.wc + v];
                const GLB float *ar = Ar + (size_t)(u * S + v) * 32;
#pragma unroll
                for (int r = 0; r < 32; r++) acc[r] = __builtin_fmaf(ar[r], d, acc[r]);
            }
        }
        if (!pv) continue;
        const float S1 = (float)S1i, S2 = (float)S2i;
#pragma unroll
        for (int r = 0; r < 32; r++) {
            const int i = r & 3;
            const uint32_t iso = r >> 2;
            uint32_t qa, qb;
            float rms;
            fit_f32(s0, S1, S2, T0[i], acc[r], T2[i], (float)c.max_alfa, c.bits_alfa,
                    c.bits_beta, &qa, &qb, &rms);
            uint32_t key = p * 8 + iso;
            if (rms < brms[i] || (rms == brms[i] && key < bkey[i])) {
                brms[i] = rms;
                bkey[i] = key;
            }
        }
    }
#pragma unroll
    for (int i = 0; i < 4; i++) {
        s_best[wave][lane][i].rms = (double)brms[i];
        s_best[wave][lane][i].key = bkey[i];
    }
    __syncthreads();
    if (threadIdx.x < 4) {
        uint32_t i = threadIdx.x, r = g * 4 + i;
        if (r < T.nr) {
            double b = __longlong_as_double(0x7ff0000000000000ll);
            uint32_t k = 0xffffffffu;
            for (int wv = 0; wv < 4; wv++)
                for (int l = 0; l < 64; l++) {
                    const BestRec &e = s_best[wv][l][i];
                    if (better(e.rms, e.key, b, k)) {
                        b = e.rms;
                        k = e.key;
                    }
                }
            st_partial(T.partial, (size_t)r * T.nchunks[1] + chunk, b, k, 0);
        }
    }
}

// ---------------------------------------------------------------- merge

__device__ __forceinline__ mars_hip_candidate winner_record(const TableG &T, const Common &cm,
                                                            uint32_t r, uint32_t k, int64_t t1)
{
    mars_hip_candidate c = {};
    const uint32_t p = k >> 3, iso = k & 7, S = T.size;
    const uint32_t pi = p / T.npc, pj = p - pi * T.npc;
    c.dom_row = pi * T.shift;
    c.dom_col = pj * T.shift;
    c.isometry = iso;
    c.s0 = (int64_t)S * S;
    c.s1_x4 = T.posx[2 * p];
    c.s2_x16 = T.posx[2 * p + 1];
    c.t0 = T.rstat[2 * r];
    c.t1_x4 = t1;
    c.t2 = T.rstat[2 * r + 1];
    uint32_t qa, qb;
    double rms, sum;
    fit_f64(c.s0, c.s1_x4, c.s2_x16, c.t0, c.t1_x4, c.t2, cm.max_alfa, cm.bits_alfa, cm.bits_beta,
            &qa, &qb, &rms, &sum);
    c.qalfa = qa;
    c.qbeta = qb;
    c.rms = rms;
    return c;
}

__device__ __forceinline__ void write_winner(const TableG &T, const Common &cm, uint32_t r,
                                             uint32_t k, int64_t t1)
{
    st_cand(T.out, r, winner_record(T, cm, r, k, t1));
    T.found[r] = 1;
}

// Ranked merge, work item (table, range), one wave: the range's chunk lists
// (each ascending by (rms, key), c.topk long, padded with key ~0) merged into
// its best c.topk -- lane j holds chunk j's head (at most 64 chunks) and the
// wave takes the smallest head c.topk times -- each written with its exact
// RawMoments and fit to out[r * topk + i]; found[r] = how many exist.
__global__ __launch_bounds__(64) void k_merge_ranked(const TableDesc *__restrict__ tabs,
                                                     const Work2 *__restrict__ work, Common cm)
{
    const Work2 wk = work[blockIdx.x];
    const TableG T = view(tabs[wk.x]);
    const uint32_t r = wk.y, lane = threadIdx.x, kk = cm.topk;
    const uint32_t nchunks = T.nchunks[0];
    const bool mine = lane < nchunks && T.P > 0;
    const size_t base = ((size_t)r * nchunks + lane) * kk;
    uint32_t head = 0;
    uint32_t nfound = 0;
    for (uint32_t i = 0; i < kk; i++) {
        double rms = __longlong_as_double(0x7ff0000000000000ll);
        uint32_t key = 0xffffffffu;
        int32_t t1 = 0;
        if (mine && head < kk) {
            const Partial e = ld_partial(T.partial, base + head);
            rms = e.rms;
            key = e.key;
            t1 = e.t1;
        }
        uint32_t who = lane;
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) {
            const double orms = __shfl_xor(rms, off, 64);
            const uint32_t okey = __shfl_xor(key, off, 64), owho = __shfl_xor(who, off, 64);
            const int32_t ot1 = __shfl_xor(t1, off, 64);
            if (better(orms, okey, rms, key)) {
                rms = orms;
                key = okey;
                who = owho;
                t1 = ot1;
            }
        }
        if (key == 0xffffffffu) break; // every list is exhausted
        if (lane == who) head++;
        if (lane == 0) st_cand(T.out, (size_t)r * kk + i, winner_record(T, cm, r, key, t1));
        nfound++;
    }
    if (lane == 0) {
        T.found[r] = (uint8_t)nfound;
        for (uint32_t i = nfound; i < kk; i++) st_cand(T.out, (size_t)r * kk + i, mars_hip_candidate{});
    }
}

// Work item (table, block of 128 ranges), one thread per range, for partials
// that carry the winner's exact t1 (the MFMA search): min over position
// chunks by (rms, key), then the winner's RawMoments and (qalfa, qbeta, rms)
// through the binary64 fit.
__global__ void k_merge_exact(const TableDesc *__restrict__ tabs, const Work2 *__restrict__ work,
                              Common cm)
{
    const Work2 wk = work[blockIdx.x];
    const TableG T = view(tabs[wk.x]);
    const uint32_t r = wk.y * blockDim.x + threadIdx.x;
    if (r >= T.nr) return;
    if (T.active && !T.active[r]) { // walk-only, not visited: its partials are stale
        T.score[r] = __longlong_as_double(0x7ff0000000000000ll);
        st_cand(T.out, r, mars_hip_candidate{});
        T.found[r] = 0;
        return;
    }
    const uint32_t nchunks = T.nchunks[0];
    double b = __longlong_as_double(0x7ff0000000000000ll);
    uint32_t k = 0xffffffffu;
    int32_t t1 = 0;
    if (T.P > 0)
        for (uint32_t ch = 0; ch < nchunks; ch++) {
            const Partial e = ld_partial(T.partial, (size_t)r * nchunks + ch);
            if (better(e.rms, e.key, b, k)) {
                b = e.rms;
                k = e.key;
                t1 = e.t1;
            }
        }
    T.score[r] = b;
    if (k == 0xffffffffu) {
        st_cand(T.out, r, mars_hip_candidate{});
        T.found[r] = 0;
        return;
    }
    write_winner(T, cm, r, k, t1);
}

// Work item (table, range), one 64-thread block, for partials without an
// exact t1 (the FP32 baseline): min over position chunks by (rms, key), the
// winner's t1 by a direct window correlation, then its RawMoments and
// (qalfa, qbeta, rms) through the binary64 fit, so the table holds the
// reference's Candidate (encode.rs:234-244) field for field.
__global__ void k_merge(const TableDesc *__restrict__ tabs, const Work2 *__restrict__ work,
                        Common cm)
{
    __shared__ double s_rms[64];
    __shared__ uint32_t s_key[64];
    __shared__ int32_t s_t1[64];
    __shared__ long long s_acc;
    const Work2 wk = work[blockIdx.x];
    const TableG T = view(tabs[wk.x]);
    const uint32_t r = wk.y, t = threadIdx.x, nchunks = T.nchunks[1];
    if (T.active && !T.active[r]) { // walk-only, not visited (block-uniform)
        if (t == 0) {
            T.score[r] = __longlong_as_double(0x7ff0000000000000ll);
            st_cand(T.out, r, mars_hip_candidate{});
            T.found[r] = 0;
        }
        return;
    }
    double b = __longlong_as_double(0x7ff0000000000000ll);
    uint32_t k = 0xffffffffu;
    int32_t t1 = 0;
    if (T.P > 0)
        for (uint32_t ch = t; ch < nchunks; ch += 64) {
            const Partial e = ld_partial(T.partial, (size_t)r * nchunks + ch);
            if (better(e.rms, e.key, b, k)) {
                b = e.rms;
                k = e.key;
                t1 = e.t1;
            }
        }
    s_rms[t] = b;
    s_key[t] = k;
    s_t1[t] = t1;
    if (t == 0) s_acc = 0;
    __syncthreads();
    if (t == 0) {
        for (int l = 1; l < 64; l++)
            if (better(s_rms[l], s_key[l], b, k)) {
                b = s_rms[l];
                k = s_key[l];
                t1 = s_t1[l];
            }
        s_key[0] = k;
        s_t1[0] = t1;
    }
    if (t == 0) s_rms[0] = b;
    __syncthreads();
    k = s_key[0];
    if (t == 0) T.score[r] = s_rms[0];
    if (k == 0xffffffffu) {
        if (t == 0) {
            st_cand(T.out, r, mars_hip_candidate{});
            T.found[r] = 0;
        }
        return;
    }
    const uint32_t p = k >> 3, iso = k & 7, S = T.size;
    const uint32_t pi = p / T.npc, pj = p - pi * T.npc;
    const uint32_t dr = (pi * T.shift) >> 1, dc = (pj * T.shift) >> 1;
    {
        const uint32_t row = T.rows[r], col = T.cols[r];
        long long part = 0;
        for (uint32_t e = t; e < S * S; e += 64) {
            uint32_t u = e / S, v = e - u * S, i, j;
            iso_map(iso, u, v, S, &i, &j);
            part += (long long)T.px[(size_t)(row + i) * T.w + col + j] *
                    (long long)T.d[(size_t)(dr + u) * T.wc + dc + v];
        }
        atomicAdd((unsigned long long *)&s_acc, (unsigned long long)part);
        __syncthreads();
    }
    if (t != 0) return;
    write_winner(T, cm, r, k, (int64_t)s_acc);
}

// Plain copy of n 16-byte words: the known-traffic reference for the HBM
// counters (it reads and writes exactly 16*n bytes).
__global__ void k_copy_probe(const uint4 *__restrict__ src, uint4 *__restrict__ dst, size_t n)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

void launch_copy_probe(const void *src, void *dst, size_t bytes, hipStream_t st)
{
    hipLaunchKernelGGL(k_copy_probe, dim3(4096), dim3(256), 0, st, (const uint4 *)src,
                       (uint4 *)dst, bytes / 16);
}

__global__ void k_fit_selftest(const int64_t *m, uint32_t n, double max_alfa, uint32_t bits_alfa,
                               uint32_t bits_beta, uint32_t *qa, uint32_t *qb, double *rms)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int64_t *x = m + 6 * (size_t)i;
    double sum;
    fit_f64(x[0], x[1], x[2], x[3], x[4], x[5], max_alfa, bits_alfa, bits_beta, &qa[i], &qb[i],
            &rms[i], &sum);
}

// Set-up of a fixed-grid search whose buffers the caller did not allocate:
// one grid-stride pass writes every list, coordinate and table that
// prepare_batch would build on the host (same values, so the shared kernels
// behave exactly as for a prepared table of one plane and one size).
__global__ void k_grid_setup(GridSetup g)
{
    const uint32_t nsearch = g.ngroups * g.nchunks;
    const uint32_t nperm = g.iperm ? g.P : 0;
    const uint32_t total = max(max(max(g.n2, nsearch), max(g.nr, nperm)), max(g.nq, 1u));
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < total; i += stride) {
        if (i == 0) {
            *g.plane_dst = g.plane;
            *g.table_dst = g.table;
        }
        if (i < g.n2) g.w2[i] = make_uint2(0u, i);
        if (i < nsearch) g.search[i] = make_uint3(0u, i % g.nchunks, i / g.nchunks);
        if (g.prepass && i < g.ngroups) g.prepass[i] = make_uint3(0u, 0u, i);
        if (i < g.nr) {
            const uint32_t gi = i / g.grid_cols;
            g.rows[i] = gi * g.size;
            g.cols[i] = (i - gi * g.grid_cols) * g.size;
        }
        if (i < nperm) {
            // Positions ordered by domain row, then by column residue modulo
            // order_m, then by column: a residue class c starts after the
            // ceil((npc - c') / m) columns of every smaller residue c'.
            const uint32_t pi = i / g.npc, pj = i - pi * g.npc, m = g.order_m;
            const uint32_t c = pj % m;
            uint32_t off = 0;
            for (uint32_t c2 = 0; c2 < c; c2++) off += (g.npc - c2 + m - 1) / m;
            g.iperm[i] = pi * g.npc + off + pj / m;
        }
        if (i < g.nq) {
            const double a2 = (double)i / (double)(1u << g.bits_alfa) * g.max_alfa;
            g.yscale[i] = g.qb_max / ((1.0 + fabs(a2)) * 255.0);
        }
    }
}

__global__ void k_zero_bytes(uint4 *p, size_t n)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        p[i] = make_uint4(0u, 0u, 0u, 0u);
}

// ---------------------------------------------------------------- launchers

void launch_grid_setup(const GridSetup &g, hipStream_t st)
{
    hipLaunchKernelGGL(k_grid_setup, dim3(256), dim3(256), 0, st, g);
}

void launch_zero_bytes(void *p, size_t bytes, hipStream_t st)
{
    const size_t n = bytes / 16;
    if (n == 0) return;
    const size_t blocks = std::min<size_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(k_zero_bytes, dim3((uint32_t)blocks), dim3(256), 0, st, (uint4 *)p, n);
}

template <int S, int KCAP>
static void launch_mfma(const TableDesc *t, const Work3 *work, uint32_t n, const Common &c,
                        hipStream_t st)
{
    hipLaunchKernelGGL((k_search_mfma<S, KCAP>), dim3(n), dim3(256), 0, st, t, work, c);
}
template <int S>
static void launch_fp32(const TableDesc *t, const Work3 *work, uint32_t n, const Common &c,
                        hipStream_t st)
{
    hipLaunchKernelGGL(k_search_fp32<S>, dim3(n), dim3(256), 0, st, t, work, c);
}

int launch_search(int kernel, uint32_t size, const TableDesc *t, const Work3 *work,
                  uint32_t n, const Common &c, hipStream_t st)
{
    if (n == 0) return 0;
    if (kernel == 0 && c.topk > 8) {
        switch (size) {
        case 2: launch_mfma<2, 32>(t, work, n, c, st); return 0;
        case 4: launch_mfma<4, 32>(t, work, n, c, st); return 0;
        case 8: launch_mfma<8, 32>(t, work, n, c, st); return 0;
        case 16: launch_mfma<16, 32>(t, work, n, c, st); return 0;
        case 32: launch_mfma<32, 32>(t, work, n, c, st); return 0;
        }
    } else if (kernel == 0 && c.topk > 1) {
        switch (size) {
        case 2: launch_mfma<2, 8>(t, work, n, c, st); return 0;
        case 4: launch_mfma<4, 8>(t, work, n, c, st); return 0;
        case 8: launch_mfma<8, 8>(t, work, n, c, st); return 0;
        case 16: launch_mfma<16, 8>(t, work, n, c, st); return 0;
        case 32: launch_mfma<32, 8>(t, work, n, c, st); return 0;
        }
    } else if (kernel == 0) {
        switch (size) {
        case 2: launch_mfma<2, 0>(t, work, n, c, st); return 0;
        case 4: launch_mfma<4, 0>(t, work, n, c, st); return 0;
        case 8: launch_mfma<8, 0>(t, work, n, c, st); return 0;
        case 16: launch_mfma<16, 0>(t, work, n, c, st); return 0;
        case 32: launch_mfma<32, 0>(t, work, n, c, st); return 0;
        }
    } else if (c.topk <= 1) {
        switch (size) {
        case 2: launch_fp32<2>(t, work, n, c, st); return 0;
        case 4: launch_fp32<4>(t, work, n, c, st); return 0;
        case 8: launch_fp32<8>(t, work, n, c, st); return 0;
        case 16: launch_fp32<16>(t, work, n, c, st); return 0;
        case 32: launch_fp32<32>(t, work, n, c, st); return 0;
        }
    }
    return -1;
}

void launch_contract(const PlaneDesc *planes, const Work2 *work, uint32_t n,
                     unsigned long long *counters, uint32_t n_counters, hipStream_t st)
{
    if (n == 0) return;
    hipLaunchKernelGGL(k_contract, dim3(n), dim3(256), 0, st, planes, work, counters, n_counters);
}

void launch_sat(const PlaneDesc *planes, const Work2 *rows_work, uint32_t n_rows,
                const Work2 *cols_work, uint32_t n_cols, hipStream_t st)
{
    if (n_rows) hipLaunchKernelGGL(k_sat_rows, dim3(n_rows), dim3(256), 0, st, planes, rows_work);
    if (n_cols) hipLaunchKernelGGL(k_sat_cols, dim3(n_cols), dim3(256), 0, st, planes, cols_work);
}

void launch_possums(const TableDesc *t, const Work2 *work, uint32_t n, hipStream_t st)
{
    if (n) hipLaunchKernelGGL(k_possums, dim3(n), dim3(256), 0, st, t, work);
}

void launch_activate(const TableDesc *t, const Work2 *work, uint32_t n, double t_rms,
                     uint32_t min_size, Work3 *wq, uint32_t *wq_count, hipStream_t st)
{
    if (n)
        hipLaunchKernelGGL(k_activate, dim3(n), dim3(256), 0, st, t, work, t_rms, min_size, wq,
                           wq_count);
}

void launch_zero_u32(uint32_t *p, uint32_t n, hipStream_t st)
{
    if (n) hipLaunchKernelGGL(k_zero_u32, dim3((n + 255) / 256), dim3(256), 0, st, p, n);
}

void launch_seed(const TableDesc *t, const Work2 *work, uint32_t n, const Common &c,
                 hipStream_t st)
{
    if (n) hipLaunchKernelGGL(k_seed, dim3(n), dim3(256), 0, st, t, work, c);
}

void launch_ranges(const TableDesc *t, const Work2 *work, uint32_t n, bool fp32, hipStream_t st)
{
    if (n) hipLaunchKernelGGL(k_ranges, dim3(n), dim3(256), 0, st, t, work, fp32 ? 1 : 0);
}



void launch_merge_ranked(const TableDesc *t, const Work2 *work, uint32_t n, const Common &c,
                         hipStream_t st)
{
    if (n) hipLaunchKernelGGL(k_merge_ranked, dim3(n), dim3(64), 0, st, t, work, c);
}

void launch_merge(const TableDesc *t, const Work2 *work, uint32_t n, bool t1_exact,
                  const Common &c, hipStream_t st)
{
    if (n == 0) return;
    if (t1_exact)
        hipLaunchKernelGGL(k_merge_exact, dim3(n), dim3(128), 0, st, t, work, c);
    else
        hipLaunchKernelGGL(k_merge, dim3(n), dim3(64), 0, st, t, work, c);
}

void launch_fit_selftest(const int64_t *m, uint32_t n, double max_alfa, uint32_t bits_alfa,
                         uint32_t bits_beta, uint32_t *qa, uint32_t *qb, double *rms,
                         hipStream_t st)
{
    hipLaunchKernelGGL(k_fit_selftest, dim3((n + 255) / 256), dim3(256), 0, st, m, n, max_alfa,
                       bits_alfa, bits_beta, qa, qb, rms);
}

} // namespace mars
