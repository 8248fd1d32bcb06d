// Device helpers shared by the search kernels (kernels.hip) and the Fisher
// classified search (fisher.hip): address-space-1 views of the descriptors,
// the binary64 and binary32 fits, the candidate order and the isometry map.
// Included by device code only; compiled with -ffp-contract=off.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "kernels.h"

#pragma clang fp contract(off)

namespace mars {

// Global-memory pointers.  Pointers read out of a descriptor in device
// memory are generic to the compiler, which then emits flat_* accesses (they
// also count against the LDS wait counter); the device views below carry
// address space 1 so every descriptor access compiles to global_*.
#define GLB __attribute__((address_space(1)))

template <class X> __device__ __forceinline__ GLB X *glb(X *p) { return (GLB X *)p; }

// binary16 bits of an integer of magnitude <= 2048 (exact: an 11-bit
// significand holds it).
__device__ __forceinline__ uint16_t f16_bits(int32_t v)
{
    return __builtin_bit_cast(uint16_t, (_Float16)(float)v);
}

// binary16 bits of a float that binary16 holds exactly.
__device__ __forceinline__ uint16_t f16_bits_of(float x)
{
    return __builtin_bit_cast(uint16_t, (_Float16)x);
}

// binary16 bits of x >= 0 rounded toward zero (the largest binary16 <= x;
// 65504 above the range).
__device__ __forceinline__ uint16_t f16_rd_bits(float x)
{
    const _Float16 h = (_Float16)x;
    const uint16_t b = __builtin_bit_cast(uint16_t, h);
    return (float)h > x ? (uint16_t)(b - 1) : b;
}

// binary16 bits of x >= 0 rounded up (the smallest binary16 >= x, x <= 65504).
__device__ __forceinline__ uint16_t f16_ru_bits(float x)
{
    const _Float16 h = (_Float16)x;
    const uint16_t b = __builtin_bit_cast(uint16_t, h);
    return (float)h < x ? (uint16_t)(b + 1) : b;
}

// Class types (uint4, Partial, mars_hip_candidate) cannot be copied to or
// from address space 1 in C++, so the views hold them as plain vectors and
// words and these helpers move them.
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ uint4 ld_u4(const GLB u32x4 *p, size_t i)
{
    const u32x4 v = p[i];
    return make_uint4(v.x, v.y, v.z, v.w);
}
// A fresh read of an LDS word that other waves update (a relaxed atomic
// load stays a ds_read; a volatile one would become a flat load, which also
// waits for every global load in flight).
__device__ __forceinline__ uint32_t lds_load(const uint32_t *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// lane % 32, computed afresh at each call: the asm is volatile, so the
// compiler can neither hoist it out of a loop nor keep an older copy live.
__device__ __forceinline__ uint32_t lane_col()
{
    uint32_t l;
    asm volatile("v_mbcnt_lo_u32_b32 %0, -1, 0\n\t"
                 "v_mbcnt_hi_u32_b32 %0, -1, %0\n\t"
                 "v_and_b32 %0, 31, %0"
                 : "=v"(l));
    return l;
}

// The first three words of a record: a 12-byte load, so no destination
// register of an in-flight load is dead (the compiler would reuse it and
// wait for the load).
// Record i of an array of 16-byte records as the array's base plus a 32-bit
// byte offset, so a load takes the base in SGPRs and one VGPR offset (no
// 64-bit address arithmetic per lane).  Callers keep 16 i below 2^32: a
// table holds at most 2^27 - 32 positions, two records each at most.
__device__ __forceinline__ const GLB u32x4 *rec16(const GLB u32x4 *p, uint32_t i)
{
    return (const GLB u32x4 *)((const GLB char *)p + 16u * i);
}
__device__ __forceinline__ uint3 ld_u3(const GLB u32x4 *p, size_t i)
{
    typedef uint32_t u32x3 __attribute__((ext_vector_type(3)));
    const u32x3 v = *(const GLB u32x3 *)(p + i);
    return make_uint3(v.x, v.y, v.z);
}
__device__ __forceinline__ Partial ld_partial(const GLB Partial *p, size_t i)
{
    Partial e;
    e.rms = p[i].rms;
    e.key = p[i].key;
    e.t1 = p[i].t1;
    return e;
}
__device__ __forceinline__ void st_partial(GLB Partial *p, size_t i, double rms, uint32_t key,
                                           int32_t t1)
{
    p[i].rms = rms;
    p[i].key = key;
    p[i].t1 = t1;
}
__device__ __forceinline__ void st_cand(GLB mars_hip_candidate *p, size_t i,
                                        const mars_hip_candidate &c)
{
    static_assert(sizeof(mars_hip_candidate) % 8 == 0, "candidate is whole words");
    GLB uint64_t *d = (GLB uint64_t *)(p + i);
    const uint64_t *w = (const uint64_t *)&c;
#pragma unroll
    for (size_t k = 0; k < sizeof(mars_hip_candidate) / 8; k++) d[k] = w[k];
}

struct PlaneG {
    const GLB uint8_t *px;
    GLB int32_t *d;
    GLB float *df;
    GLB int8_t *lo, *hi;
    GLB uint16_t *wf;
    GLB int64_t *sat1, *sat2;
    uint32_t w, h, wc, hc, pitch, copy_stride;
};

__device__ __forceinline__ PlaneG view(const PlaneDesc &p)
{
    return PlaneG{glb(p.px),  glb(p.d),    glb(p.df), glb(p.lo), glb(p.hi), glb(p.wf),
                  glb(p.sat1), glb(p.sat2), p.w,       p.h,       p.wc,      p.hc,
                  p.pitch,     p.copy_stride};
}

struct TableG {
    const GLB uint8_t *px;
    const GLB int32_t *d;
    const GLB float *df;
    const GLB int8_t *lo, *hi;
    const GLB uint16_t *wf;
    const GLB int64_t *sat1, *sat2;
    uint32_t w, wc, pitch, copy_stride;
    const GLB uint32_t *rows, *cols;
    GLB int8_t *rng8;
    GLB float *rngf;
    GLB int32_t *rstat;
    GLB int32_t *rpos;
    GLB uint32_t *gthr;
    GLB u32x4 *posinfo;
    GLB u32x4 *posbias;
    GLB int32_t *posx;
    const GLB uint32_t *iperm;
    GLB double *posd;
    GLB Partial *partial;
    GLB mars_hip_candidate *out;
    GLB uint8_t *found;
    GLB double *score;
    GLB uint8_t *active;
    const GLB int32_t *parent;
    uint32_t parent_tab;
    uint32_t size, shift, P, P_pad, npc, nr, nr_pad, kmem;
    uint32_t nchunks[2], tiles_per_chunk[2];
    GLB uint16_t *fcls;
    GLB uint32_t *fbuck;
    GLB uint32_t *flist;
    GLB uint16_t *frcls;
    GLB mars_hip_candidate *fout;
    GLB uint8_t *ffound;
    GLB uint32_t *fevals;
};

__device__ __forceinline__ TableG view(const TableDesc &t)
{
    TableG g;
    g.px = glb(t.px);
    g.d = glb(t.d);
    g.df = glb(t.df);
    g.lo = glb(t.lo);
    g.hi = glb(t.hi);
    g.wf = glb(t.wf);
    g.sat1 = glb(t.sat1);
    g.sat2 = glb(t.sat2);
    g.w = t.w;
    g.wc = t.wc;
    g.pitch = t.pitch;
    g.copy_stride = t.copy_stride;
    g.rows = glb(t.rows);
    g.cols = glb(t.cols);
    g.rng8 = glb(t.rng8);
    g.rngf = glb(t.rngf);
    g.rstat = glb(t.rstat);
    g.rpos = glb(t.rpos);
    g.gthr = glb(t.gthr);
    g.posinfo = (GLB u32x4 *)t.posinfo;
    g.posbias = (GLB u32x4 *)t.posbias;
    g.posx = glb(t.posx);
    g.iperm = glb(t.iperm);
    g.posd = glb(t.posd);
    g.partial = glb(t.partial);
    g.out = glb(t.out);
    g.found = glb(t.found);
    g.score = glb(t.score);
    g.active = glb(t.active);
    g.parent = glb(t.parent);
    g.parent_tab = t.parent_tab;
    g.size = t.size;
    g.shift = t.shift;
    g.P = t.P;
    g.P_pad = t.P_pad;
    g.npc = t.npc;
    g.nr = t.nr;
    g.nr_pad = t.nr_pad;
    g.kmem = t.kmem;
    g.nchunks[0] = t.nchunks[0];
    g.nchunks[1] = t.nchunks[1];
    g.tiles_per_chunk[0] = t.tiles_per_chunk[0];
    g.tiles_per_chunk[1] = t.tiles_per_chunk[1];
    g.fcls = glb(t.fcls);
    g.fbuck = glb(t.fbuck);
    g.flist = glb(t.flist);
    g.frcls = glb(t.frcls);
    g.fout = glb(t.fout);
    g.ffound = glb(t.ffound);
    g.fevals = glb(t.fevals);
    return g;
}

// ---------------------------------------------------------------- fits

// encode.rs:446-448 quantise_f64
__device__ __forceinline__ uint32_t quantise_f64(double x, uint32_t max)
{
    double t = trunc(0.5 + x);
    t = t < 0.0 ? 0.0 : t;
    t = t > (double)max ? (double)max : t;
    return (uint32_t)t;
}

__device__ __forceinline__ uint32_t quantise_f32(float x, uint32_t max)
{
    float t = truncf(0.5f + x);
    t = t < 0.0f ? 0.0f : t;
    t = t > (float)max ? (float)max : t;
    return (uint32_t)t;
}

// encode.rs:247-289 fit_f64, same operation order.  Also returns `sum`
// (the SSE before /s0) for the pruning threshold.
__device__ __forceinline__ void fit_f64(int64_t ms0, int64_t ms1_x4, int64_t ms2_x16, int64_t mt0,
                                     int64_t mt1_x4, int64_t mt2, double max_alfa,
                                     uint32_t bits_alfa, uint32_t bits_beta, uint32_t *qa_out,
                                     uint32_t *qb_out, double *rms_out, double *sum_out)
{
    double s0 = (double)ms0;
    double s1 = (double)ms1_x4 / 4.0;
    double s2 = (double)ms2_x16 / 16.0;
    double t0 = (double)mt0;
    double t1 = (double)mt1_x4 / 4.0;
    double t2 = (double)mt2;

    double det = s0 * s2 - s1 * s1;
    double alfa = det == 0.0 ? 0.0 : (s0 * t1 - s1 * t0) / det;
    if (alfa < 0.0) alfa = 0.0;

    uint32_t max_qalfa = (1u << bits_alfa) - 1;
    uint32_t qalfa = quantise_f64(alfa / max_alfa * (double)(1u << bits_alfa), max_qalfa);
    // qalfa / 2^bits_alfa: a division by a power of two is exact, so the
    // scaling by ldexp gives the same binary64 bits without a divide.
    double alfa2 = ldexp((double)qalfa, -(int)bits_alfa) * max_alfa;

    double beta = (t0 - alfa2 * s1) / s0;
    if (alfa2 > 0.0) beta += alfa2 * 255.0;
    uint32_t max_qbeta = (1u << bits_beta) - 1;
    uint32_t qbeta =
        quantise_f64(beta / ((1.0 + fabs(alfa2)) * 255.0) * (double)max_qbeta, max_qbeta);
    double beta2 = (double)qbeta / (double)max_qbeta * ((1.0 + fabs(alfa2)) * 255.0);
    if (alfa2 > 0.0) beta2 -= alfa2 * 255.0;

    double sum = t2 - 2.0 * alfa2 * t1 - 2.0 * beta2 * t0 + alfa2 * alfa2 * s2 +
                 2.0 * alfa2 * beta2 * s1 + s0 * beta2 * beta2;
    double q = sum / s0;
    *rms_out = sqrt(q > 0.0 ? q : 0.0);
    *qa_out = qalfa;
    *qb_out = qbeta;
    *sum_out = sum;
}

// encode.rs:295-337 fit_f32 (the reference's own binary32 twin), used by the
// plain FP32 baseline kernel.
__device__ __forceinline__ void fit_f32(float s0, float s1_x4, float s2_x16, float t0,
                                        float t1_x4, float t2, float max_alfa,
                                        uint32_t bits_alfa, uint32_t bits_beta, uint32_t *qa_out,
                                        uint32_t *qb_out, float *rms_out)
{
    float s1 = s1_x4 / 4.0f;
    float s2 = s2_x16 / 16.0f;
    float t1 = t1_x4 / 4.0f;
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
    *qa_out = qalfa;
    *qb_out = qbeta;
}

// (rms, key) total order of the reference scan: strict `<` on rms keeps the
// first enumerated candidate, and enumeration order is the key order
// (dom_row, dom_col, isometry) -- encode.rs:676-685, 744-786.
__device__ __forceinline__ bool better(double rms, uint32_t key, double brms, uint32_t bkey)
{
    return rms < brms || (rms == brms && key < bkey);
}

__device__ __forceinline__ int32_t sel32(const int32_t (&v)[4], int i)
{
    return i == 0 ? v[0] : i == 1 ? v[1] : i == 2 ? v[2] : v[3];
}
__device__ __forceinline__ int64_t sel64(const int64_t (&v)[4], int i)
{
    return i == 0 ? v[0] : i == 1 ? v[1] : i == 2 ? v[2] : v[3];
}
__device__ __forceinline__ double sel_d(const double (&v)[4], int i)
{
    return i == 0 ? v[0] : i == 1 ? v[1] : i == 2 ? v[2] : v[3];
}

// isometry.rs:7-19: the source pixel (i, j) of cell (u, v) under isometry
// k = 0..7, i.e. (u, v), (s-1-v, u), (v, s-1-u), (s-1-u, s-1-v),
// (u, s-1-v), (s-1-u, v), (v, u), (s-1-v, s-1-u).  Written without
// branches -- a transpose (k in {1, 2, 6, 7}) then a flip of the row
// (k odd) and of the column (k in {2, 3, 4, 7}) -- so lanes of one wave
// with different k do not run the eight cases one after another.
__device__ __forceinline__ void iso_map(uint32_t k, uint32_t u, uint32_t v, uint32_t size,
                                        uint32_t *i, uint32_t *j)
{
    const bool tr = (0xC6u >> k) & 1u;
    const uint32_t a = tr ? v : u, b = tr ? u : v;
    *i = (k & 1u) ? size - 1 - a : a;
    *j = ((0x9Cu >> k) & 1u) ? size - 1 - b : b;
}

} // namespace mars
