// Shared declarations between the device code (kernels.hip) and the host
// side of the C ABI (api.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "mars_hip.h"

namespace mars {

// Per-(range, position chunk) partial winner.
struct Partial {
    double rms;
    uint32_t key; // (position index * 8 + isometry); 0xffffffff = none
    int32_t t1;   // exact t1_x4 of the winner (MFMA kernel), 0 otherwise
};

struct SearchArgs {
    const int8_t *rng8;      // [nr_pad][8][kmem] r - 128
    const float *rngf;       // [nr_pad/4][K][32] r (FP32 baseline)
    const int32_t *rstat;    // [nr_pad][2] t0, t2
    uint32_t *gthr;          // [nr_pad] pruning threshold (float bits, SSE units)
    const uint4 *posinfo;    // [P] {dr | dc << 16, S1, S2, (float)(n*S2 - S1^2)}
    const double *posd;      // [P][2] Di = n*S2 - S1^2, 1/Di (binary64, stage 2)
    const int8_t *lo, *hi;   // split planes, 16 shifted copies each
    size_t copy_stride;      // bytes between copies
    uint32_t pitch;          // bytes per contracted row in a copy
    const float *dplanef;    // contracted plane as f32 (FP32 baseline)
    uint32_t wc;             // contracted width
    uint32_t P, npc, shift;  // positions, positions per row, stride
    uint32_t nr, ngroups, nchunks, tiles_per_chunk;
    double max_alfa;
    uint32_t bits_alfa, bits_beta;
    Partial *partial;        // [nr_pad][nchunks]
    unsigned long long *counters;
    uint32_t counter_slot;   // log2(size): counters[slot] stage-1 survivors, [8+slot] stage 3
    // Stage-2 constants (valid when stage2 != 0: max_alfa is a power of two).
    int stage2;
    double alfa_c;           // 2^(bits_alfa+2) / max_alfa: x = alfa_c * Ni / Di
    double alfa_unit;        // max_alfa / 2^bits_alfa: alfa2 = qalfa * alfa_unit
    double qa_max, qb_max;   // 2^bits_alfa - 1, 2^bits_beta - 1
    const double *yscale;    // [2^bits_alfa] (2^bits_beta-1) / ((1 + alfa2) * 255)
};

struct MergeArgs {
    const Partial *partial;
    int t1_exact;
    uint32_t nchunks, nr, P, npc, shift, size;
    const uint32_t *rows, *cols;
    const uint8_t *px;
    uint32_t w;
    const int32_t *d;
    uint32_t wc;
    const uint4 *posinfo;
    const int32_t *rstat;
    double max_alfa;
    uint32_t bits_alfa, bits_beta;
    mars_hip_candidate *out;
    uint8_t *found;
};

int launch_search(int kernel, uint32_t size, const SearchArgs &a, hipStream_t st);
void launch_contract(const uint8_t *px, uint32_t w, uint32_t wc, uint32_t hc, uint32_t pitch,
                     size_t copy_stride, int32_t *d, float *df, int8_t *lo, int8_t *hi,
                     unsigned long long *counters, uint32_t n_counters, hipStream_t st);
void launch_sat(const int32_t *d, uint32_t wc, uint32_t hc, int64_t *sat1, int64_t *sat2,
                hipStream_t st);
void launch_possums(const int64_t *sat1, const int64_t *sat2, uint32_t wc, uint32_t size,
                    uint32_t shift, uint32_t npc, uint32_t P, uint4 *posinfo, double *posd,
                    hipStream_t st);
void launch_seed(const int8_t *rng8, const int32_t *rstat, const int32_t *d, uint32_t wc,
                 uint32_t size, uint32_t kmem, uint32_t npc, uint32_t P, const uint4 *posinfo,
                 uint32_t nr, double max_alfa, uint32_t bits_alfa, uint32_t bits_beta,
                 uint32_t *gthr, hipStream_t st);
void launch_ranges(const uint8_t *px, uint32_t w, const uint32_t *rows, const uint32_t *cols,
                   uint32_t nr, uint32_t nr_pad, uint32_t size, uint32_t kmem, int8_t *rng8,
                   float *rngf, int32_t *rstat, uint32_t *gthr, hipStream_t st);
void launch_merge(const MergeArgs &a, hipStream_t st);
void launch_copy_probe(const void *src, void *dst, size_t bytes, hipStream_t st);
void launch_fit_selftest(const int64_t *m, uint32_t n, double max_alfa, uint32_t bits_alfa,
                         uint32_t bits_beta, uint32_t *qa, uint32_t *qb, double *rms,
                         hipStream_t st);

} // namespace mars
