// Shared declarations between the device code (kernels.hip) and the host
// side of the C ABI (api.hip).
//
// A step covers a batch of planes.  Each plane has a pool (contracted plane,
// split i8 planes, summed-area tables); each (plane, range size, stride)
// has a search table.  Kernels take arrays of plane or table descriptors in
// device memory plus a flat work list, so one launch covers every plane (and
// every table of a size) of the batch.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "mars_hip.h"

namespace mars {

// Most candidates a ranked table keeps per range (search_top_k's k).
constexpr uint32_t kMaxRanked = 32;
// Smallest range size whose MFMA search runs a threshold pre-pass (every
// sample-th tile) before the full pass; smaller sizes always search with
// sample 1.
constexpr uint32_t kPrepassMinSize = 32;

// Per-(range, position chunk) partial winner.
struct Partial {
    double rms;
    uint32_t key; // (position index * 8 + isometry); 0xffffffff = none
    int32_t t1;   // exact t1_x4 of the winner (MFMA kernel), 0 otherwise
};

// One plane's pool.
struct PlaneDesc {
    const uint8_t *px;
    int32_t *d;          // contracted plane (i32)
    float *df;           // contracted plane (f32, FP32 baseline)
    int8_t *lo, *hi;     // split planes D % 4, D / 4 - 128; 16 byte-shifted copies each
    // D as binary16 (exact: D <= 1020) in the same 16-copy layout, 2-byte
    // elements (element offset = the split planes' byte offset); null when no
    // table of the plane has size 4 or 8
    uint16_t *wf;
    int64_t *sat1, *sat2;
    uint32_t w, h, wc, hc, pitch, copy_stride;
};

// One (plane, range size, stride) search table.
struct TableDesc {
    // the plane
    const uint8_t *px;
    const int32_t *d;
    const float *df;
    const int8_t *lo, *hi;
    const uint16_t *wf;  // binary16 plane (sizes 4 and 8), as PlaneDesc::wf
    const int64_t *sat1, *sat2;
    uint32_t w, wc, pitch, copy_stride;
    // the table
    const uint32_t *rows, *cols;
    int8_t *rng8;        // [nr_pad][8][kmem] r - 128
    float *rngf;         // [nr_pad/4][K][32] r (FP32 baseline)
    int32_t *rstat;      // [nr_pad][2] t0, t2
    int32_t *rpos;       // [nr_pad] 3 * sum max(r - 128, 0)
    uint32_t *gthr;      // [nr_pad] pruning threshold (float bits, SSE units)
    // Search-order records, P_pad = 32 * ceil(P / 32) slots:
    // {window byte offset, S1, p, sqrt(n*S2 - S1^2) as f32 rounded down},
    // at sizes 4 and 8 {window byte offset, p, S1, S2}
    // (p = ~0 in padding slots).  Slot of position p: iperm[p], or p when
    // iperm is null.
    uint4 *posinfo;
    // Sizes 4 and 8 (null otherwise), per search slot, two records: the
    // position's factors of the stage-1 bound as sixteen binary16 values (the
    // B operand of one more MFMA K-step), {S1h 256/n, S1l/n, dh, dl, 2^11, 1,
    // 1, 1 | dh, 0, ...} with S1 = 256 S1h + S1l and dh + dl <= sqrt(Di)/n.
    uint4 *posbias;
    const uint32_t *iperm;
    int32_t *posx;       // [P][2] S1, S2
    double *posd;        // [P][2] Di = n*S2 - S1^2, 1/Di (binary64, stage 2)
    Partial *partial;    // [nr_pad][nchunks][topk]
    mars_hip_candidate *out; // [nr][topk]
    uint8_t *found;          // [nr] candidates found (0..topk)
    double *score;       // [nr] the search kernel's own score of the winner (FP32: binary32 rms)
    // Walk-only tables (both null otherwise): only the ranges a threshold
    // walk visits are searched.  active[r]: visited this step; parent[r]:
    // the range's parent in table parent_tab, -1 when the parent splits
    // without a search (visited whatever the results), -2 when no walk asks.
    uint8_t *active;         // [nr_pad]
    const int32_t *parent;   // [nr]
    uint32_t parent_tab;
    uint32_t size, shift, P, P_pad, npc, nr, nr_pad, kmem;
    // Position chunking of the search: [0] MFMA kernel (32-position
    // tiles), [1] FP32 baseline (64-position tiles).  partial holds
    // nr_pad x nchunks[kernel] entries.
    uint32_t nchunks[2], tiles_per_chunk[2];
    // Fisher classified search (all null when the table has none).
    uint16_t *fcls;          // [P] bucket * 8 + isometry of each domain position
    uint32_t *fbuck;         // [2][72] start, count of each bucket's candidate list
    uint32_t *flist;         // [P] p * 8 + isometry, grouped by bucket
    uint16_t *frcls;         // [nr] bucket * 8 + isometry of each range
    mars_hip_candidate *fout; // [nr] Fisher winners
    uint8_t *ffound;          // [nr]
    uint32_t *fevals;         // [nr] candidates fitted (the bucket's size)
};

// Step-wide constants.
struct Common {
    double max_alfa;
    uint32_t bits_alfa, bits_beta;
    // Stage-2 constants (valid when stage2 != 0: max_alfa is a power of two).
    int stage2;
    double alfa_c;        // 2^(bits_alfa+2) / max_alfa: x = alfa_c * Ni / Di
    double alfa_unit;     // max_alfa / 2^bits_alfa: alfa2 = qalfa * alfa_unit
    double qa_max, qb_max; // 2^bits_alfa - 1, 2^bits_beta - 1
    const double *yscale; // [2^bits_alfa] (2^bits_beta-1) / ((1 + alfa2) * 255)
    // Stage-1 contrast cap: cap_u >= the largest quantised alfa / 4 (alfa
    // applies to the domain's pixel average, D / 4), cap_k <= 1 / (2 cap_u),
    // at most 8 so the integer limits of sizes <= 8 stay inside int32.
    float cap_u, cap_k;
    unsigned long long *counters; // [16]: [log2 s] stage-1 survivors, [8 + log2 s] stage-3 fits
    uint32_t sample;              // MFMA search: 1, or the tile stride of a threshold pre-pass
    uint32_t topk;                // candidates kept per range: 1, or a ranked table's k (<= 32)
    // MFMA search over a list compacted on the device (walk-only batches):
    // the number of valid work items; null for a static list.
    const uint32_t *nwork;
};

// Work items: x = descriptor index, y/z = kernel-specific block coordinates.
typedef uint2 Work2;
typedef uint3 Work3;

void launch_contract(const PlaneDesc *planes, const Work2 *work, uint32_t nwork,
                     unsigned long long *counters, uint32_t n_counters, hipStream_t st);
void launch_sat(const PlaneDesc *planes, const Work2 *rows_work, uint32_t n_rows_work,
                const Work2 *cols_work, uint32_t n_cols_work, hipStream_t st);
void launch_possums(const TableDesc *t, const Work2 *work, uint32_t nwork, hipStream_t st);
void launch_ranges(const TableDesc *t, const Work2 *work, uint32_t nwork, bool fp32,
                   hipStream_t st);
// Walk-only tables, work item (table, block of 256 ranges): active[r] from
// the parent's merged result (split when its rms > t_rms and its size >
// min_size, the threshold walk's rule).
// Also appends a search work item (table, chunk, range group) for every
// chunk of every range group with a visited range to wq, counting in
// *wq_count (zeroed by launch_zero_u32 beforehand).
void launch_activate(const TableDesc *t, const Work2 *work, uint32_t nwork, double t_rms,
                     uint32_t min_size, Work3 *wq, uint32_t *wq_count, hipStream_t st);
void launch_zero_u32(uint32_t *p, uint32_t n, hipStream_t st);
// Work item (table, first of four ranges).
// Ranges per work item of the threshold seed (work item y covers ranges
// [y * this, ...)): 32 for the winner-only tables of sizes <= 8, whose
// waves seed 8 ranges each, else 4 (a wave per range).
inline __host__ __device__ uint32_t seed_ranges_per_block(uint32_t size, uint32_t topk)
{
    return size <= 8 && topk <= 1 ? 32u : 4u;
}
void launch_seed(const TableDesc *t, const Work2 *work, uint32_t nwork, const Common &c,
                 hipStream_t st);
int launch_search(int kernel, uint32_t size, const TableDesc *t, const Work3 *work,
                  uint32_t nwork, const Common &c, hipStream_t st);
void launch_merge(const TableDesc *t, const Work2 *work, uint32_t nwork, bool t1_exact,
                  const Common &c, hipStream_t st);
// Ranked tables (c.topk > 1, MFMA search only): work item (table, range).
void launch_merge_ranked(const TableDesc *t, const Work2 *work, uint32_t nwork, const Common &c,
                         hipStream_t st);
// Fisher classified search: classification of every domain position
// (dom_work: table, block of 256 positions) and range (rng_work: table,
// block of 4 ranges), the buckets (tab_work: table, 0), then every range's
// bucket scan (range_work: table, range).  Needs the pool, k_possums,
// and k_ranges of those tables.
void launch_fisher(const TableDesc *t, const Work2 *dom_work, uint32_t n_dom,
                   const Work2 *rng_work, uint32_t n_rng, const Work2 *tab_work, uint32_t n_tab,
                   const Work2 *range_work, uint32_t n_range, const Common &c, hipStream_t st);
// Fixed-grid search into caller-free memory (mars_hip_search_grid_noalloc):
// what prepare_batch builds on the host and copies over -- the descriptors,
// the work lists, the range coordinates, the search order and the quantiser
// scale table -- is written on the device by one kernel from these values.
struct GridSetup {
    PlaneDesc plane;
    TableDesc table;
    PlaneDesc *plane_dst;
    TableDesc *table_dst;
    Work2 *w2;        // [n2]: (0, i), the one list every Work2 launch reads a prefix of
    uint32_t n2;
    Work3 *search;    // [ngroups * nchunks]: (0, chunk, range group), group-major
    Work3 *prepass;   // [ngroups]: (0, 0, range group), or null
    uint32_t ngroups, nchunks;
    uint32_t *rows, *cols; // [nr]: the grid's ranges, row-major
    uint32_t grid_cols, nr, size;
    uint32_t *iperm;  // [P]: search slot of each position, or null (natural order)
    uint32_t P, npc, order_m;
    double *yscale;   // [nq]: (2^bits_beta - 1) / ((1 + alfa2) * 255) per alfa code
    uint32_t nq, bits_alfa;
    double max_alfa, qb_max;
};
void launch_grid_setup(const GridSetup &g, hipStream_t st);
// Zero `bytes` (a multiple of 16) at p with vector stores.
void launch_zero_bytes(void *p, size_t bytes, hipStream_t st);

void launch_copy_probe(const void *src, void *dst, size_t bytes, hipStream_t st);
void launch_fit_selftest(const int64_t *m, uint32_t n, double max_alfa, uint32_t bits_alfa,
                         uint32_t bits_beta, uint32_t *qa, uint32_t *qb, double *rms,
                         hipStream_t st);

} // namespace mars
