// Host side of the C ABI declared in include/mars_hip.h.
//
// One context owns one HIP stream (plus side streams for concurrent
// per-size searches) and the device copy of a batch of planes: each plane's
// pool (contracted plane, split i8 planes, summed-area tables) and one search
// table per (plane, range size, domain stride).  All device buffers of a
// prepared batch live in one arena.  prepare_batch() enumerates the range
// nodes a quadtree walk can request on every plane, uploads the planes, lays
// out per-plane and per-table descriptors plus flat work lists, runs the
// step and downloads the winners; search() is a lookup into that table.
//
// A step is: one launch each for the pool and range preparation of every
// plane and table, then per (size, stride) one search launch over all
// planes' tables and one merge launch, the per-size chains on side streams
// (forked after the preparation, joined at the end), the whole step
// replayable from one captured HIP graph.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <unordered_map>
#include <vector>

#include "kernels.h"
#include "mars_codec.h"
#include "mars_hip.h"

// Device memory of mars_hip_search_grid_noalloc: every buffer of its step
// beyond the caller's input plane and output records (pool, range operands,
// position records, partial winners, descriptors, work lists, counters).
// Each GPU the library runs on holds its own copy until the process ends.
// Host code never passes this name to a kernel or copy (on the host it is a
// stand-in, not a device address): mars_hip_open looks up each device's
// address once with hipGetSymbolAddress and the entry point uses that.
static __device__ __align__(256) unsigned char scratch[MARS_HIP_NOALLOC_SCRATCH_BYTES];

namespace {

// Device address of `scratch` per HIP device (null until looked up).
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned char *g_scratch_ptr[kMaxDevices] = {};

// Look up `scratch` on the current device (`device`), once per device.
bool bind_scratch(int device)
{
    if (device < 0 || device >= kMaxDevices) return false;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[device]) return true;
    unsigned char *ptr = nullptr;
    if (hipGetSymbolAddress((void **)&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) return false;
    g_scratch_ptr[device] = ptr;
    return true;
}

unsigned char *scratch_of(int device)
{
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    return g_scratch_ptr[device];
}

} // namespace

namespace {

using mars::Common;
using mars::Partial;
using mars::PlaneDesc;
using mars::TableDesc;
using mars::Work2;
using mars::Work3;

constexpr uint32_t kMaxSize = 32;
constexpr uint32_t kCounters = 16;
constexpr int kSideStreams = 4;
constexpr uint32_t kMaxPlanes = 65536;
constexpr size_t kAlign = 256;

// Search chunking shape per range size: position chunks per range group
// are chosen so a launch has about `target_blocks` blocks, with at least
// `min_tiles` position tiles per chunk.  Results do not depend on it.
struct Shape {
    uint32_t target_blocks = 2048; // >> 256 CUs: fill the chip per launch
    uint32_t min_tiles = 8;        // >= 2 tiles per wave
};
constexpr uint32_t kShapeTargets[] = {256, 512, 1024, 2048, 4096, 8192, 16384};
constexpr uint32_t kShapeMinTiles[] = {4, 8, 16};
constexpr uint32_t kMaxTarget = 16384, kMinTiles = 4; // bounds of the autotuner's space
// Walk-only batches: position chunks per range group of the MFMA search (its
// work list holds only the visited groups, so it can split them finer than
// a list of every group could afford).
constexpr uint32_t kWalkChunks = 8;
uint32_t walk_chunks(uint32_t ntiles) { return std::max<uint32_t>(1, std::min(ntiles, kWalkChunks)); }
// Sizes whose MFMA search first runs a threshold pre-pass over a sample of
// the position tiles: their stage-1 bound lets many candidates
// through until the per-range thresholds are close to final, so tightening
// them on a sample first cuts the exact fits of the full pass.  The pre-pass
// runs one block per range group (all of the group's sampled tiles), so the
// group starts cold once, not once per position chunk; the launch is small
// and overlaps the other sizes' searches in the step.  (Not for walk-only
// batches, whose few searched groups the pre-pass would delay more than it
// saves.)  Size 32 only: at 512 x 512 without it the s = 32 search has
// 2.37 M stage-1 survivors and 1.07 M exact fits (126 k and 19 k with it),
// while at s = 16 the seed's thresholds leave 438 k survivors and the
// search alone (0.115 ms) is quicker than pre-pass and search (0.141 ms),
// whose 128-block pre-pass also ran last in the step.  (kPrepassMinSize,
// 32, is in kernels.h: the search kernel fixes its sample stride to 1 below
// that size.)
// The pre-pass samples every s-th tile at size s: at 512 x 512 the sparser
// sample (against every 8th) leaves the s = 32 survivors at 126 k (from 83 k)
// but finishes the launch early enough for the full pass to overlap the
// other sizes' searches (step 0.675 -> 0.661 ms).
uint32_t prepass_stride(uint32_t s) { return s; }

// Sizes whose search correlates in binary16 (the kernel's Geo<S>::F16).
bool uses_f16(uint32_t s) { return s == 4 || s == 8; }
// Positions of one table: the kernels address its 16-byte search-order
// records (two per slot at sizes 4 and 8) with 32-bit byte offsets.
constexpr uint32_t kMaxPositions = (1u << 27) - 32;
// The binary16 plane: copies 0..7 of the split planes' layout (all that
// sizes 4 and 8 address), 2-byte elements.
size_t f16_plane_bytes(size_t copy_stride) { return 8 * copy_stride * sizeof(uint16_t); }
// Window offsets are 32-bit: byte offsets into the 16 split-plane copies and
// element offsets into the binary16 plane (byte offset twice that, below
// 16 copy strides).
bool copies_addressable(size_t copy_stride) { return 16 * copy_stride < ((size_t)1 << 32); }

struct Plane {
    uint32_t w = 0, h = 0, wc = 0, hc = 0, pitch = 0;
    size_t copy_stride = 0;
    std::vector<uint8_t> host; // the caller's pixels, for the host-side encoder
    uint8_t *d_px = nullptr;
    int32_t *d_d = nullptr;
    float *d_df = nullptr;
    int8_t *d_lo = nullptr, *d_hi = nullptr;
    uint16_t *d_wf = nullptr; // binary16 plane, when a table has size 4 or 8
    int64_t *d_sat1 = nullptr, *d_sat2 = nullptr;
};

struct Table {
    uint32_t plane = 0, size = 0, shift = 0;
    std::vector<uint32_t> rows, cols;
    std::unordered_map<uint64_t, uint32_t> index;
    uint32_t nr = 0, nr_pad = 0, P = 0, npr = 0, npc = 0, kmem = 0;
    uint32_t nch_cap[2] = {1, 1};                   // partial capacity per range
    uint32_t nchunks[2] = {1, 1}, tpc[2] = {1, 1}; // current chunking
    uint32_t *d_rows = nullptr, *d_cols = nullptr;
    int8_t *d_rng8 = nullptr;
    float *d_rngf = nullptr;
    int32_t *d_rstat = nullptr;
    int32_t *d_rpos = nullptr;
    uint32_t *d_gthr = nullptr;
    uint4 *d_posinfo = nullptr;
    uint4 *d_posbias = nullptr; // sizes 4 and 8
    uint32_t *d_iperm = nullptr;
    int32_t *d_posx = nullptr;
    double *d_posd = nullptr;
    uint32_t P_pad = 0;
    std::vector<uint32_t> iperm; // search slot of each position (empty: identity)
    Partial *d_partial = nullptr;
    mars_hip_candidate *d_out = nullptr; // [nr][topk]
    uint8_t *d_found = nullptr;          // [nr] candidates kept (0..topk)
    double *d_score = nullptr;
    std::vector<mars_hip_candidate> h_out; // [nr][topk]
    std::vector<uint8_t> h_found;
    uint32_t topk = 1;
    // Walk-only batches: the threshold walk's visits (see TableDesc)
    uint8_t *d_active = nullptr;
    int32_t *d_parent = nullptr;
    uint32_t parent_tab = 0;
    std::vector<int32_t> h_parent;
    std::vector<uint8_t> h_active;
    // Fisher classified search (method MARS_HIP_METHOD_FISHER only)
    bool fisher = false;
    uint16_t *d_fcls = nullptr, *d_frcls = nullptr;
    uint32_t *d_fbuck = nullptr, *d_flist = nullptr, *d_fevals = nullptr;
    mars_hip_candidate *d_fout = nullptr;
    uint8_t *d_ffound = nullptr;
    std::vector<mars_hip_candidate> h_fout;
    std::vector<uint8_t> h_ffound;
    std::vector<uint32_t> h_fevals;
};

// Slice of a work list in the device work buffer.
struct Slice {
    uint32_t off = 0, n = 0;
};

// All tables of one (size, stride): one search launch and one merge launch.
struct Group {
    uint32_t size = 0, shift = 0;
    std::vector<uint32_t> tables; // longest searches first
    Slice search[2], merge[2];    // per kernel
    Slice prepass;                // MFMA threshold pre-pass: (table, 0, range group)
    Slice merge_ranked;           // (table, range) items of the ranked merge
    Slice seed, activate;         // walk-only batches: (table, 4 ranges), (table, 256 ranges)
    // Walk-only batches: the MFMA search's work list, compacted on the device
    // to the chunks of the range groups the walk visits (capacity: every chunk).
    Work3 *d_wq = nullptr;
    uint32_t wq_cap = 0;
};

} // namespace

struct mars_hip_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipStream_t side[kSideStreams] = {}; // default priority, like `stream`
    hipEvent_t fork_ev = nullptr, join_ev[kSideStreams] = {};
    hipEvent_t prep_fork_ev = nullptr, prep_join_ev = nullptr; // range operands beside the pool
    mutable std::mutex err_mu;
    std::string err;
    int kernel = MARS_HIP_KERNEL_MFMA_I8;
    uint32_t method = MARS_HIP_METHOD_EXHAUSTIVE; // of the prepared batch
    uint32_t ranked = 1;   // candidates per range the next prepare keeps (search_top_k's k)
    uint32_t topk = 1;     // ... and the prepared batch keeps
    // Threshold of the walk the next Fisher prepare serves its exhaustive
    // tables to (0: full tables), and that of the prepared batch.
    double walk_t = 0.0, walk_prepared = 0.0;
    uint32_t *d_wq_count = nullptr; // per group: items in its compacted work list
    bool prepared = false;
    mars_hip_params params{};
    std::vector<Plane> planes;
    std::vector<Table> tables;
    std::vector<Group> groups;
    std::map<std::tuple<uint32_t, uint32_t, uint32_t>, uint32_t> table_of; // (plane, size, shift)
    Shape shape[6];                                                        // per log2(size) - 1
    void *arena = nullptr;
    size_t arena_bytes = 0;
    PlaneDesc *d_planes = nullptr;
    TableDesc *d_tabs = nullptr;
    Work2 *d_work2 = nullptr;
    Work3 *d_work3 = nullptr;
    size_t work2_cap = 0, work3_cap = 0;
    Slice wl_contract, wl_sat_rows, wl_sat_cols, wl_possums, wl_ranges, wl_seed;
    Slice wl_fdom, wl_frng, wl_ftab, wl_frange; // Fisher: positions, range blocks, tables, ranges
    unsigned long long *d_counters = nullptr;
    double *d_yscale = nullptr;
    Common common{};
    hipGraphExec_t graph_exec = nullptr;
    hipGraph_t graph = nullptr;
    int graph_kernel = -1;
};

namespace {

int fail(mars_hip_ctx *ctx, int code, const std::string &msg)
{
    if (ctx) {
        std::lock_guard<std::mutex> lk(ctx->err_mu);
        ctx->err = msg;
    }
    return code;
}

#define HIP_TRY(ctx, expr)                                                                     \
    do {                                                                                       \
        hipError_t e_ = (expr);                                                                \
        if (e_ != hipSuccess)                                                                  \
            return fail(ctx, MARS_HIP_E_DEVICE,                                               \
                        std::string(#expr) + ": " + hipGetErrorString(e_));                    \
    } while (0)

#define RC_TRY(expr)                                                                           \
    do {                                                                                       \
        int rc_ = (expr);                                                                      \
        if (rc_) return rc_;                                                                   \
    } while (0)

uint32_t ceil_log2(uint32_t q)
{
    if (q <= 1) return 0;
    return 32 - __builtin_clz(q - 1);
}

uint32_t cdiv(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

typedef std::map<uint32_t, std::vector<std::pair<uint32_t, uint32_t>>> NodeMap;

// The nodes a walk can search (encode.rs:1330-1362 walk, 1865-1919 walk_rd):
// forced nodes (larger than max_size or crossing the border) split without a
// search; a size-1 node never searches; every other node is searched and
// splits further while size > min_size.
void collect(uint32_t row, uint32_t col, uint32_t size, uint32_t w, uint32_t h,
             const mars_hip_params &p, NodeMap &out)
{
    if (row >= h || col >= w) return;
    bool forced = size > p.max_size || row + size > h || col + size > w;
    if (forced) {
        uint32_t half = size / 2;
        collect(row, col, half, w, h, p, out);
        collect(row + half, col, half, w, h, p, out);
        collect(row, col + half, half, w, h, p, out);
        collect(row + half, col + half, half, w, h, p, out);
        return;
    }
    if (size == 1) return;
    out[size].emplace_back(row, col);
    if (size > p.min_size) {
        uint32_t half = size / 2;
        collect(row, col, half, w, h, p, out);
        collect(row + half, col, half, w, h, p, out);
        collect(row, col + half, half, w, h, p, out);
        collect(row + half, col + half, half, w, h, p, out);
    }
}

void drop_graph(mars_hip_ctx *ctx)
{
    if (ctx->graph_exec) (void)hipGraphExecDestroy(ctx->graph_exec);
    if (ctx->graph) (void)hipGraphDestroy(ctx->graph);
    ctx->graph_exec = nullptr;
    ctx->graph = nullptr;
    ctx->graph_kernel = -1;
}

void free_batch(mars_hip_ctx *ctx)
{
    drop_graph(ctx);
    ctx->planes.clear();
    ctx->tables.clear();
    ctx->groups.clear();
    ctx->table_of.clear();
    (void)hipFree(ctx->arena);
    (void)hipFree(ctx->d_work2);
    (void)hipFree(ctx->d_work3);
    (void)hipFree(ctx->d_yscale);
    ctx->arena = nullptr;
    ctx->arena_bytes = 0;
    ctx->d_planes = nullptr;
    ctx->d_tabs = nullptr;
    ctx->d_work2 = nullptr;
    ctx->d_work3 = nullptr;
    ctx->work2_cap = ctx->work3_cap = 0;
    ctx->d_yscale = nullptr;
    ctx->prepared = false;
    ctx->walk_prepared = 0.0;
}

// Position chunks for a table of `ntiles` tiles in a launch that covers
// `ngroups` range groups in total.
void chunking(uint32_t ntiles, uint32_t ngroups, uint32_t target, uint32_t min_tiles,
              uint32_t *nchunks, uint32_t *tpc)
{
    if (ntiles == 0) {
        *nchunks = 1;
        *tpc = 1;
        return;
    }
    uint32_t want = std::max<uint32_t>(1, target / std::max<uint32_t>(1, ngroups));
    uint32_t maxc = std::max<uint32_t>(1, ntiles / min_tiles);
    uint32_t c = std::min(want, maxc);
    uint32_t t = cdiv(ntiles, c);
    *tpc = t;
    *nchunks = cdiv(ntiles, t);
}

// Tiles and range groups of a table for kernel k (MFMA: 32 positions x 8
// ranges; FP32: 64 positions x 4 ranges).
uint32_t tiles_of(const Table &t, int k) { return cdiv(t.P, k == 0 ? 32 : 64); }
uint32_t groups_of(const Table &t, int k) { return t.nr_pad / (k == 0 ? 8 : 4); }

// Bytes of one window-row load of the MFMA search at size s (0: s = 2,
// whose loads are single bytes).
uint32_t row_load_bytes(uint32_t s) { return s <= 2 ? 0 : s == 4 ? 4 : s == 8 ? 8 : 16; }

// Search order of a table's positions.  The MFMA search loads each lane's
// window row segments (wb = 4, 8 or 16 bytes at sizes 4, 8, >= 16) straight
// from the copy of the split planes in which they start wb-aligned; in
// natural order the 32 lanes of a tile take windows from several copies, so
// a load touches several times the cache lines it needs (at s = 8 the L1
// stalled the address path for half the launch).  A tile whose 32 lanes
// take positions of one domain row whose window origins lie wb bytes apart
// reads one contiguous run per half-wave instead.  Positions are therefore
// ordered by domain row, then by column residue modulo 2*wb/shift, then by
// column (when shift divides 2*wb; otherwise the natural order is kept).
// (s = 8 launch at 512 x 512: 0.201 -> 0.142 ms.)
void search_order(Table &t)
{
    t.iperm.clear();
    const uint32_t wb = row_load_bytes(t.size);
    if (wb == 0 || t.P == 0 || (2 * wb) % t.shift != 0) return;
    const uint32_t m = 2 * wb / t.shift; // column step whose window origins are wb bytes apart
    if (m == 1) return;              // already contiguous
    t.iperm.resize(t.P);
    uint32_t q = 0;
    for (uint32_t pi = 0; pi < t.npr; pi++)
        for (uint32_t c = 0; c < m; c++)
            for (uint32_t pj = c; pj < t.npc; pj += m) t.iperm[pi * t.npc + pj] = q++;
}

// Bump allocator over one device allocation.
struct ArenaPlan {
    struct Item {
        void **dst;
        size_t bytes;
    };
    std::vector<Item> items;
    size_t total = 0;
    template <class X> void add(X **dst, size_t bytes)
    {
        bytes = std::max<size_t>(bytes, 1);
        items.push_back({(void **)dst, bytes});
        total += (bytes + kAlign - 1) / kAlign * kAlign;
    }
    void bind(char *base)
    {
        size_t off = 0;
        for (const Item &it : items) {
            *it.dst = base + off;
            off += (it.bytes + kAlign - 1) / kAlign * kAlign;
        }
    }
};

// Chunking of every table from the current shapes, then descriptors and work
// lists built and uploaded.  Invalidates the captured graph.
int layout(mars_hip_ctx *ctx)
{
    drop_graph(ctx);
    // chunking per group: the launch covers all tables of the group
    for (Group &g : ctx->groups) {
        // Walk-only batches search few of their groups, each whole: a finer
        // split of the positions than the autotuned one keeps more CUs busy.
        const Shape walk_shape{8192, 8};
        const Shape &sh = ctx->walk_prepared > 0.0 ? walk_shape : ctx->shape[ceil_log2(g.size) - 1];
        for (int k = 0; k < 2; k++) {
            uint32_t ng = 0;
            for (uint32_t ti : g.tables) ng += groups_of(ctx->tables[ti], k);
            for (uint32_t ti : g.tables) {
                Table &t = ctx->tables[ti];
                chunking(tiles_of(t, k), ng, sh.target_blocks, sh.min_tiles, &t.nchunks[k],
                         &t.tpc[k]);
                if (ctx->walk_prepared > 0.0 && k == 0) { // fixed chunks, listed on the device
                    t.tpc[0] = std::max<uint32_t>(1, cdiv(tiles_of(t, 0), walk_chunks(tiles_of(t, 0))));
                    t.nchunks[0] = std::max<uint32_t>(1, cdiv(tiles_of(t, 0), t.tpc[0]));
                }
                if (t.nchunks[k] > t.nch_cap[k]) { // outside the allocated shape space
                    t.tpc[k] = cdiv(tiles_of(t, k), t.nch_cap[k]);
                    t.nchunks[k] = cdiv(tiles_of(t, k), t.tpc[k]);
                }
            }
        }
    }
    // descriptors
    std::vector<PlaneDesc> pd(ctx->planes.size());
    for (size_t i = 0; i < pd.size(); i++) {
        const Plane &p = ctx->planes[i];
        pd[i] = PlaneDesc{p.d_px, p.d_d, p.d_df,  p.d_lo,  p.d_hi,  p.d_wf, p.d_sat1,
                          p.d_sat2, p.w, p.h, p.wc, p.hc, p.pitch, (uint32_t)p.copy_stride};
    }
    std::vector<TableDesc> td(ctx->tables.size());
    for (size_t i = 0; i < td.size(); i++) {
        const Table &t = ctx->tables[i];
        const Plane &p = ctx->planes[t.plane];
        TableDesc &d = td[i];
        d.px = p.d_px;
        d.d = p.d_d;
        d.df = p.d_df;
        d.lo = p.d_lo;
        d.hi = p.d_hi;
        d.wf = p.d_wf;
        d.sat1 = p.d_sat1;
        d.sat2 = p.d_sat2;
        d.w = p.w;
        d.wc = p.wc;
        d.pitch = p.pitch;
        d.copy_stride = (uint32_t)p.copy_stride;
        d.rows = t.d_rows;
        d.cols = t.d_cols;
        d.rng8 = t.d_rng8;
        d.rngf = t.d_rngf;
        d.rstat = t.d_rstat;
        d.rpos = t.d_rpos;
        d.gthr = t.d_gthr;
        d.posinfo = t.d_posinfo;
        d.posbias = t.d_posbias;
        d.iperm = t.iperm.empty() ? nullptr : t.d_iperm;
        d.posx = t.d_posx;
        d.posd = t.d_posd;
        d.partial = t.d_partial;
        d.out = t.d_out;
        d.found = t.d_found;
        d.score = t.d_score;
        d.active = t.d_active;
        d.parent = t.d_parent;
        d.parent_tab = t.parent_tab;
        d.size = t.size;
        d.shift = t.shift;
        d.P = t.P;
        d.P_pad = t.P_pad;
        d.npc = t.npc;
        d.nr = t.nr;
        d.nr_pad = t.nr_pad;
        d.kmem = t.kmem;
        for (int k = 0; k < 2; k++) {
            d.nchunks[k] = t.nchunks[k];
            d.tiles_per_chunk[k] = t.tpc[k];
        }
        d.fcls = t.d_fcls;
        d.fbuck = t.d_fbuck;
        d.flist = t.d_flist;
        d.frcls = t.d_frcls;
        d.fout = t.d_fout;
        d.ffound = t.d_ffound;
        d.fevals = t.d_fevals;
    }
    // work lists
    std::vector<Work2> w2;
    std::vector<Work3> w3;
    auto begin2 = [&](Slice &s) { s.off = (uint32_t)w2.size(); };
    auto end2 = [&](Slice &s) { s.n = (uint32_t)w2.size() - s.off; };
    begin2(ctx->wl_contract);
    for (uint32_t i = 0; i < ctx->planes.size(); i++) {
        const Plane &p = ctx->planes[i];
        for (uint32_t b = 0; b < cdiv((uint64_t)p.wc * p.hc, 256); b++) w2.push_back(make_uint2(i, b));
    }
    if (w2.size() == ctx->wl_contract.off) w2.push_back(make_uint2(0, 0)); // counters reset
    end2(ctx->wl_contract);
    begin2(ctx->wl_sat_rows);
    for (uint32_t i = 0; i < ctx->planes.size(); i++)
        for (uint32_t r = 0; r <= ctx->planes[i].hc; r++) w2.push_back(make_uint2(i, r));
    end2(ctx->wl_sat_rows);
    begin2(ctx->wl_sat_cols);
    for (uint32_t i = 0; i < ctx->planes.size(); i++)
        for (uint32_t b = 0; b < cdiv(ctx->planes[i].wc + 1, 16); b++) w2.push_back(make_uint2(i, b));
    end2(ctx->wl_sat_cols);
    begin2(ctx->wl_possums);
    for (uint32_t i = 0; i < ctx->tables.size(); i++)
        for (uint32_t b = 0; b < cdiv(ctx->tables[i].P_pad, 256); b++) w2.push_back(make_uint2(i, b));
    end2(ctx->wl_possums);
    begin2(ctx->wl_ranges);
    for (uint32_t i = 0; i < ctx->tables.size(); i++) {
        const Table &t = ctx->tables[i];
        if (t.P == 0) continue;
        for (uint32_t b = 0; b < cdiv((uint64_t)t.nr_pad * 8 * t.kmem / 4, 256); b++)
            w2.push_back(make_uint2(i, b));
    }
    end2(ctx->wl_ranges);
    begin2(ctx->wl_seed);
    for (uint32_t i = 0; i < ctx->tables.size(); i++) {
        const Table &t = ctx->tables[i];
        if (t.P == 0) continue;
        const uint32_t per = mars::seed_ranges_per_block(t.size, ctx->topk);
        for (uint32_t b = 0; b < cdiv(t.nr, per); b++) w2.push_back(make_uint2(i, b));
    }
    end2(ctx->wl_seed);
    // Fisher (tables with an empty pool keep their zeroed results)
    begin2(ctx->wl_fdom);
    for (uint32_t i = 0; i < ctx->tables.size(); i++) {
        const Table &t = ctx->tables[i];
        if (!t.fisher || t.P == 0) continue;
        for (uint32_t b = 0; b < cdiv(t.P, 256); b++) w2.push_back(make_uint2(i, b));
    }
    end2(ctx->wl_fdom);
    begin2(ctx->wl_frng);
    for (uint32_t i = 0; i < ctx->tables.size(); i++) {
        const Table &t = ctx->tables[i];
        if (!t.fisher || t.P == 0) continue;
        for (uint32_t b = 0; b < cdiv(t.nr, 4); b++) w2.push_back(make_uint2(i, b));
    }
    end2(ctx->wl_frng);
    begin2(ctx->wl_ftab);
    for (uint32_t i = 0; i < ctx->tables.size(); i++) {
        const Table &t = ctx->tables[i];
        if (t.fisher && t.P > 0) w2.push_back(make_uint2(i, 0));
    }
    end2(ctx->wl_ftab);
    begin2(ctx->wl_frange);
    for (uint32_t i = 0; i < ctx->tables.size(); i++) {
        const Table &t = ctx->tables[i];
        if (!t.fisher || t.P == 0) continue;
        for (uint32_t r = 0; r < t.nr; r++) w2.push_back(make_uint2(i, r));
    }
    end2(ctx->wl_frange);
    for (Group &g : ctx->groups) {
        for (int k = 0; k < 2; k++) {
            g.search[k].off = (uint32_t)w3.size();
            for (uint32_t ti : g.tables) {
                const Table &t = ctx->tables[ti];
                // (walk-only batches list the MFMA search's items on the device)
                if (t.P == 0 || (k == 0 && ctx->walk_prepared > 0.0)) continue;
                for (uint32_t gr = 0; gr < groups_of(t, k); gr++)
                    for (uint32_t ch = 0; ch < t.nchunks[k]; ch++) w3.push_back(make_uint3(ti, ch, gr));
            }
            g.search[k].n = (uint32_t)w3.size() - g.search[k].off;
            if (k == 0) {
                g.prepass.off = (uint32_t)w3.size();
                if (g.size >= mars::kPrepassMinSize && ctx->walk_prepared <= 0.0)
                    for (uint32_t ti : g.tables) {
                        const Table &t = ctx->tables[ti];
                        if (t.P == 0) continue;
                        for (uint32_t gr = 0; gr < groups_of(t, 0); gr++) w3.push_back(make_uint3(ti, 0, gr));
                    }
                g.prepass.n = (uint32_t)w3.size() - g.prepass.off;
            }
            begin2(g.merge[k]);
            for (uint32_t ti : g.tables) {
                const Table &t = ctx->tables[ti];
                const uint32_t nb = k == 0 ? cdiv(t.nr, 128) : t.nr;
                for (uint32_t b = 0; b < nb; b++) w2.push_back(make_uint2(ti, b));
            }
            end2(g.merge[k]);
        }
        begin2(g.merge_ranked);
        if (ctx->topk > 1)
            for (uint32_t ti : g.tables)
                for (uint32_t r = 0; r < ctx->tables[ti].nr; r++) w2.push_back(make_uint2(ti, r));
        end2(g.merge_ranked);
        // walk-only batches seed and activate each group after its parent's merge
        const bool walk = ctx->walk_prepared > 0.0;
        begin2(g.seed);
        for (uint32_t ti : g.tables) {
            const Table &t = ctx->tables[ti];
            if (!walk || t.P == 0) continue;
            const uint32_t per = mars::seed_ranges_per_block(t.size, ctx->topk);
            for (uint32_t b = 0; b < cdiv(t.nr, per); b++) w2.push_back(make_uint2(ti, b));
        }
        end2(g.seed);
        begin2(g.activate);
        for (uint32_t ti : g.tables)
            for (uint32_t b = 0; walk && b < cdiv(ctx->tables[ti].nr_pad, 256); b++)
                w2.push_back(make_uint2(ti, b));
        end2(g.activate);
    }
    // every launch must stay below 2^32 threads
    if (w2.size() >= (1u << 24) || w3.size() >= (1u << 24))
        return fail(ctx, MARS_HIP_E_UNSUPPORTED, "batch too large for one step");
    if (w2.size() > ctx->work2_cap) {
        (void)hipFree(ctx->d_work2);
        ctx->d_work2 = nullptr;
        HIP_TRY(ctx, hipMalloc((void **)&ctx->d_work2, sizeof(Work2) * w2.size()));
        ctx->work2_cap = w2.size();
    }
    if (std::max<size_t>(w3.size(), 1) > ctx->work3_cap) {
        (void)hipFree(ctx->d_work3);
        ctx->d_work3 = nullptr;
        HIP_TRY(ctx, hipMalloc((void **)&ctx->d_work3, sizeof(Work3) * std::max<size_t>(w3.size(), 1)));
        ctx->work3_cap = std::max<size_t>(w3.size(), 1);
    }
    HIP_TRY(ctx, hipMemcpy(ctx->d_work2, w2.data(), sizeof(Work2) * w2.size(), hipMemcpyHostToDevice));
    if (!w3.empty())
        HIP_TRY(ctx, hipMemcpy(ctx->d_work3, w3.data(), sizeof(Work3) * w3.size(), hipMemcpyHostToDevice));
    HIP_TRY(ctx, hipMemcpy(ctx->d_planes, pd.data(), sizeof(PlaneDesc) * pd.size(), hipMemcpyHostToDevice));
    if (!td.empty())
        HIP_TRY(ctx, hipMemcpy(ctx->d_tabs, td.data(), sizeof(TableDesc) * td.size(), hipMemcpyHostToDevice));
    return MARS_HIP_OK;
}

// Pool and range preparation of every plane and table.
// The step's preparation.  The range operands and sums read only the
// pixels, so with `fork` they run on a side stream beside the domain pool
// (contraction, summed-area tables, position sums) and join before the seed,
// which needs both; without, everything runs in order on st.
int enqueue_prep(mars_hip_ctx *ctx, int kernel, hipStream_t st, bool fork,
                 hipEvent_t *pool_ev = nullptr)
{
    const Work2 *w2 = ctx->d_work2;
    hipStream_t rs = st;
    if (fork) {
        rs = ctx->side[kSideStreams - 1];
        HIP_TRY(ctx, hipEventRecord(ctx->prep_fork_ev, st));
        HIP_TRY(ctx, hipStreamWaitEvent(rs, ctx->prep_fork_ev, 0));
    }
    mars::launch_ranges(ctx->d_tabs, w2 + ctx->wl_ranges.off, ctx->wl_ranges.n,
                        kernel == MARS_HIP_KERNEL_FP32, rs);
    if (fork) HIP_TRY(ctx, hipEventRecord(ctx->prep_join_ev, rs));
    if (pool_ev) HIP_TRY(ctx, hipEventRecord(pool_ev[0], st));
    mars::launch_contract(ctx->d_planes, w2 + ctx->wl_contract.off, ctx->wl_contract.n,
                          ctx->d_counters, kCounters, st);
    mars::launch_sat(ctx->d_planes, w2 + ctx->wl_sat_rows.off, ctx->wl_sat_rows.n,
                     w2 + ctx->wl_sat_cols.off, ctx->wl_sat_cols.n, st);
    mars::launch_possums(ctx->d_tabs, w2 + ctx->wl_possums.off, ctx->wl_possums.n, st);
    if (pool_ev) HIP_TRY(ctx, hipEventRecord(pool_ev[1], st));
    if (fork) HIP_TRY(ctx, hipStreamWaitEvent(st, ctx->prep_join_ev, 0));
    if (kernel == MARS_HIP_KERNEL_MFMA_I8 && ctx->walk_prepared <= 0.0)
        mars::launch_seed(ctx->d_tabs, w2 + ctx->wl_seed.off, ctx->wl_seed.n, ctx->common, st);
    return MARS_HIP_OK;
}

int enqueue_search(mars_hip_ctx *ctx, const Group &g, int kernel, hipStream_t st)
{
    const int k = kernel == MARS_HIP_KERNEL_MFMA_I8 ? 0 : 1;
    Common c = ctx->common;
    c.sample = 1;
    if (k == 0 && g.prepass.n > 0) {
        Common pre = c;
        pre.sample = prepass_stride(g.size);
        if (mars::launch_search(kernel, g.size, ctx->d_tabs, ctx->d_work3 + g.prepass.off,
                                g.prepass.n, pre, st) != 0)
            return fail(ctx, MARS_HIP_E_UNSUPPORTED, "no search kernel for this size");
    }
    if (mars::launch_search(kernel, g.size, ctx->d_tabs, ctx->d_work3 + g.search[k].off,
                            g.search[k].n, c, st) != 0)
        return fail(ctx, MARS_HIP_E_UNSUPPORTED, "no search kernel for this size");
    return MARS_HIP_OK;
}

void enqueue_merge(mars_hip_ctx *ctx, const Group &g, int kernel, hipStream_t st)
{
    const int k = kernel == MARS_HIP_KERNEL_MFMA_I8 ? 0 : 1;
    if (ctx->topk > 1) {
        mars::launch_merge_ranked(ctx->d_tabs, ctx->d_work2 + g.merge_ranked.off, g.merge_ranked.n,
                                  ctx->common, st);
        return;
    }
    mars::launch_merge(ctx->d_tabs, ctx->d_work2 + g.merge[k].off, g.merge[k].n, k == 0,
                       ctx->common, st);
}

// The Fisher search of every table (after the preparation).
void enqueue_fisher(mars_hip_ctx *ctx, hipStream_t st)
{
    const Work2 *w2 = ctx->d_work2;
    mars::launch_fisher(ctx->d_tabs, w2 + ctx->wl_fdom.off, ctx->wl_fdom.n, w2 + ctx->wl_frng.off,
                        ctx->wl_frng.n, w2 + ctx->wl_ftab.off, ctx->wl_ftab.n,
                        w2 + ctx->wl_frange.off, ctx->wl_frange.n, ctx->common, st);
}

// Enqueue one whole step.  With ev (6 + n_groups events: before the pool,
// after the preparation, after each group's search, after the merges, after
// the Fisher search) the step runs serially on the context stream so each
// phase can be timed; without, each group's search and merge run on a side
// stream forked after the preparation and joined at the end, the Fisher
// search on the context stream meanwhile.
int enqueue_step(mars_hip_ctx *ctx, int kernel, hipEvent_t *ev)
{
    hipStream_t st = ctx->stream;
    if (ev) HIP_TRY(ctx, hipEventRecord(ev[0], st));
    RC_TRY(enqueue_prep(ctx, kernel, st, ev == nullptr, ev ? ev + 4 + ctx->groups.size() : nullptr));
    if (ctx->walk_prepared > 0.0) {
        // Walk-only batch: the sizes in turn, largest first, each group's
        // ranges activated from its parent group's merged winners, then
        // seeded, searched and merged.  That chain keeps few blocks busy, so
        // it runs on a side stream beside the Fisher search; with ev it runs
        // serially on the context stream, group i's interval covering all
        // four steps.
        if (ev) HIP_TRY(ctx, hipEventRecord(ev[1], st));
        hipStream_t ws = st;
        if (!ev) {
            HIP_TRY(ctx, hipEventRecord(ctx->fork_ev, st));
            ws = ctx->side[0];
            HIP_TRY(ctx, hipStreamWaitEvent(ws, ctx->fork_ev, 0));
        }
        const Work2 *w2 = ctx->d_work2;
        mars::launch_zero_u32(ctx->d_wq_count, (uint32_t)ctx->groups.size(), ws);
        for (size_t i = 0; i < ctx->groups.size(); i++) {
            const Group &g = ctx->groups[i];
            mars::launch_activate(ctx->d_tabs, w2 + g.activate.off, g.activate.n, ctx->walk_prepared,
                                  ctx->params.min_size, g.d_wq, ctx->d_wq_count + i, ws);
            if (kernel == MARS_HIP_KERNEL_MFMA_I8)
                mars::launch_seed(ctx->d_tabs, w2 + g.seed.off, g.seed.n, ctx->common, ws);
            if (kernel == MARS_HIP_KERNEL_MFMA_I8) {
                Common c = ctx->common;
                c.sample = 1;
                c.nwork = ctx->d_wq_count + i;
                if (mars::launch_search(kernel, g.size, ctx->d_tabs, g.d_wq, g.wq_cap, c, ws) != 0)
                    return fail(ctx, MARS_HIP_E_UNSUPPORTED, "no search kernel for this size");
            } else {
                RC_TRY(enqueue_search(ctx, g, kernel, ws));
            }
            enqueue_merge(ctx, g, kernel, ws);
            if (ev) HIP_TRY(ctx, hipEventRecord(ev[2 + i], ws));
        }
        if (ev) HIP_TRY(ctx, hipEventRecord(ev[2 + ctx->groups.size()], st));
        enqueue_fisher(ctx, st);
        if (ev) HIP_TRY(ctx, hipEventRecord(ev[3 + ctx->groups.size()], st));
        if (!ev) {
            HIP_TRY(ctx, hipEventRecord(ctx->join_ev[0], ws));
            HIP_TRY(ctx, hipStreamWaitEvent(st, ctx->join_ev[0], 0));
        }
        HIP_TRY(ctx, hipGetLastError());
        return MARS_HIP_OK;
    }
    if (!ev) {
        HIP_TRY(ctx, hipEventRecord(ctx->fork_ev, st));
        const int nside = std::min<int>(kSideStreams, (int)ctx->groups.size());
        for (int j = 0; j < nside; j++) HIP_TRY(ctx, hipStreamWaitEvent(ctx->side[j], ctx->fork_ev, 0));
        // Smallest size (the most work) first, so its blocks are dispatched
        // ahead of the shorter chains'.
        for (size_t i = ctx->groups.size(); i-- > 0;) {
            hipStream_t ss = ctx->side[i % kSideStreams];
            RC_TRY(enqueue_search(ctx, ctx->groups[i], kernel, ss));
            enqueue_merge(ctx, ctx->groups[i], kernel, ss);
        }
        enqueue_fisher(ctx, st);
        for (int j = 0; j < nside; j++) {
            HIP_TRY(ctx, hipEventRecord(ctx->join_ev[j], ctx->side[j]));
            HIP_TRY(ctx, hipStreamWaitEvent(st, ctx->join_ev[j], 0));
        }
        HIP_TRY(ctx, hipGetLastError());
        return MARS_HIP_OK;
    }
    HIP_TRY(ctx, hipEventRecord(ev[1], st));
    for (size_t i = 0; i < ctx->groups.size(); i++) {
        RC_TRY(enqueue_search(ctx, ctx->groups[i], kernel, st));
        HIP_TRY(ctx, hipEventRecord(ev[2 + i], st));
    }
    for (const Group &g : ctx->groups) enqueue_merge(ctx, g, kernel, st);
    HIP_TRY(ctx, hipEventRecord(ev[2 + ctx->groups.size()], st));
    enqueue_fisher(ctx, st);
    HIP_TRY(ctx, hipEventRecord(ev[3 + ctx->groups.size()], st));
    HIP_TRY(ctx, hipGetLastError());
    return MARS_HIP_OK;
}

int download(mars_hip_ctx *ctx)
{
    for (Table &t : ctx->tables) {
        t.h_out.resize((size_t)t.nr * t.topk);
        t.h_found.resize(t.nr);
        if (t.nr == 0) continue;
        HIP_TRY(ctx, hipMemcpyAsync(t.h_out.data(), t.d_out,
                                    sizeof(mars_hip_candidate) * t.nr * t.topk,
                                    hipMemcpyDeviceToHost, ctx->stream));
        HIP_TRY(ctx, hipMemcpyAsync(t.h_found.data(), t.d_found, t.nr, hipMemcpyDeviceToHost,
                                    ctx->stream));
        if (t.d_active) {
            t.h_active.resize(t.nr);
            HIP_TRY(ctx, hipMemcpyAsync(t.h_active.data(), t.d_active, t.nr, hipMemcpyDeviceToHost,
                                        ctx->stream));
        }
        if (!t.fisher) continue;
        t.h_fout.resize(t.nr);
        t.h_ffound.resize(t.nr);
        t.h_fevals.resize(t.nr);
        HIP_TRY(ctx, hipMemcpyAsync(t.h_fout.data(), t.d_fout, sizeof(mars_hip_candidate) * t.nr,
                                    hipMemcpyDeviceToHost, ctx->stream));
        HIP_TRY(ctx, hipMemcpyAsync(t.h_ffound.data(), t.d_ffound, t.nr, hipMemcpyDeviceToHost,
                                    ctx->stream));
        HIP_TRY(ctx, hipMemcpyAsync(t.h_fevals.data(), t.d_fevals, 4 * (size_t)t.nr,
                                    hipMemcpyDeviceToHost, ctx->stream));
    }
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return MARS_HIP_OK;
}

uint64_t table_evals(const Table &t) { return (uint64_t)t.P * 8; }

// One eager step with events: per-phase times (ms) and counters.
int timed_step(mars_hip_ctx *ctx, mars_hip_timing *tm, std::vector<double> *group_ms)
{
    const size_t ng = ctx->groups.size();
    std::vector<hipEvent_t> ev(6 + ng);
    for (auto &e : ev) HIP_TRY(ctx, hipEventCreate(&e));
    int rc = enqueue_step(ctx, ctx->kernel, ev.data());
    if (rc == 0) {
        hipError_t e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) rc = fail(ctx, MARS_HIP_E_DEVICE, hipGetErrorString(e));
    }
    float ms = 0;
    if (rc == 0 && tm) {
        *tm = mars_hip_timing{};
        (void)hipEventElapsedTime(&ms, ev[0], ev[1]);
        tm->ms_pool = ms;
        (void)hipEventElapsedTime(&ms, ev[4 + ng], ev[5 + ng]);
        tm->ms_pool_build = ms;
    }
    if (group_ms) group_ms->assign(ng, 0.0);
    for (size_t i = 0; rc == 0 && i < ng; i++) {
        (void)hipEventElapsedTime(&ms, ev[1 + i], ev[2 + i]);
        if (group_ms) (*group_ms)[i] = ms;
        if (tm) {
            tm->ms_search += ms;
            uint32_t lg = ceil_log2(ctx->groups[i].size);
            if (lg >= 1 && lg <= 6) tm->ms_by_size[lg - 1] += ms;
        }
    }
    if (rc == 0 && tm) {
        unsigned long long cnt[kCounters] = {};
        hipError_t e = hipMemcpy(cnt, ctx->d_counters, sizeof(cnt), hipMemcpyDeviceToHost);
        if (e != hipSuccess) rc = fail(ctx, MARS_HIP_E_DEVICE, hipGetErrorString(e));
        for (int i = 0; i < 6; i++) {
            tm->survivors_by_size[i] = cnt[i + 1];
            tm->exact_fits_by_size[i] = cnt[8 + i + 1];
            tm->survivors += cnt[i + 1];
        }
        for (const Table &t : ctx->tables) {
            // walk-only tables: the visited nodes (the same every step)
            uint64_t nr = t.nr;
            if (t.d_active) nr = std::count(t.h_active.begin(), t.h_active.end(), (uint8_t)1);
            tm->evals += nr * table_evals(t);
        }
        (void)hipEventElapsedTime(&ms, ev[2 + ng], ev[3 + ng]);
        tm->ms_fisher = ms;
        for (const Table &t : ctx->tables)
            for (uint32_t e : t.h_fevals) tm->fisher_evals += e;
    }
    for (auto &e : ev) (void)hipEventDestroy(e);
    return rc;
}

// Step constants of one set of search params (the yscale table, counters,
// sample and topk are the caller's to set).  Stage 2 is exact only when
// max_alfa is a power of two: then the alfa scaling of encode.rs:264 is exact
// in binary64.
void fill_common(const mars_hip_params &p, Common &c)
{
    int e = 0;
    const double mant = std::frexp(p.max_alfa, &e);
    c = Common{};
    c.max_alfa = p.max_alfa;
    c.bits_alfa = p.bits_alfa;
    c.bits_beta = p.bits_beta;
    c.stage2 = (mant == 0.5 && p.bits_alfa <= 8 && p.bits_beta <= 12) ? 1 : 0;
    c.alfa_c = std::ldexp(1.0, (int)p.bits_alfa + 2) / p.max_alfa;
    c.alfa_unit = p.max_alfa / std::ldexp(1.0, (int)p.bits_alfa);
    c.qa_max = (double)((1u << p.bits_alfa) - 1);
    c.qb_max = (double)((1u << p.bits_beta) - 1);
    // Largest alfa2 of encode.rs:266 (qalfa <= qa_max), over 4, rounded up;
    // 1 / (2 cap_u) rounded down and capped.
    const double u = c.qa_max * c.alfa_unit * 0.25 * (1.0 + 1e-12);
    c.cap_u = std::nextafter((float)u, INFINITY);
    const double k = u > 0.0 ? 1.0 / (2.0 * (double)c.cap_u) : 8.0;
    c.cap_k = std::nextafter((float)(k < 8.0 ? k : 8.0), 0.0f);
}

bool valid_params(const mars_hip_params *p)
{
    return p->bits_alfa >= 1 && p->bits_alfa <= 16 && p->bits_beta >= 1 && p->bits_beta <= 16 &&
           p->max_alfa > 0.0 && p->min_size >= 1 && p->max_size >= p->min_size;
}

int prepare_batch(mars_hip_ctx *ctx, const mars_hip_plane *in, uint32_t n_planes,
                  const mars_hip_params *p, const uint32_t *shifts, uint32_t n_shifts)
{
    for (uint32_t i = 0; i < n_planes; i++) {
        if (!in[i].px) return fail(ctx, MARS_HIP_E_ARG, "null plane");
        if (in[i].w == 0 || in[i].h == 0 || in[i].w > 16384 || in[i].h > 16384)
            return fail(ctx, MARS_HIP_E_ARG, "bad plane size");
    }
    if (!valid_params(p)) return fail(ctx, MARS_HIP_E_ARG, "bad params");
    if (p->method != MARS_HIP_METHOD_EXHAUSTIVE && p->method != MARS_HIP_METHOD_FISHER)
        return fail(ctx, MARS_HIP_E_ARG, "unknown retrieval method");
    for (uint32_t i = 0; i < n_shifts; i++)
        if (shifts[i] == 0) return fail(ctx, MARS_HIP_E_ARG, "shift must be > 0");
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    if (ctx->ranked > 1 && ctx->kernel != MARS_HIP_KERNEL_MFMA_I8)
        return fail(ctx, MARS_HIP_E_UNSUPPORTED, "ranked tables need the MFMA search kernel");
    free_batch(ctx);
    ctx->params = *p;
    ctx->method = p->method;
    ctx->topk = ctx->ranked;

    std::vector<uint32_t> sh(shifts, shifts + n_shifts);
    std::sort(sh.begin(), sh.end());
    sh.erase(std::unique(sh.begin(), sh.end()), sh.end());

    // host geometry of planes and tables
    ctx->planes.resize(n_planes);
    for (uint32_t pi = 0; pi < n_planes; pi++) {
        Plane &pl = ctx->planes[pi];
        pl.w = in[pi].w;
        pl.h = in[pi].h;
        pl.wc = pl.w / 2;
        pl.hc = pl.h / 2;
        pl.pitch = ((pl.wc + 15) / 16) * 16 + 16;
        pl.copy_stride = (size_t)pl.pitch * std::max<uint32_t>(pl.hc, 1);
        if (!copies_addressable(pl.copy_stride))
            return fail(ctx, MARS_HIP_E_UNSUPPORTED, "plane too large for 32-bit window offsets");
        pl.host.assign(in[pi].px, in[pi].px + (size_t)pl.w * pl.h);
        NodeMap nodes;
        collect(0, 0, 1u << ceil_log2(std::max(pl.w, pl.h)), pl.w, pl.h, *p, nodes);
        for (auto &kv : nodes)
            if (kv.first > kMaxSize)
                return fail(ctx, MARS_HIP_E_UNSUPPORTED, "range sizes above 32 are not supported");
        for (uint32_t shift : sh) {
            for (auto it = nodes.rbegin(); it != nodes.rend(); ++it) {
                Table t;
                t.plane = pi;
                t.size = it->first;
                t.shift = shift;
                for (auto &rc : it->second) {
                    t.index[((uint64_t)rc.first << 32) | rc.second] = (uint32_t)t.rows.size();
                    t.rows.push_back(rc.first);
                    t.cols.push_back(rc.second);
                }
                t.nr = (uint32_t)t.rows.size();
                t.nr_pad = (t.nr + 7) / 8 * 8;
                const uint32_t s = t.size, K = s * s;
                t.kmem = K < 16 ? 16 : K;
                if (pl.h >= 2 * s && pl.w >= 2 * s) {
                    t.npr = (pl.h - 2 * s) / shift + 1;
                    t.npc = (pl.w - 2 * s) / shift + 1;
                }
                t.P = t.npr * t.npc;
                if (t.P > kMaxPositions) return fail(ctx, MARS_HIP_E_UNSUPPORTED, "too many positions");
                t.P_pad = (t.P + 31) / 32 * 32;
                t.fisher = p->method == MARS_HIP_METHOD_FISHER;
                t.topk = ctx->topk;
                search_order(t);
                ctx->table_of[std::make_tuple(pi, t.size, shift)] = (uint32_t)ctx->tables.size();
                ctx->tables.push_back(std::move(t));
            }
        }
    }
    // groups (size descending, then stride), longest searches first
    {
        std::map<std::pair<uint32_t, uint32_t>, std::vector<uint32_t>> by;
        for (uint32_t i = 0; i < ctx->tables.size(); i++)
            by[{ctx->tables[i].size, ctx->tables[i].shift}].push_back(i);
        for (auto it = by.rbegin(); it != by.rend(); ++it) {
            Group g;
            g.size = it->first.first;
            g.shift = it->first.second;
            g.tables = it->second;
            std::stable_sort(g.tables.begin(), g.tables.end(), [&](uint32_t a, uint32_t b) {
                return ctx->tables[a].P > ctx->tables[b].P;
            });
            ctx->groups.push_back(std::move(g));
        }
    }
    // Walk-only tables (Fisher batches whose exhaustive tables only serve a
    // threshold walk, the RD warm-up): every range's parent, at the base
    // stride (the walk's); other strides are never asked.
    ctx->walk_prepared = 0.0;
    if (ctx->walk_t > 0.0 && p->method == MARS_HIP_METHOD_FISHER && ctx->topk == 1) {
        ctx->walk_prepared = ctx->walk_t;
        for (Table &t : ctx->tables) {
            t.h_parent.assign(t.nr, -2);
            if (t.shift != p->shift) continue;
            const uint32_t ps = 2 * t.size;
            auto pt = ctx->table_of.find(std::make_tuple(t.plane, ps, t.shift));
            const Table *u = ps <= p->max_size && pt != ctx->table_of.end() ? &ctx->tables[pt->second]
                                                                           : nullptr;
            t.parent_tab = u ? pt->second : 0;
            for (uint32_t r = 0; r < t.nr; r++) {
                // A parent absent from its size's table is one the walk splits
                // without a search (larger than max_size or across the border).
                int32_t par = -1;
                if (u) {
                    auto it = u->index.find(((uint64_t)(t.rows[r] & ~(ps - 1)) << 32) |
                                            (t.cols[r] & ~(ps - 1)));
                    if (it != u->index.end()) par = (int32_t)it->second;
                }
                t.h_parent[r] = par;
            }
        }
    }
    // partial capacity: the most chunks any shape of the autotuner's space gives
    for (Group &g : ctx->groups)
        for (int k = 0; k < 2; k++) {
            uint32_t ng = 0;
            for (uint32_t ti : g.tables) ng += groups_of(ctx->tables[ti], k);
            for (uint32_t ti : g.tables) {
                Table &t = ctx->tables[ti];
                uint32_t tpc;
                chunking(tiles_of(t, k), ng, kMaxTarget, kMinTiles, &t.nch_cap[k], &tpc);
                // the ranked merge gives each chunk list one lane of a wave
                if (ctx->topk > 1) t.nch_cap[k] = std::min<uint32_t>(t.nch_cap[k], 64);
                if (ctx->walk_prepared > 0.0 && k == 0)
                    t.nch_cap[0] = std::max(t.nch_cap[0], walk_chunks(tiles_of(t, 0)));
            }
            if (ctx->walk_prepared > 0.0 && k == 0) {
                g.wq_cap = 0;
                for (uint32_t ti : g.tables)
                    g.wq_cap += groups_of(ctx->tables[ti], 0) * walk_chunks(tiles_of(ctx->tables[ti], 0));
            }
        }

    // one arena for every buffer of the batch
    ArenaPlan plan;
    plan.add(&ctx->d_planes, sizeof(PlaneDesc) * std::max<size_t>(ctx->planes.size(), 1));
    plan.add(&ctx->d_tabs, sizeof(TableDesc) * std::max<size_t>(ctx->tables.size(), 1));
    std::vector<char> plane_f16(ctx->planes.size(), 0);
    for (const Table &t : ctx->tables)
        if (uses_f16(t.size)) plane_f16[t.plane] = 1;
    for (size_t pi = 0; pi < ctx->planes.size(); pi++) {
        Plane &pl = ctx->planes[pi];
        const size_t nc = (size_t)pl.wc * pl.hc, nsat = (size_t)(pl.wc + 1) * (pl.hc + 1);
        plan.add(&pl.d_px, (size_t)pl.w * pl.h);
        plan.add(&pl.d_d, nc * sizeof(int32_t));
        plan.add(&pl.d_df, nc * sizeof(float));
        plan.add(&pl.d_lo, 16 * pl.copy_stride);
        plan.add(&pl.d_hi, 16 * pl.copy_stride);
        pl.d_wf = nullptr;
        if (plane_f16[pi]) plan.add(&pl.d_wf, f16_plane_bytes(pl.copy_stride));
        plan.add(&pl.d_sat1, nsat * sizeof(int64_t));
        plan.add(&pl.d_sat2, nsat * sizeof(int64_t));
    }
    for (Table &t : ctx->tables) {
        const size_t K = (size_t)t.size * t.size;
        plan.add(&t.d_rows, 4 * (size_t)t.nr);
        plan.add(&t.d_cols, 4 * (size_t)t.nr);
        plan.add(&t.d_rng8, (size_t)t.nr_pad * 8 * t.kmem);
        plan.add(&t.d_rngf, (size_t)t.nr_pad / 4 * K * 32 * sizeof(float));
        plan.add(&t.d_rstat, sizeof(int32_t) * 2 * t.nr_pad);
        plan.add(&t.d_rpos, sizeof(int32_t) * t.nr_pad);
        plan.add(&t.d_gthr, sizeof(uint32_t) * t.nr_pad);
        plan.add(&t.d_posinfo, sizeof(uint4) * (size_t)t.P_pad);
        t.d_posbias = nullptr;
        if (uses_f16(t.size)) plan.add(&t.d_posbias, 2 * sizeof(uint4) * (size_t)t.P_pad);
        plan.add(&t.d_iperm, sizeof(uint32_t) * t.iperm.size());
        plan.add(&t.d_posx, 2 * sizeof(int32_t) * (size_t)t.P);
        plan.add(&t.d_posd, 2 * sizeof(double) * (size_t)t.P);
        plan.add(&t.d_partial, sizeof(Partial) * (size_t)t.nr_pad *
                                   std::max(t.nch_cap[0], t.nch_cap[1]) * t.topk);
        plan.add(&t.d_out, sizeof(mars_hip_candidate) * (size_t)t.nr * t.topk);
        plan.add(&t.d_found, (size_t)t.nr);
        plan.add(&t.d_score, sizeof(double) * (size_t)t.nr);
        if (ctx->walk_prepared > 0.0) {
            plan.add(&t.d_active, (size_t)t.nr_pad);
            plan.add(&t.d_parent, sizeof(int32_t) * (size_t)t.nr);
        }
        if (t.fisher) {
            plan.add(&t.d_fcls, sizeof(uint16_t) * (size_t)t.P);
            plan.add(&t.d_fbuck, sizeof(uint32_t) * 2 * 72);
            plan.add(&t.d_flist, sizeof(uint32_t) * (size_t)t.P);
            plan.add(&t.d_frcls, sizeof(uint16_t) * (size_t)t.nr);
            plan.add(&t.d_fout, sizeof(mars_hip_candidate) * (size_t)t.nr);
            plan.add(&t.d_ffound, (size_t)t.nr);
            plan.add(&t.d_fevals, sizeof(uint32_t) * (size_t)t.nr);
        }
    }
    if (ctx->walk_prepared > 0.0) {
        for (Group &g : ctx->groups) plan.add(&g.d_wq, sizeof(Work3) * std::max<uint32_t>(g.wq_cap, 1));
        plan.add(&ctx->d_wq_count, sizeof(uint32_t) * ctx->groups.size());
    }
    HIP_TRY(ctx, hipMalloc(&ctx->arena, plan.total));
    ctx->arena_bytes = plan.total;
    plan.bind((char *)ctx->arena);
    HIP_TRY(ctx, hipMemsetAsync(ctx->arena, 0, plan.total, ctx->stream));
    for (uint32_t pi = 0; pi < n_planes; pi++) {
        Plane &pl = ctx->planes[pi];
        HIP_TRY(ctx, hipMemcpyAsync(pl.d_px, in[pi].px, (size_t)pl.w * pl.h, hipMemcpyHostToDevice,
                                    ctx->stream));
    }
    for (Table &t : ctx->tables) {
        HIP_TRY(ctx, hipMemcpyAsync(t.d_rows, t.rows.data(), 4 * (size_t)t.nr, hipMemcpyHostToDevice, ctx->stream));
        HIP_TRY(ctx, hipMemcpyAsync(t.d_cols, t.cols.data(), 4 * (size_t)t.nr, hipMemcpyHostToDevice, ctx->stream));
        if (!t.iperm.empty())
            HIP_TRY(ctx, hipMemcpyAsync(t.d_iperm, t.iperm.data(), 4 * t.iperm.size(),
                                        hipMemcpyHostToDevice, ctx->stream));
        if (t.d_parent && t.nr)
            HIP_TRY(ctx, hipMemcpyAsync(t.d_parent, t.h_parent.data(), 4 * (size_t)t.nr,
                                        hipMemcpyHostToDevice, ctx->stream));
    }
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));

    {
        Common &c = ctx->common;
        fill_common(*p, c);
        const uint32_t nq = 1u << p->bits_alfa;
        std::vector<double> ys(nq);
        for (uint32_t q = 0; q < nq; q++) {
            double a2 = (double)q / (double)(1u << p->bits_alfa) * p->max_alfa;
            ys[q] = c.qb_max / ((1.0 + std::fabs(a2)) * 255.0);
        }
        HIP_TRY(ctx, hipMalloc((void **)&ctx->d_yscale, sizeof(double) * nq));
        HIP_TRY(ctx, hipMemcpy(ctx->d_yscale, ys.data(), sizeof(double) * nq, hipMemcpyHostToDevice));
        c.yscale = ctx->d_yscale;
        c.counters = ctx->d_counters;
        c.sample = 1;
        c.topk = ctx->topk;
    }
    RC_TRY(layout(ctx));
    RC_TRY(enqueue_step(ctx, ctx->kernel, nullptr));
    RC_TRY(download(ctx));
    ctx->prepared = true;
    return MARS_HIP_OK;
}

int lookup(const mars_hip_ctx *ctx, uint32_t plane, uint32_t method, uint32_t row, uint32_t col,
           uint32_t size, uint32_t shift, uint32_t k, mars_hip_candidate *out, uint32_t *n_out,
           uint64_t *evals)
{
    mars_hip_ctx *mctx = const_cast<mars_hip_ctx *>(ctx);
    if (!ctx || !out || !n_out || !evals) return fail(mctx, MARS_HIP_E_ARG, "null argument");
    if (k < 1) return fail(mctx, MARS_HIP_E_ARG, "k must be >= 1");
    if (!ctx->prepared) return fail(mctx, MARS_HIP_E_NOT_PREPARED, "prepare() first");
    // A ranked request beyond what the table keeps cannot be served; Fisher
    // ranks nothing (the trait's default search_candidates: the winner).
    if (method == MARS_HIP_METHOD_EXHAUSTIVE && k > ctx->topk)
        return fail(mctx, MARS_HIP_E_UNSUPPORTED,
                    "more ranked candidates than prepared (mars_hip_set_ranked)");
    if (method != MARS_HIP_METHOD_EXHAUSTIVE &&
        !(method == MARS_HIP_METHOD_FISHER && ctx->method == MARS_HIP_METHOD_FISHER))
        return fail(mctx, MARS_HIP_E_ARG, "method not prepared");
    auto ti = ctx->table_of.find(std::make_tuple(plane, size, shift));
    if (ti != ctx->table_of.end()) {
        const Table &t = ctx->tables[ti->second];
        auto it = t.index.find(((uint64_t)row << 32) | col);
        if (it != t.index.end()) {
            const uint32_t i = it->second;
            const bool fisher = method == MARS_HIP_METHOD_FISHER;
            if (fisher) {
                *evals = t.h_fevals[i];
                *n_out = t.h_ffound[i] ? 1 : 0;
                if (*n_out) out[0] = t.h_fout[i];
            } else {
                if (t.d_active && !t.h_active[i]) {
                    char buf[192];
                    snprintf(buf, sizeof buf,
                             "plane %u range (%u,%u) size %u shift %u: the prepared walk-only "
                             "tables hold only the nodes the threshold walk visits",
                             plane, row, col, size, shift);
                    return fail(mctx, MARS_HIP_E_NOT_IN_TABLE, buf);
                }
                *evals = table_evals(t);
                const uint32_t n = std::min<uint32_t>(k, t.h_found[i]);
                for (uint32_t j = 0; j < n; j++) out[j] = t.h_out[(size_t)i * t.topk + j];
                *n_out = n;
            }
            return MARS_HIP_OK;
        }
    }
    char buf[192];
    snprintf(buf, sizeof buf, "plane %u range (%u,%u) size %u shift %u was not prepared", plane,
             row, col, size, shift);
    return fail(mctx, MARS_HIP_E_NOT_IN_TABLE, buf);
}

struct PlaneRef {
    const mars_hip_ctx *ctx;
    uint32_t plane;
    uint32_t method;
};

int plane_search_cb(void *user, uint32_t row, uint32_t col, uint32_t size, uint32_t shift,
                    uint32_t max, mars_hip_candidate *out, uint32_t *n_out, uint64_t *evals)
{
    const PlaneRef *r = (const PlaneRef *)user;
    // Fisher keeps the trait's default search_candidates (the winner only).
    const uint32_t k = r->method == MARS_HIP_METHOD_FISHER ? 1 : max;
    return lookup(r->ctx, r->plane, r->method, row, col, size, shift, k, out, n_out, evals);
}

} // namespace

extern "C" {

int mars_hip_open(int device, mars_hip_ctx **out)
{
    if (!out) return MARS_HIP_E_ARG;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n)
        return MARS_HIP_E_DEVICE;
    mars_hip_ctx *ctx = new mars_hip_ctx();
    ctx->device = device;
    bool ok = hipSetDevice(device) == hipSuccess && hipStreamCreate(&ctx->stream) == hipSuccess &&
              hipMalloc((void **)&ctx->d_counters, sizeof(unsigned long long) * kCounters) ==
                  hipSuccess &&
              hipEventCreateWithFlags(&ctx->fork_ev, hipEventDisableTiming) == hipSuccess &&
              hipEventCreateWithFlags(&ctx->prep_fork_ev, hipEventDisableTiming) == hipSuccess &&
              hipEventCreateWithFlags(&ctx->prep_join_ev, hipEventDisableTiming) == hipSuccess;
    for (int j = 0; ok && j < kSideStreams; j++)
        ok = hipStreamCreate(&ctx->side[j]) == hipSuccess &&
             hipEventCreateWithFlags(&ctx->join_ev[j], hipEventDisableTiming) == hipSuccess;
    if (!ok) {
        mars_hip_close(ctx);
        return MARS_HIP_E_DEVICE;
    }
    ctx->err.reserve(512); // error messages then fit without a host allocation
    // The no-alloc search's scratch address on this device; a failed lookup
    // (status checked in bind_scratch) leaves it null and only that entry
    // point then refuses to run.
    (void)bind_scratch(device);
    *out = ctx;
    return MARS_HIP_OK;
}

void mars_hip_close(mars_hip_ctx *ctx)
{
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    for (int j = 0; j < kSideStreams; j++)
        if (ctx->side[j]) (void)hipStreamSynchronize(ctx->side[j]);
    free_batch(ctx);
    (void)hipFree(ctx->d_counters);
    for (int j = 0; j < kSideStreams; j++) {
        if (ctx->side[j]) (void)hipStreamDestroy(ctx->side[j]);
        if (ctx->join_ev[j]) (void)hipEventDestroy(ctx->join_ev[j]);
    }
    if (ctx->fork_ev) (void)hipEventDestroy(ctx->fork_ev);
    if (ctx->prep_fork_ev) (void)hipEventDestroy(ctx->prep_fork_ev);
    if (ctx->prep_join_ev) (void)hipEventDestroy(ctx->prep_join_ev);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

const char *mars_hip_last_error(const mars_hip_ctx *ctx)
{
    if (!ctx) return "null context";
    std::lock_guard<std::mutex> lk(ctx->err_mu);
    return ctx->err.c_str();
}

int mars_hip_set_kernel(mars_hip_ctx *ctx, int kernel)
{
    if (!ctx) return MARS_HIP_E_ARG;
    if (kernel != MARS_HIP_KERNEL_MFMA_I8 && kernel != MARS_HIP_KERNEL_FP32)
        return fail(ctx, MARS_HIP_E_ARG, "unknown kernel");
    ctx->kernel = kernel;
    return MARS_HIP_OK;
}

int mars_hip_set_ranked(mars_hip_ctx *ctx, uint32_t k)
{
    if (!ctx) return MARS_HIP_E_ARG;
    if (k < 1 || k > mars::kMaxRanked) return fail(ctx, MARS_HIP_E_ARG, "k must be 1..32");
    ctx->ranked = k;
    return MARS_HIP_OK;
}

int mars_hip_set_walk_tables(mars_hip_ctx *ctx, double t_rms)
{
    if (!ctx) return MARS_HIP_E_ARG;
    if (!(t_rms == t_rms)) return fail(ctx, MARS_HIP_E_ARG, "t_rms is NaN");
    ctx->walk_t = t_rms > 0.0 ? t_rms : 0.0;
    return MARS_HIP_OK;
}

int mars_hip_prepare(mars_hip_ctx *ctx, const uint8_t *px, uint32_t w, uint32_t h,
                     const mars_hip_params *p, const uint32_t *shifts, uint32_t n_shifts)
{
    if (!ctx || !px || !p || !shifts || n_shifts == 0) return fail(ctx, MARS_HIP_E_ARG, "null argument");
    const mars_hip_plane one{px, w, h};
    return prepare_batch(ctx, &one, 1, p, shifts, n_shifts);
}

int mars_hip_prepare_batch(mars_hip_ctx *ctx, const mars_hip_plane *planes, uint32_t n_planes,
                           const mars_hip_params *p, const uint32_t *shifts, uint32_t n_shifts)
{
    if (!ctx || !planes || !p || !shifts || n_shifts == 0 || n_planes == 0)
        return fail(ctx, MARS_HIP_E_ARG, "null argument");
    if (n_planes > kMaxPlanes) return fail(ctx, MARS_HIP_E_ARG, "too many planes");
    return prepare_batch(ctx, planes, n_planes, p, shifts, n_shifts);
}

int mars_hip_run(mars_hip_ctx *ctx, int iters, int use_graph, mars_hip_timing *timing)
{
    if (!ctx || iters < 1) return fail(ctx, MARS_HIP_E_ARG, "bad argument");
    if (!ctx->prepared) return fail(ctx, MARS_HIP_E_NOT_PREPARED, "prepare() first");
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    mars_hip_timing tm{};
    RC_TRY(timed_step(ctx, &tm, nullptr)); // per-phase breakdown from one eager step

    if (use_graph && (ctx->graph_exec == nullptr || ctx->graph_kernel != ctx->kernel)) {
        drop_graph(ctx);
        HIP_TRY(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
        int rc = enqueue_step(ctx, ctx->kernel, nullptr);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(ctx->stream, &g);
        if (rc || e != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            return rc ? rc : fail(ctx, MARS_HIP_E_DEVICE, "graph capture failed");
        }
        ctx->graph = g; // owned by ctx from here: drop_graph() frees it on any later error
        HIP_TRY(ctx, hipGraphInstantiate(&ctx->graph_exec, g, nullptr, nullptr, 0));
        ctx->graph_kernel = ctx->kernel;
    }
    // The two timing events are released on every return path.
    struct Events {
        hipEvent_t e[2] = {nullptr, nullptr};
        ~Events()
        {
            for (hipEvent_t x : e)
                if (x) (void)hipEventDestroy(x);
        }
    } ev;
    HIP_TRY(ctx, hipEventCreate(&ev.e[0]));
    HIP_TRY(ctx, hipEventCreate(&ev.e[1]));
    HIP_TRY(ctx, hipEventRecord(ev.e[0], ctx->stream));
    for (int it = 0; it < iters; it++) {
        if (use_graph) {
            HIP_TRY(ctx, hipGraphLaunch(ctx->graph_exec, ctx->stream));
        } else {
            RC_TRY(enqueue_step(ctx, ctx->kernel, nullptr));
        }
    }
    HIP_TRY(ctx, hipEventRecord(ev.e[1], ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    float ms = 0;
    HIP_TRY(ctx, hipEventElapsedTime(&ms, ev.e[0], ev.e[1]));
    tm.ms_step = ms / iters;
    RC_TRY(download(ctx));
    if (timing) *timing = tm;
    return MARS_HIP_OK;
}

int mars_hip_autotune(mars_hip_ctx *ctx, uint32_t reps, mars_hip_tuning *out)
{
    if (!ctx || reps < 1) return fail(ctx, MARS_HIP_E_ARG, "bad argument");
    if (!ctx->prepared) return fail(ctx, MARS_HIP_E_NOT_PREPARED, "prepare() first");
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    // Every candidate shape is applied to all sizes at once; each size keeps
    // the shape whose summed search time (its groups, median of reps) is
    // lowest.
    double best[6];
    Shape chosen[6];
    for (int i = 0; i < 6; i++) {
        best[i] = 1e30;
        chosen[i] = ctx->shape[i];
    }
    for (uint32_t target : kShapeTargets)
        for (uint32_t mt : kShapeMinTiles) {
            for (int i = 0; i < 6; i++) ctx->shape[i] = Shape{target, mt};
            RC_TRY(layout(ctx));
            std::vector<std::vector<double>> samples(6);
            for (uint32_t r = 0; r < reps; r++) {
                std::vector<double> gms;
                RC_TRY(timed_step(ctx, nullptr, &gms));
                double by[6] = {};
                for (size_t g = 0; g < gms.size(); g++) by[ceil_log2(ctx->groups[g].size) - 1] += gms[g];
                for (int i = 0; i < 6; i++) samples[i].push_back(by[i]);
            }
            for (int i = 0; i < 6; i++) {
                std::sort(samples[i].begin(), samples[i].end());
                const double med = samples[i][samples[i].size() / 2];
                if (med < best[i]) {
                    best[i] = med;
                    chosen[i] = Shape{target, mt};
                }
            }
        }
    for (int i = 0; i < 6; i++) ctx->shape[i] = chosen[i];
    RC_TRY(layout(ctx));
    RC_TRY(enqueue_step(ctx, ctx->kernel, nullptr));
    RC_TRY(download(ctx));
    if (out) {
        *out = mars_hip_tuning{};
        for (int i = 0; i < 6; i++) {
            out->target_blocks[i] = chosen[i].target_blocks;
            out->min_tiles[i] = chosen[i].min_tiles;
            out->ms[i] = best[i] < 1e29 ? best[i] : 0.0;
        }
    }
    return MARS_HIP_OK;
}

int mars_hip_search(const mars_hip_ctx *ctx, uint32_t row, uint32_t col, uint32_t size,
                    uint32_t shift, uint32_t k, mars_hip_candidate *out, uint32_t *n_out,
                    uint64_t *evals)
{
    return lookup(ctx, 0, ctx ? ctx->method : 0, row, col, size, shift, k, out, n_out, evals);
}

int mars_hip_search_plane(const mars_hip_ctx *ctx, uint32_t plane, uint32_t row, uint32_t col,
                          uint32_t size, uint32_t shift, uint32_t k, mars_hip_candidate *out,
                          uint32_t *n_out, uint64_t *evals)
{
    return lookup(ctx, plane, ctx ? ctx->method : 0, row, col, size, shift, k, out, n_out, evals);
}

int mars_hip_search_method(const mars_hip_ctx *ctx, uint32_t plane, uint32_t method, uint32_t row,
                           uint32_t col, uint32_t size, uint32_t shift, uint32_t k,
                           mars_hip_candidate *out, uint32_t *n_out, uint64_t *evals)
{
    return lookup(ctx, plane, method, row, col, size, shift, k, out, n_out, evals);
}

int mars_hip_table_count(const mars_hip_ctx *ctx, uint32_t *n_tables)
{
    if (!ctx || !n_tables) return MARS_HIP_E_ARG;
    *n_tables = (uint32_t)ctx->tables.size();
    return MARS_HIP_OK;
}

int mars_hip_table_info(const mars_hip_ctx *ctx, uint32_t i, uint32_t *size, uint32_t *shift,
                        uint32_t *n_ranges, uint32_t *n_positions)
{
    if (!ctx || i >= ctx->tables.size()) return MARS_HIP_E_ARG;
    const Table &t = ctx->tables[i];
    if (size) *size = t.size;
    if (shift) *shift = t.shift;
    if (n_ranges) *n_ranges = t.nr;
    if (n_positions) *n_positions = t.P;
    return MARS_HIP_OK;
}

int mars_hip_table_plane(const mars_hip_ctx *ctx, uint32_t i, uint32_t *plane)
{
    if (!ctx || !plane || i >= ctx->tables.size()) return MARS_HIP_E_ARG;
    *plane = ctx->tables[i].plane;
    return MARS_HIP_OK;
}

int mars_hip_table_read(const mars_hip_ctx *ctx, uint32_t i, uint32_t *rows, uint32_t *cols,
                        mars_hip_candidate *out, uint8_t *found)
{
    if (!ctx || i >= ctx->tables.size()) return MARS_HIP_E_ARG;
    if (!ctx->prepared) return MARS_HIP_E_NOT_PREPARED;
    const Table &t = ctx->tables[i];
    if (rows) memcpy(rows, t.rows.data(), 4 * t.nr);
    if (cols) memcpy(cols, t.cols.data(), 4 * t.nr);
    for (uint32_t r = 0; r < t.nr; r++) {
        if (out) out[r] = t.h_out[(size_t)r * t.topk];
        if (found) found[r] = t.h_found[r] ? 1 : 0;
    }
    return MARS_HIP_OK;
}

int mars_hip_table_read_ranked(const mars_hip_ctx *ctx, uint32_t i, uint32_t *k,
                               mars_hip_candidate *out, uint8_t *counts)
{
    if (!ctx || i >= ctx->tables.size()) return MARS_HIP_E_ARG;
    if (!ctx->prepared) return MARS_HIP_E_NOT_PREPARED;
    const Table &t = ctx->tables[i];
    if (k) *k = t.topk;
    if (out) memcpy(out, t.h_out.data(), sizeof(mars_hip_candidate) * t.nr * t.topk);
    if (counts) memcpy(counts, t.h_found.data(), t.nr);
    return MARS_HIP_OK;
}

int mars_hip_table_read_method(const mars_hip_ctx *ctx, uint32_t i, uint32_t method,
                               mars_hip_candidate *out, uint8_t *found, uint64_t *evals)
{
    if (!ctx || i >= ctx->tables.size()) return MARS_HIP_E_ARG;
    if (!ctx->prepared) return MARS_HIP_E_NOT_PREPARED;
    const Table &t = ctx->tables[i];
    if (method == MARS_HIP_METHOD_EXHAUSTIVE) {
        for (uint32_t r = 0; r < t.nr; r++) {
            if (out) out[r] = t.h_out[(size_t)r * t.topk];
            if (found) found[r] = t.h_found[r] ? 1 : 0;
        }
        if (evals)
            for (uint32_t r = 0; r < t.nr; r++) evals[r] = table_evals(t);
        return MARS_HIP_OK;
    }
    if (method != MARS_HIP_METHOD_FISHER || !t.fisher) return MARS_HIP_E_ARG;
    if (out) memcpy(out, t.h_fout.data(), sizeof(mars_hip_candidate) * t.nr);
    if (found) memcpy(found, t.h_ffound.data(), t.nr);
    if (evals)
        for (uint32_t r = 0; r < t.nr; r++) evals[r] = t.h_fevals[r];
    return MARS_HIP_OK;
}

int mars_hip_table_scores(mars_hip_ctx *ctx, uint32_t i, double *score)
{
    if (!ctx || !score || i >= ctx->tables.size()) return fail(ctx, MARS_HIP_E_ARG, "bad argument");
    if (!ctx->prepared) return fail(ctx, MARS_HIP_E_NOT_PREPARED, "prepare() first");
    const Table &t = ctx->tables[i];
    if (t.nr == 0) return MARS_HIP_OK;
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    HIP_TRY(ctx, hipMemcpy(score, t.d_score, sizeof(double) * t.nr, hipMemcpyDeviceToHost));
    return MARS_HIP_OK;
}

int mars_hip_fit_f64(mars_hip_ctx *ctx, const int64_t *moments, uint32_t n, double max_alfa,
                     uint32_t bits_alfa, uint32_t bits_beta, uint32_t *qalfa, uint32_t *qbeta,
                     double *rms)
{
    if (!ctx || !moments || !qalfa || !qbeta || !rms) return fail(ctx, MARS_HIP_E_ARG, "null argument");
    if (n == 0) return MARS_HIP_OK;
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    int64_t *dm = nullptr;
    uint32_t *dqa = nullptr, *dqb = nullptr;
    double *dr = nullptr;
    HIP_TRY(ctx, hipMalloc((void **)&dm, sizeof(int64_t) * 6 * n));
    HIP_TRY(ctx, hipMalloc((void **)&dqa, 4 * n));
    HIP_TRY(ctx, hipMalloc((void **)&dqb, 4 * n));
    HIP_TRY(ctx, hipMalloc((void **)&dr, 8 * (size_t)n));
    HIP_TRY(ctx, hipMemcpy(dm, moments, sizeof(int64_t) * 6 * n, hipMemcpyHostToDevice));
    mars::launch_fit_selftest(dm, n, max_alfa, bits_alfa, bits_beta, dqa, dqb, dr, ctx->stream);
    HIP_TRY(ctx, hipGetLastError());
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    HIP_TRY(ctx, hipMemcpy(qalfa, dqa, 4 * n, hipMemcpyDeviceToHost));
    HIP_TRY(ctx, hipMemcpy(qbeta, dqb, 4 * n, hipMemcpyDeviceToHost));
    HIP_TRY(ctx, hipMemcpy(rms, dr, 8 * (size_t)n, hipMemcpyDeviceToHost));
    (void)hipFree(dm);
    (void)hipFree(dqa);
    (void)hipFree(dqb);
    (void)hipFree(dr);
    return MARS_HIP_OK;
}

int mars_hip_search_grid_noalloc(mars_hip_ctx *ctx, const uint8_t *d_px, uint32_t w, uint32_t h,
                                 const mars_hip_params *p, mars_hip_candidate *d_out,
                                 uint8_t *d_found, void *stream)
{
    if (!ctx || !d_px || !p || !d_out || !d_found) return fail(ctx, MARS_HIP_E_ARG, "null argument");
    if (w == 0 || h == 0 || w > 16384 || h > 16384) return fail(ctx, MARS_HIP_E_ARG, "bad plane size");
    if (!valid_params(p) || p->shift == 0) return fail(ctx, MARS_HIP_E_ARG, "bad params");
    if (p->method != MARS_HIP_METHOD_EXHAUSTIVE)
        return fail(ctx, MARS_HIP_E_ARG, "the no-alloc search is exhaustive only");
    const uint32_t s = p->max_size;
    if (p->min_size != s || s < 2 || s > kMaxSize || (s & (s - 1)))
        return fail(ctx, MARS_HIP_E_ARG,
                    "the no-alloc search takes one range size: min_size == max_size, a power of "
                    "two 2..32");
    unsigned char *const base = scratch_of(ctx->device);
    if (!base)
        return fail(ctx, MARS_HIP_E_DEVICE,
                    "hipGetSymbolAddress found no scratch address for this device");

    // Geometry, as prepare_batch lays out one table of one plane.
    const uint32_t shift = p->shift, gw = w / s, gh = h / s, nr = gw * gh;
    if (nr == 0) return MARS_HIP_OK;
    const uint32_t nr_pad = (nr + 7) / 8 * 8, K = s * s, kmem = K < 16 ? 16 : K;
    uint32_t npr = 0, npc = 0;
    if (h >= 2 * s && w >= 2 * s) {
        npr = (h - 2 * s) / shift + 1;
        npc = (w - 2 * s) / shift + 1;
    }
    const uint64_t P64 = (uint64_t)npr * npc;
    if (P64 > kMaxPositions) return fail(ctx, MARS_HIP_E_UNSUPPORTED, "too many positions");
    const uint32_t P = (uint32_t)P64, P_pad = (P + 31) / 32 * 32;
    const uint32_t wc = w / 2, hc = h / 2, pitch = ((wc + 15) / 16) * 16 + 16;
    const size_t copy_stride = (size_t)pitch * std::max<uint32_t>(hc, 1);
    if (!copies_addressable(copy_stride))
        return fail(ctx, MARS_HIP_E_UNSUPPORTED, "plane too large for 32-bit window offsets");
    const size_t nc = (size_t)wc * hc, nsat = (size_t)(wc + 1) * (hc + 1);
    // the search order of search_order()
    const uint32_t wb = row_load_bytes(s);
    const uint32_t order_m =
        (wb && P > 0 && (2 * wb) % shift == 0 && 2 * wb / shift > 1) ? 2 * wb / shift : 0;
    // chunking of the search with the context's shape for this size
    const uint32_t ngroups = nr_pad / 8;
    const Shape &sh = ctx->shape[ceil_log2(s) - 1];
    uint32_t nchunks = 1, tpc = 1;
    chunking(cdiv(P, 32), ngroups, sh.target_blocks, sh.min_tiles, &nchunks, &tpc);
    const bool prepass = s >= mars::kPrepassMinSize && P > 0;
    const uint32_t n_contract = std::max<uint32_t>(1, cdiv(nc, 256));
    const uint32_t n_sat_rows = hc + 1, n_sat_cols = cdiv(wc + 1, 16), n_possums = cdiv(P_pad, 256);
    const uint32_t n_ranges = cdiv((uint64_t)nr_pad * 8 * kmem / 4, 256);
    const uint32_t n_merge = cdiv(nr, 128);
    const uint32_t n_seed = cdiv(nr, mars::seed_ranges_per_block(s, 1));
    const uint32_t n2 = std::max({n_contract, n_sat_rows, n_sat_cols, n_possums, n_ranges, n_seed,
                                  n_merge});
    const uint64_t nsearch = (uint64_t)ngroups * nchunks;
    if (n2 >= (1u << 24) || nsearch >= (1u << 24))
        return fail(ctx, MARS_HIP_E_UNSUPPORTED, "plane too large for one step");
    const uint32_t nq = 1u << p->bits_alfa;

    // Carve the scratch (no allocation: fixed offsets from its base).
    size_t off = 0;
    auto take = [&](size_t bytes) {
        unsigned char *q = base + off;
        off += (std::max<size_t>(bytes, 1) + kAlign - 1) / kAlign * kAlign;
        return (void *)q;
    };
    mars::GridSetup g{};
    g.plane_dst = (PlaneDesc *)take(sizeof(PlaneDesc));
    g.table_dst = (TableDesc *)take(sizeof(TableDesc));
    g.w2 = (Work2 *)take(sizeof(Work2) * n2);
    g.search = (Work3 *)take(sizeof(Work3) * std::max<uint64_t>(nsearch, 1));
    g.prepass = prepass ? (Work3 *)take(sizeof(Work3) * ngroups) : nullptr;
    auto *counters = (unsigned long long *)take(sizeof(unsigned long long) * kCounters);
    g.yscale = (double *)take(sizeof(double) * nq);
    g.rows = (uint32_t *)take(4 * (size_t)nr);
    g.cols = (uint32_t *)take(4 * (size_t)nr);
    PlaneDesc &pd = g.plane;
    pd.px = d_px;
    pd.d = (int32_t *)take(nc * sizeof(int32_t));
    pd.df = (float *)take(nc * sizeof(float));
    pd.lo = (int8_t *)take(16 * copy_stride);
    pd.hi = (int8_t *)take(16 * copy_stride);
    pd.wf = uses_f16(s) ? (uint16_t *)take(f16_plane_bytes(copy_stride)) : nullptr;
    pd.sat1 = (int64_t *)take(nsat * sizeof(int64_t));
    pd.sat2 = (int64_t *)take(nsat * sizeof(int64_t));
    pd.w = w;
    pd.h = h;
    pd.wc = wc;
    pd.hc = hc;
    pd.pitch = pitch;
    pd.copy_stride = (uint32_t)copy_stride;
    TableDesc &td = g.table;
    td.px = pd.px;
    td.d = pd.d;
    td.df = pd.df;
    td.lo = pd.lo;
    td.hi = pd.hi;
    td.wf = pd.wf;
    td.sat1 = pd.sat1;
    td.sat2 = pd.sat2;
    td.w = w;
    td.wc = wc;
    td.pitch = pitch;
    td.copy_stride = (uint32_t)copy_stride;
    td.rows = g.rows;
    td.cols = g.cols;
    td.rng8 = (int8_t *)take((size_t)nr_pad * 8 * kmem);
    td.rngf = nullptr; // the FP32 baseline's operand: not used here
    td.rstat = (int32_t *)take(sizeof(int32_t) * 2 * nr_pad);
    td.rpos = (int32_t *)take(sizeof(int32_t) * nr_pad);
    td.gthr = (uint32_t *)take(sizeof(uint32_t) * nr_pad);
    td.posinfo = (uint4 *)take(sizeof(uint4) * (size_t)P_pad);
    td.posbias = uses_f16(s) ? (uint4 *)take(2 * sizeof(uint4) * (size_t)P_pad) : nullptr;
    g.iperm = order_m ? (uint32_t *)take(sizeof(uint32_t) * (size_t)P) : nullptr;
    td.iperm = g.iperm;
    td.posx = (int32_t *)take(2 * sizeof(int32_t) * (size_t)P);
    td.posd = (double *)take(2 * sizeof(double) * (size_t)P);
    td.partial = (Partial *)take(sizeof(Partial) * (size_t)nr_pad * nchunks);
    td.out = d_out;
    td.found = d_found;
    td.score = (double *)take(sizeof(double) * (size_t)nr);
    td.size = s;
    td.shift = shift;
    td.P = P;
    td.P_pad = P_pad;
    td.npc = npc;
    td.nr = nr;
    td.nr_pad = nr_pad;
    td.kmem = kmem;
    td.nchunks[0] = nchunks;
    td.tiles_per_chunk[0] = tpc;
    td.nchunks[1] = td.tiles_per_chunk[1] = 1;
    if (off > (size_t)MARS_HIP_NOALLOC_SCRATCH_BYTES) {
        char buf[160];
        snprintf(buf, sizeof buf,
                 "%ux%u plane, size %u, stride %u needs %zu bytes of scratch; the no-alloc "
                 "search has %zu",
                 w, h, s, shift, off, (size_t)MARS_HIP_NOALLOC_SCRATCH_BYTES);
        return fail(ctx, MARS_HIP_E_UNSUPPORTED, buf);
    }
    g.n2 = n2;
    g.ngroups = ngroups;
    g.nchunks = nchunks;
    g.grid_cols = gw;
    g.nr = nr;
    g.size = s;
    g.P = P;
    g.npc = npc;
    g.order_m = order_m;
    g.nq = nq;
    g.bits_alfa = p->bits_alfa;
    g.max_alfa = p->max_alfa;
    Common c;
    fill_common(*p, c);
    g.qb_max = c.qb_max;
    c.yscale = g.yscale;
    c.counters = counters;
    c.sample = 1;
    c.topk = 1;

    // The step: everything the shared kernels read starts from zero, as a
    // freshly prepared arena does (the previous call's contents are stale).
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    hipStream_t st = stream ? (hipStream_t)stream : ctx->stream;
    const PlaneDesc *planes = g.plane_dst;
    const TableDesc *tabs = g.table_dst;
    const Work2 *w2 = g.w2;
    mars::launch_zero_bytes(base, off, st);
    mars::launch_grid_setup(g, st);
    if (P > 0) {
        mars::launch_ranges(tabs, w2, n_ranges, false, st);
    }
    mars::launch_contract(planes, w2, n_contract, counters, kCounters, st);
    mars::launch_sat(planes, w2, n_sat_rows, w2, n_sat_cols, st);
    mars::launch_possums(tabs, w2, n_possums, st);
    if (P > 0) {
        mars::launch_seed(tabs, w2, n_seed, c, st);
        if (prepass) {
            Common pre = c;
            pre.sample = prepass_stride(s);
            if (mars::launch_search(MARS_HIP_KERNEL_MFMA_I8, s, tabs, g.prepass, ngroups, pre, st))
                return fail(ctx, MARS_HIP_E_UNSUPPORTED, "no search kernel for this size");
        }
        if (mars::launch_search(MARS_HIP_KERNEL_MFMA_I8, s, tabs, g.search, (uint32_t)nsearch, c, st))
            return fail(ctx, MARS_HIP_E_UNSUPPORTED, "no search kernel for this size");
    }
    mars::launch_merge(tabs, w2, n_merge, true, c, st);
    HIP_TRY(ctx, hipGetLastError());
    return MARS_HIP_OK;
}

int mars_hip_copy_probe(mars_hip_ctx *ctx, size_t bytes, double *ms)
{
    if (!ctx || bytes == 0 || bytes % 16) return fail(ctx, MARS_HIP_E_ARG, "bytes must be a positive multiple of 16");
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    void *src = nullptr, *dst = nullptr;
    HIP_TRY(ctx, hipMalloc(&src, bytes));
    HIP_TRY(ctx, hipMalloc(&dst, bytes));
    HIP_TRY(ctx, hipMemsetAsync(src, 1, bytes, ctx->stream));
    hipEvent_t e0, e1;
    HIP_TRY(ctx, hipEventCreate(&e0));
    HIP_TRY(ctx, hipEventCreate(&e1));
    HIP_TRY(ctx, hipEventRecord(e0, ctx->stream));
    mars::launch_copy_probe(src, dst, bytes, ctx->stream);
    HIP_TRY(ctx, hipEventRecord(e1, ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    float t = 0;
    HIP_TRY(ctx, hipEventElapsedTime(&t, e0, e1));
    if (ms) *ms = t;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(src);
    (void)hipFree(dst);
    return MARS_HIP_OK;
}

namespace {

mars_hip_params search_params(const mars_encode_params *p, uint32_t method)
{
    mars_hip_params hp{};
    hp.min_size = p->min_size;
    hp.max_size = p->max_size;
    hp.shift = p->shift;
    hp.bits_alfa = p->bits_alfa;
    hp.bits_beta = p->bits_beta;
    hp.method = method;
    hp.max_alfa = p->max_alfa;
    return hp;
}

// The walks request stride `shift`, and 2*shift for flat blocks when
// adaptive density is on (encode.rs:432-437, RD path only).
uint32_t walk_shifts(const mars_encode_params *p, const mars_encode_options *o, uint32_t *shifts)
{
    shifts[0] = p->shift;
    shifts[1] = 2 * p->shift;
    return (o->adaptive_density && p->has_lambda) ? 2 : 1;
}

} // namespace

int mars_hip_encode_plane(mars_hip_ctx *ctx, const uint8_t *px, uint32_t w, uint32_t h,
                          const mars_encode_params *p, const mars_encode_options *o,
                          uint8_t **out, size_t *out_len, mars_encode_stats *stats)
{
    return mars_hip_encode_plane_method(ctx, px, w, h, p, o, MARS_HIP_METHOD_EXHAUSTIVE, out,
                                        out_len, stats);
}

int mars_hip_encode_plane_method(mars_hip_ctx *ctx, const uint8_t *px, uint32_t w, uint32_t h,
                                 const mars_encode_params *p, const mars_encode_options *o,
                                 uint32_t method, uint8_t **out, size_t *out_len,
                                 mars_encode_stats *stats)
{
    if (!ctx || !p || !o) return fail(ctx, MARS_HIP_E_ARG, "null argument");
    const mars_hip_params hp = search_params(p, method);
    uint32_t shifts[2];
    const uint32_t ns = walk_shifts(p, o, shifts);
    // P5c: the exhaustive provider ranks rd_candidates per node.
    const uint32_t want = method == MARS_HIP_METHOD_EXHAUSTIVE && p->has_lambda
                              ? std::max<uint32_t>(1, std::min<uint32_t>(o->rd_candidates, mars::kMaxRanked))
                              : 1;
    const uint32_t keep = ctx->ranked;
    const double keep_walk = ctx->walk_t;
    RC_TRY(mars_hip_set_ranked(ctx, want));
    // With the Fisher retriever the exhaustive tables serve only the warm-up.
    ctx->walk_t = method == MARS_HIP_METHOD_FISHER && p->has_lambda ? MARS_WARMUP_T_RMS : 0.0;
    const int prc = mars_hip_prepare(ctx, px, w, h, &hp, shifts, ns);
    ctx->ranked = keep;
    ctx->walk_t = keep_walk;
    RC_TRY(prc);
    return mars_hip_encode_prepared(ctx, 0, p, o, out, out_len, stats);
}

int mars_hip_encode_prepared(mars_hip_ctx *ctx, uint32_t plane, const mars_encode_params *p,
                             const mars_encode_options *o, uint8_t **out, size_t *out_len,
                             mars_encode_stats *stats)
{
    if (!ctx || !p || !o) return fail(ctx, MARS_HIP_E_ARG, "null argument");
    if (!ctx->prepared) return fail(ctx, MARS_HIP_E_NOT_PREPARED, "prepare() first");
    if (plane >= ctx->planes.size()) return fail(ctx, MARS_HIP_E_ARG, "no such plane");
    const mars_hip_params &hp = ctx->params;
    if (p->min_size != hp.min_size || p->max_size != hp.max_size || p->bits_alfa != hp.bits_alfa ||
        p->bits_beta != hp.bits_beta || p->max_alfa != hp.max_alfa)
        return fail(ctx, MARS_HIP_E_ARG, "encode params differ from the prepared search params");
    const Plane &pl = ctx->planes[plane];
    // The final walk asks the prepared method's provider; the RD warm-up is
    // exhaustive whatever the method (encode.rs:1050, 1139-1169).
    PlaneRef search{ctx, plane, ctx->method};
    PlaneRef warmup{ctx, plane, MARS_HIP_METHOD_EXHAUSTIVE};
    int rc = mars_encode_plane(pl.host.data(), pl.w, pl.h, p, o, plane_search_cb, &search,
                               plane_search_cb, &warmup, out, out_len, stats);
    if (rc) {
        std::lock_guard<std::mutex> lk(ctx->err_mu);
        if (ctx->err.empty()) ctx->err = "encode failed";
    }
    return rc;
}

int mars_hip_encode_rgb(mars_hip_ctx *ctx, const uint8_t *rgb, uint32_t w, uint32_t h,
                        uint32_t subsampling, const mars_encode_params *y_params,
                        const mars_encode_params *chroma_params, const mars_encode_options *o,
                        uint32_t method, uint8_t **out, size_t *out_len, mars_color_stats *stats)
{
    if (!ctx || !rgb || !y_params || !chroma_params || !o || !out || !out_len)
        return fail(ctx, MARS_HIP_E_ARG, "null argument");
    if (w == 0 || h == 0) return fail(ctx, MARS_HIP_E_ARG, "empty image");
    if (subsampling != MARS_COLOR_RGB_444 && subsampling != MARS_COLOR_RGB_420)
        return fail(ctx, MARS_HIP_E_ARG, "subsampling must be 4:4:4 or 4:2:0");
    *out = nullptr;
    *out_len = 0;
    // color.rs:429-463: YCbCr, then chroma subsampled for 4:2:0.
    const size_t n = (size_t)w * h;
    std::vector<uint8_t> y(n), cb(n), cr(n);
    RC_TRY(mars_ycbcr(rgb, n, y.data(), cb.data(), cr.data()));
    uint32_t cw = w, ch = h;
    if (subsampling == MARS_COLOR_RGB_420) {
        cw = (w + 1) / 2;
        ch = (h + 1) / 2;
        std::vector<uint8_t> a((size_t)cw * ch), b((size_t)cw * ch);
        RC_TRY(mars_downsample_box(cb.data(), w, h, a.data()));
        RC_TRY(mars_downsample_box(cr.data(), w, h, b.data()));
        cb.swap(a);
        cr.swap(b);
    }
    // color.rs:465-476: Y with the luma params, Cb and Cr with the chroma
    // params (one batch: same params and strides).
    mars_color_stats st{};
    uint8_t *ys = nullptr, *cbs = nullptr, *crs = nullptr;
    size_t yl = 0, cbl = 0, crl = 0;
    int rc = mars_hip_encode_plane_method(ctx, y.data(), w, h, y_params, o, method, &ys, &yl, &st.y);
    if (rc == 0) {
        const mars_hip_params hp = search_params(chroma_params, method);
        uint32_t shifts[2];
        const uint32_t ns = walk_shifts(chroma_params, o, shifts);
        const mars_hip_plane planes[2] = {{cb.data(), cw, ch}, {cr.data(), cw, ch}};
        const uint32_t keep = ctx->ranked;
        const double keep_walk = ctx->walk_t;
        ctx->ranked = method == MARS_HIP_METHOD_EXHAUSTIVE && chroma_params->has_lambda
                          ? std::max<uint32_t>(1, std::min<uint32_t>(o->rd_candidates, mars::kMaxRanked))
                          : 1;
        ctx->walk_t = method == MARS_HIP_METHOD_FISHER && chroma_params->has_lambda
                          ? MARS_WARMUP_T_RMS
                          : 0.0;
        rc = mars_hip_prepare_batch(ctx, planes, 2, &hp, shifts, ns);
        ctx->ranked = keep;
        ctx->walk_t = keep_walk;
    }
    if (rc == 0) rc = mars_hip_encode_prepared(ctx, 0, chroma_params, o, &cbs, &cbl, &st.cb);
    if (rc == 0) rc = mars_hip_encode_prepared(ctx, 1, chroma_params, o, &crs, &crl, &st.cr);
    if (rc == 0) {
        const uint8_t *streams[3] = {ys, cbs, crs};
        const size_t lens[3] = {yl, cbl, crl};
        rc = mars_write_container((uint8_t)subsampling, streams, lens, 3, out, out_len);
    }
    mars_free(ys);
    mars_free(cbs);
    mars_free(crs);
    if (rc) return rc;
    st.y_bytes = yl;
    st.cb_bytes = cbl;
    st.cr_bytes = crl;
    if (stats) *stats = st;
    return MARS_HIP_OK;
}

} // extern "C"
