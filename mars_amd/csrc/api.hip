// Host side of the C ABI declared in include/mars_hip.h.
//
// One context owns one HIP stream and the device copy of one plane: its pool
// (contracted plane, split i8 planes, per-size window sums) and one search
// table per (range size, domain stride).  prepare() enumerates the range
// nodes a quadtree walk can request, uploads the plane, runs the step and
// downloads the winners; search() is a lookup into that table.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "kernels.h"
#include "mars_codec.h"
#include "mars_hip.h"

namespace {

using mars::MergeArgs;
using mars::Partial;
using mars::SearchArgs;

constexpr uint32_t kMaxSize = 32;
constexpr uint32_t kTargetBlocks = 2048; // >> 256 CUs: fill the chip per launch
constexpr uint32_t kCounters = 16;

struct Table {
    uint32_t size = 0, shift = 0;
    std::vector<uint32_t> rows, cols;
    std::unordered_map<uint64_t, uint32_t> index;
    uint32_t nr = 0, nr_pad = 0, P = 0, npr = 0, npc = 0, kmem = 0;
    uint32_t nchunks_mfma = 1, tpc_mfma = 1, nchunks_fp32 = 1, tpc_fp32 = 1;
    uint32_t *d_rows = nullptr, *d_cols = nullptr;
    int8_t *d_rng8 = nullptr;
    float *d_rngf = nullptr;
    int32_t *d_rstat = nullptr;
    uint32_t *d_gthr = nullptr;
    uint4 *d_posinfo = nullptr;
    double *d_posd = nullptr;
    Partial *d_partial = nullptr;
    mars_hip_candidate *d_out = nullptr;
    uint8_t *d_found = nullptr;
    std::vector<mars_hip_candidate> h_out;
    std::vector<uint8_t> h_found;
};

} // namespace

constexpr int kSideStreams = 4;

struct mars_hip_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    // Per-size search chains run on side streams forked from `stream` after
    // the pool build (all at the default priority).
    hipStream_t side[kSideStreams] = {};
    hipEvent_t fork_ev = nullptr, join_ev[kSideStreams] = {};
    mutable std::mutex err_mu;
    std::string err;
    int kernel = MARS_HIP_KERNEL_MFMA_I8;
    bool prepared = false;
    uint32_t w = 0, h = 0, wc = 0, hc = 0, pitch = 0;
    size_t copy_stride = 0;
    mars_hip_params params{};
    uint8_t *d_px = nullptr;
    int32_t *d_d = nullptr;
    float *d_df = nullptr;
    int8_t *d_lo = nullptr, *d_hi = nullptr;
    int64_t *d_sat1 = nullptr, *d_sat2 = nullptr;
    unsigned long long *d_counters = nullptr;
    double *d_yscale = nullptr;
    int stage2 = 0;
    double alfa_c = 0, alfa_unit = 0;
    std::vector<Table> tables;
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

uint32_t ceil_log2(uint32_t q)
{
    if (q <= 1) return 0;
    return 32 - __builtin_clz(q - 1);
}

// The nodes a walk can search (encode.rs:1330-1362 walk, 1865-1919 walk_rd):
// forced nodes (larger than max_size or crossing the border) split without a
// search; a size-1 node never searches; every other node is searched and
// splits further while size > min_size.
void collect(uint32_t row, uint32_t col, uint32_t size, uint32_t w, uint32_t h,
             const mars_hip_params &p, std::map<uint32_t, std::vector<std::pair<uint32_t, uint32_t>>> &out)
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

void free_tables(mars_hip_ctx *ctx)
{
    for (Table &t : ctx->tables) {
        (void)hipFree(t.d_rows);
        (void)hipFree(t.d_cols);
        (void)hipFree(t.d_rng8);
        (void)hipFree(t.d_rngf);
        (void)hipFree(t.d_rstat);
        (void)hipFree(t.d_gthr);
        (void)hipFree(t.d_posinfo);
        (void)hipFree(t.d_posd);
        (void)hipFree(t.d_partial);
        (void)hipFree(t.d_out);
        (void)hipFree(t.d_found);
    }
    ctx->tables.clear();
    (void)hipFree(ctx->d_px);
    (void)hipFree(ctx->d_d);
    (void)hipFree(ctx->d_df);
    (void)hipFree(ctx->d_lo);
    (void)hipFree(ctx->d_hi);
    (void)hipFree(ctx->d_yscale);
    (void)hipFree(ctx->d_sat1);
    (void)hipFree(ctx->d_sat2);
    ctx->d_yscale = nullptr;
    ctx->d_sat1 = nullptr;
    ctx->d_sat2 = nullptr;
    ctx->d_px = nullptr;
    ctx->d_d = nullptr;
    ctx->d_df = nullptr;
    ctx->d_lo = nullptr;
    ctx->d_hi = nullptr;
    if (ctx->graph_exec) (void)hipGraphExecDestroy(ctx->graph_exec);
    if (ctx->graph) (void)hipGraphDestroy(ctx->graph);
    ctx->graph_exec = nullptr;
    ctx->graph = nullptr;
    ctx->graph_kernel = -1;
    ctx->prepared = false;
}

void chunking(uint32_t ntiles, uint32_t ngroups, uint32_t *nchunks, uint32_t *tpc)
{
    if (ntiles == 0) {
        *nchunks = 1;
        *tpc = 1;
        return;
    }
    uint32_t want = std::max<uint32_t>(1, kTargetBlocks / std::max<uint32_t>(1, ngroups));
    uint32_t maxc = std::max<uint32_t>(1, ntiles / 8); // >= 2 tiles per wave
    uint32_t c = std::min(want, maxc);
    uint32_t t = (ntiles + c - 1) / c;
    *tpc = t;
    *nchunks = (ntiles + t - 1) / t;
}

SearchArgs search_args(const mars_hip_ctx *ctx, const Table &t, int kernel)
{
    SearchArgs a{};
    a.rng8 = t.d_rng8;
    a.rngf = t.d_rngf;
    a.rstat = t.d_rstat;
    a.gthr = t.d_gthr;
    a.posinfo = t.d_posinfo;
    a.posd = t.d_posd;
    a.lo = ctx->d_lo;
    a.hi = ctx->d_hi;
    a.copy_stride = ctx->copy_stride;
    a.pitch = ctx->pitch;
    a.dplanef = ctx->d_df;
    a.wc = ctx->wc;
    a.P = t.P;
    a.npc = t.npc;
    a.shift = t.shift;
    a.nr = t.nr;
    a.max_alfa = ctx->params.max_alfa;
    a.bits_alfa = ctx->params.bits_alfa;
    a.bits_beta = ctx->params.bits_beta;
    a.partial = t.d_partial;
    a.counters = ctx->d_counters;
    a.counter_slot = ceil_log2(t.size);
    a.stage2 = ctx->stage2;
    a.alfa_c = ctx->alfa_c;
    a.alfa_unit = ctx->alfa_unit;
    a.qa_max = (double)((1u << ctx->params.bits_alfa) - 1);
    a.qb_max = (double)((1u << ctx->params.bits_beta) - 1);
    a.yscale = ctx->d_yscale;
    if (kernel == MARS_HIP_KERNEL_MFMA_I8) {
        a.ngroups = t.nr_pad / 8;
        a.nchunks = t.nchunks_mfma;
        a.tiles_per_chunk = t.tpc_mfma;
    } else {
        a.ngroups = t.nr_pad / 4;
        a.nchunks = t.nchunks_fp32;
        a.tiles_per_chunk = t.tpc_fp32;
    }
    return a;
}

MergeArgs merge_args(const mars_hip_ctx *ctx, const Table &t, int kernel)
{
    MergeArgs m{};
    m.partial = t.d_partial;
    m.t1_exact = kernel == MARS_HIP_KERNEL_MFMA_I8 ? 1 : 0;
    m.nchunks = kernel == MARS_HIP_KERNEL_MFMA_I8 ? t.nchunks_mfma : t.nchunks_fp32;
    m.nr = t.nr;
    m.P = t.P;
    m.npc = t.npc;
    m.shift = t.shift;
    m.size = t.size;
    m.rows = t.d_rows;
    m.cols = t.d_cols;
    m.px = ctx->d_px;
    m.w = ctx->w;
    m.d = ctx->d_d;
    m.wc = ctx->wc;
    m.posinfo = t.d_posinfo;
    m.rstat = t.d_rstat;
    m.max_alfa = ctx->params.max_alfa;
    m.bits_alfa = ctx->params.bits_alfa;
    m.bits_beta = ctx->params.bits_beta;
    m.out = t.d_out;
    m.found = t.d_found;
    return m;
}

void enqueue_table(mars_hip_ctx *ctx, Table &t, int kernel, hipStream_t st)
{
    mars::launch_possums(ctx->d_sat1, ctx->d_sat2, ctx->wc, t.size, t.shift, t.npc, t.P,
                         t.d_posinfo, t.d_posd, st);
    mars::launch_ranges(ctx->d_px, ctx->w, t.d_rows, t.d_cols, t.nr, t.nr_pad, t.size, t.kmem,
                        t.d_rng8, kernel == MARS_HIP_KERNEL_FP32 ? t.d_rngf : nullptr, t.d_rstat,
                        t.d_gthr, st);
    if (kernel == MARS_HIP_KERNEL_MFMA_I8)
        mars::launch_seed(t.d_rng8, t.d_rstat, ctx->d_d, ctx->wc, t.size, t.kmem, t.npc, t.P,
                          t.d_posinfo, t.nr, ctx->params.max_alfa, ctx->params.bits_alfa,
                          ctx->params.bits_beta, t.d_gthr, st);
}

// Enqueue one whole step.  With ev (3 + n_tables events: before the pool,
// after pool/range prep, after each table's search, after the merges) the
// step runs serially on the context stream so each launch can be timed;
// without, each table's chain (prep, search, merge) runs on its own side
// stream forked after the pool build and joined at the end.
int enqueue_step(mars_hip_ctx *ctx, int kernel, hipEvent_t *ev)
{
    hipStream_t st = ctx->stream;
    if (ev) HIP_TRY(ctx, hipEventRecord(ev[0], st));
    mars::launch_contract(ctx->d_px, ctx->w, ctx->wc, ctx->hc, ctx->pitch, ctx->copy_stride,
                          ctx->d_d, ctx->d_df, ctx->d_lo, ctx->d_hi, ctx->d_counters, kCounters,
                          st);
    mars::launch_sat(ctx->d_d, ctx->wc, ctx->hc, ctx->d_sat1, ctx->d_sat2, st);
    if (!ev) {
        HIP_TRY(ctx, hipEventRecord(ctx->fork_ev, st));
        const int nside = std::min<int>(kSideStreams, (int)ctx->tables.size());
        for (int j = 0; j < nside; j++) HIP_TRY(ctx, hipStreamWaitEvent(ctx->side[j], ctx->fork_ev, 0));
        for (size_t i = 0; i < ctx->tables.size(); i++) {
            Table &t = ctx->tables[i];
            hipStream_t ss = ctx->side[i % kSideStreams];
            enqueue_table(ctx, t, kernel, ss);
            if (t.P > 0 && t.nr > 0) {
                SearchArgs a = search_args(ctx, t, kernel);
                if (mars::launch_search(kernel, t.size, a, ss) != 0)
                    return fail(ctx, MARS_HIP_E_UNSUPPORTED, "no search kernel for this size");
            }
            mars::launch_merge(merge_args(ctx, t, kernel), ss);
        }
        for (int j = 0; j < nside; j++) {
            HIP_TRY(ctx, hipEventRecord(ctx->join_ev[j], ctx->side[j]));
            HIP_TRY(ctx, hipStreamWaitEvent(st, ctx->join_ev[j], 0));
        }
        HIP_TRY(ctx, hipGetLastError());
        return MARS_HIP_OK;
    }
    for (Table &t : ctx->tables) enqueue_table(ctx, t, kernel, st);
    HIP_TRY(ctx, hipEventRecord(ev[1], st));
    for (size_t i = 0; i < ctx->tables.size(); i++) {
        Table &t = ctx->tables[i];
        if (t.P > 0 && t.nr > 0) {
            SearchArgs a = search_args(ctx, t, kernel);
            if (mars::launch_search(kernel, t.size, a, st) != 0)
                return fail(ctx, MARS_HIP_E_UNSUPPORTED, "no search kernel for this size");
        }
        HIP_TRY(ctx, hipEventRecord(ev[2 + i], st));
    }
    for (Table &t : ctx->tables) mars::launch_merge(merge_args(ctx, t, kernel), st);
    HIP_TRY(ctx, hipEventRecord(ev[2 + ctx->tables.size()], st));
    HIP_TRY(ctx, hipGetLastError());
    return MARS_HIP_OK;
}

int download(mars_hip_ctx *ctx)
{
    for (Table &t : ctx->tables) {
        t.h_out.resize(t.nr);
        t.h_found.resize(t.nr);
        if (t.nr == 0) continue;
        HIP_TRY(ctx, hipMemcpyAsync(t.h_out.data(), t.d_out, sizeof(mars_hip_candidate) * t.nr,
                                    hipMemcpyDeviceToHost, ctx->stream));
        HIP_TRY(ctx, hipMemcpyAsync(t.h_found.data(), t.d_found, t.nr, hipMemcpyDeviceToHost,
                                    ctx->stream));
    }
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return MARS_HIP_OK;
}

uint64_t table_evals(const Table &t) { return (uint64_t)t.P * 8; }

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
              hipEventCreateWithFlags(&ctx->fork_ev, hipEventDisableTiming) == hipSuccess;
    for (int j = 0; ok && j < kSideStreams; j++)
        ok = hipStreamCreate(&ctx->side[j]) == hipSuccess &&
             hipEventCreateWithFlags(&ctx->join_ev[j], hipEventDisableTiming) == hipSuccess;
    if (!ok) {
        mars_hip_close(ctx);
        return MARS_HIP_E_DEVICE;
    }
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
    free_tables(ctx);
    (void)hipFree(ctx->d_counters);
    for (int j = 0; j < kSideStreams; j++) {
        if (ctx->side[j]) (void)hipStreamDestroy(ctx->side[j]);
        if (ctx->join_ev[j]) (void)hipEventDestroy(ctx->join_ev[j]);
    }
    if (ctx->fork_ev) (void)hipEventDestroy(ctx->fork_ev);
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

int mars_hip_prepare(mars_hip_ctx *ctx, const uint8_t *px, uint32_t w, uint32_t h,
                     const mars_hip_params *p, const uint32_t *shifts, uint32_t n_shifts)
{
    if (!ctx || !px || !p || !shifts || n_shifts == 0) return fail(ctx, MARS_HIP_E_ARG, "null argument");
    if (w == 0 || h == 0 || w > 16384 || h > 16384) return fail(ctx, MARS_HIP_E_ARG, "bad plane size");
    if (p->bits_alfa < 1 || p->bits_alfa > 16 || p->bits_beta < 1 || p->bits_beta > 16 ||
        !(p->max_alfa > 0.0) || p->min_size < 1 || p->max_size < p->min_size)
        return fail(ctx, MARS_HIP_E_ARG, "bad params");
    for (uint32_t i = 0; i < n_shifts; i++)
        if (shifts[i] == 0) return fail(ctx, MARS_HIP_E_ARG, "shift must be > 0");
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    free_tables(ctx);
    ctx->params = *p;
    ctx->w = w;
    ctx->h = h;
    ctx->wc = w / 2;
    ctx->hc = h / 2;
    ctx->pitch = ((ctx->wc + 15) / 16) * 16 + 16;
    ctx->copy_stride = (size_t)ctx->pitch * std::max<uint32_t>(ctx->hc, 1);

    std::map<uint32_t, std::vector<std::pair<uint32_t, uint32_t>>> nodes;
    uint32_t vsize = 1u << ceil_log2(std::max(w, h));
    collect(0, 0, vsize, w, h, *p, nodes);
    for (auto &kv : nodes)
        if (kv.first > kMaxSize)
            return fail(ctx, MARS_HIP_E_UNSUPPORTED, "range sizes above 32 are not supported");

    HIP_TRY(ctx, hipMalloc((void **)&ctx->d_px, (size_t)w * h));
    HIP_TRY(ctx, hipMemcpy(ctx->d_px, px, (size_t)w * h, hipMemcpyHostToDevice));
    size_t nc = std::max<size_t>((size_t)ctx->wc * ctx->hc, 1);
    HIP_TRY(ctx, hipMalloc((void **)&ctx->d_d, nc * sizeof(int32_t)));
    HIP_TRY(ctx, hipMalloc((void **)&ctx->d_df, nc * sizeof(float)));
    HIP_TRY(ctx, hipMalloc((void **)&ctx->d_lo, 16 * ctx->copy_stride));
    HIP_TRY(ctx, hipMalloc((void **)&ctx->d_hi, 16 * ctx->copy_stride));
    HIP_TRY(ctx, hipMemset(ctx->d_lo, 0, 16 * ctx->copy_stride));
    HIP_TRY(ctx, hipMemset(ctx->d_hi, 0, 16 * ctx->copy_stride));
    size_t nsat = (size_t)(ctx->wc + 1) * (ctx->hc + 1);
    HIP_TRY(ctx, hipMalloc((void **)&ctx->d_sat1, nsat * sizeof(int64_t)));
    HIP_TRY(ctx, hipMalloc((void **)&ctx->d_sat2, nsat * sizeof(int64_t)));

    // Stage-2 constants: exact only when max_alfa is a power of two (then the
    // alfa scaling of encode.rs:264 is exact in binary64).
    {
        int e = 0;
        double mant = std::frexp(p->max_alfa, &e);
        ctx->stage2 = (mant == 0.5 && p->bits_alfa <= 8 && p->bits_beta <= 12) ? 1 : 0;
        ctx->alfa_c = std::ldexp(1.0, (int)p->bits_alfa + 2) / p->max_alfa;
        ctx->alfa_unit = p->max_alfa / std::ldexp(1.0, (int)p->bits_alfa);
        uint32_t nq = 1u << p->bits_alfa;
        std::vector<double> ys(nq);
        double qbm = (double)((1u << p->bits_beta) - 1);
        for (uint32_t q = 0; q < nq; q++) {
            double a2 = (double)q / (double)(1u << p->bits_alfa) * p->max_alfa;
            ys[q] = qbm / ((1.0 + std::fabs(a2)) * 255.0);
        }
        HIP_TRY(ctx, hipMalloc((void **)&ctx->d_yscale, sizeof(double) * nq));
        HIP_TRY(ctx, hipMemcpy(ctx->d_yscale, ys.data(), sizeof(double) * nq, hipMemcpyHostToDevice));
    }

    std::vector<uint32_t> sh(shifts, shifts + n_shifts);
    std::sort(sh.begin(), sh.end());
    sh.erase(std::unique(sh.begin(), sh.end()), sh.end());
    for (uint32_t shift : sh) {
        for (auto it = nodes.rbegin(); it != nodes.rend(); ++it) {
            Table t;
            t.size = it->first;
            t.shift = shift;
            for (auto &rc : it->second) {
                t.index[((uint64_t)rc.first << 32) | rc.second] = (uint32_t)t.rows.size();
                t.rows.push_back(rc.first);
                t.cols.push_back(rc.second);
            }
            t.nr = (uint32_t)t.rows.size();
            t.nr_pad = (t.nr + 7) / 8 * 8;
            uint32_t s = t.size, K = s * s;
            t.kmem = K < 16 ? 16 : K;
            if (h >= 2 * s && w >= 2 * s) {
                t.npr = (h - 2 * s) / shift + 1;
                t.npc = (w - 2 * s) / shift + 1;
            }
            t.P = t.npr * t.npc;
            if (t.P > (0xffffffffu / 8)) return fail(ctx, MARS_HIP_E_UNSUPPORTED, "too many positions");
            chunking((t.P + 31) / 32, t.nr_pad / 8, &t.nchunks_mfma, &t.tpc_mfma);
            chunking((t.P + 63) / 64, t.nr_pad / 4, &t.nchunks_fp32, &t.tpc_fp32);
            uint32_t nch = std::max(t.nchunks_mfma, t.nchunks_fp32);
            HIP_TRY(ctx, hipMalloc((void **)&t.d_rows, sizeof(uint32_t) * std::max<uint32_t>(t.nr, 1)));
            HIP_TRY(ctx, hipMalloc((void **)&t.d_cols, sizeof(uint32_t) * std::max<uint32_t>(t.nr, 1)));
            if (t.nr) {
                HIP_TRY(ctx, hipMemcpy(t.d_rows, t.rows.data(), 4 * t.nr, hipMemcpyHostToDevice));
                HIP_TRY(ctx, hipMemcpy(t.d_cols, t.cols.data(), 4 * t.nr, hipMemcpyHostToDevice));
            }
            HIP_TRY(ctx, hipMalloc((void **)&t.d_rng8, (size_t)t.nr_pad * 8 * t.kmem));
            HIP_TRY(ctx, hipMalloc((void **)&t.d_rngf, (size_t)t.nr_pad / 4 * K * 32 * sizeof(float)));
            HIP_TRY(ctx, hipMalloc((void **)&t.d_rstat, sizeof(int32_t) * 2 * t.nr_pad));
            HIP_TRY(ctx, hipMalloc((void **)&t.d_gthr, sizeof(uint32_t) * t.nr_pad));
            HIP_TRY(ctx, hipMalloc((void **)&t.d_posinfo, sizeof(uint4) * std::max<uint32_t>(t.P, 1)));
            HIP_TRY(ctx, hipMalloc((void **)&t.d_posd, 2 * sizeof(double) * std::max<uint32_t>(t.P, 1)));
            HIP_TRY(ctx, hipMalloc((void **)&t.d_partial, sizeof(Partial) * (size_t)t.nr_pad * nch));
            HIP_TRY(ctx, hipMalloc((void **)&t.d_out, sizeof(mars_hip_candidate) * std::max<uint32_t>(t.nr, 1)));
            HIP_TRY(ctx, hipMalloc((void **)&t.d_found, std::max<uint32_t>(t.nr, 1)));
            ctx->tables.push_back(std::move(t));
        }
    }
    int rc = enqueue_step(ctx, ctx->kernel, nullptr);
    if (rc) return rc;
    rc = download(ctx);
    if (rc) return rc;
    ctx->prepared = true;
    return MARS_HIP_OK;
}

int mars_hip_run(mars_hip_ctx *ctx, int iters, int use_graph, mars_hip_timing *timing)
{
    if (!ctx || iters < 1) return fail(ctx, MARS_HIP_E_ARG, "bad argument");
    if (!ctx->prepared) return fail(ctx, MARS_HIP_E_NOT_PREPARED, "prepare() first");
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    const size_t nt = ctx->tables.size();
    std::vector<hipEvent_t> ev(3 + nt);
    for (auto &e : ev) HIP_TRY(ctx, hipEventCreate(&e));
    mars_hip_timing tm{};
    // Per-phase breakdown from one eager step.
    int rc = enqueue_step(ctx, ctx->kernel, ev.data());
    if (rc) return rc;
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    float ms = 0;
    HIP_TRY(ctx, hipEventElapsedTime(&ms, ev[0], ev[1]));
    tm.ms_pool = ms;
    for (size_t i = 0; i < nt; i++) {
        HIP_TRY(ctx, hipEventElapsedTime(&ms, i == 0 ? ev[1] : ev[1 + i], ev[2 + i]));
        tm.ms_search += ms;
        uint32_t lg = ceil_log2(ctx->tables[i].size);
        if (lg >= 1 && lg <= 6) tm.ms_by_size[lg - 1] += ms;
    }
    unsigned long long cnt[kCounters] = {};
    HIP_TRY(ctx, hipMemcpy(cnt, ctx->d_counters, sizeof(cnt), hipMemcpyDeviceToHost));
    for (int i = 0; i < 6; i++) {
        tm.survivors_by_size[i] = cnt[i + 1];
        tm.exact_fits_by_size[i] = cnt[8 + i + 1];
        tm.survivors += cnt[i + 1];
    }
    for (const Table &t : ctx->tables) tm.evals += (uint64_t)t.nr * table_evals(t);

    // Timed steps.
    if (use_graph && (ctx->graph_exec == nullptr || ctx->graph_kernel != ctx->kernel)) {
        if (ctx->graph_exec) (void)hipGraphExecDestroy(ctx->graph_exec);
        if (ctx->graph) (void)hipGraphDestroy(ctx->graph);
        ctx->graph_exec = nullptr;
        ctx->graph = nullptr;
        HIP_TRY(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
        rc = enqueue_step(ctx, ctx->kernel, nullptr);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(ctx->stream, &g);
        if (rc) return rc;
        if (e != hipSuccess) return fail(ctx, MARS_HIP_E_DEVICE, "graph capture failed");
        ctx->graph = g;
        HIP_TRY(ctx, hipGraphInstantiate(&ctx->graph_exec, g, nullptr, nullptr, 0));
        ctx->graph_kernel = ctx->kernel;
    }
    HIP_TRY(ctx, hipEventRecord(ev[0], ctx->stream));
    for (int it = 0; it < iters; it++) {
        if (use_graph) {
            HIP_TRY(ctx, hipGraphLaunch(ctx->graph_exec, ctx->stream));
        } else {
            rc = enqueue_step(ctx, ctx->kernel, nullptr);
            if (rc) return rc;
        }
    }
    HIP_TRY(ctx, hipEventRecord(ev[1], ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    HIP_TRY(ctx, hipEventElapsedTime(&ms, ev[0], ev[1]));
    tm.ms_step = ms / iters;
    for (auto &e : ev) (void)hipEventDestroy(e);
    rc = download(ctx);
    if (rc) return rc;
    if (timing) *timing = tm;
    return MARS_HIP_OK;
}

int mars_hip_search(const mars_hip_ctx *ctx, uint32_t row, uint32_t col, uint32_t size,
                    uint32_t shift, uint32_t k, mars_hip_candidate *out, uint32_t *n_out,
                    uint64_t *evals)
{
    mars_hip_ctx *mctx = const_cast<mars_hip_ctx *>(ctx);
    if (!ctx || !out || !n_out || !evals) return fail(mctx, MARS_HIP_E_ARG, "null argument");
    if (k != 1) return fail(mctx, MARS_HIP_E_UNSUPPORTED, "only k == 1 is supported");
    if (!ctx->prepared) return fail(mctx, MARS_HIP_E_NOT_PREPARED, "prepare() first");
    for (const Table &t : ctx->tables) {
        if (t.size != size || t.shift != shift) continue;
        auto it = t.index.find(((uint64_t)row << 32) | col);
        if (it == t.index.end()) break;
        uint32_t i = it->second;
        *evals = table_evals(t);
        if (t.h_found[i]) {
            *out = t.h_out[i];
            *n_out = 1;
        } else {
            *n_out = 0;
        }
        return MARS_HIP_OK;
    }
    char buf[160];
    snprintf(buf, sizeof buf, "range (%u,%u) size %u shift %u was not prepared", row, col, size,
             shift);
    return fail(mctx, MARS_HIP_E_NOT_IN_TABLE, buf);
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

int mars_hip_table_read(const mars_hip_ctx *ctx, uint32_t i, uint32_t *rows, uint32_t *cols,
                        mars_hip_candidate *out, uint8_t *found)
{
    if (!ctx || i >= ctx->tables.size()) return MARS_HIP_E_ARG;
    if (!ctx->prepared) return MARS_HIP_E_NOT_PREPARED;
    const Table &t = ctx->tables[i];
    if (rows) memcpy(rows, t.rows.data(), 4 * t.nr);
    if (cols) memcpy(cols, t.cols.data(), 4 * t.nr);
    if (out) memcpy(out, t.h_out.data(), sizeof(mars_hip_candidate) * t.nr);
    if (found) memcpy(found, t.h_found.data(), t.nr);
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

static int table_search_cb(void *user, uint32_t row, uint32_t col, uint32_t size,
                           uint32_t shift, mars_hip_candidate *out, uint32_t *n_out,
                           uint64_t *evals)
{
    return mars_hip_search((const mars_hip_ctx *)user, row, col, size, shift, 1, out, n_out,
                           evals);
}

int mars_hip_encode_plane(mars_hip_ctx *ctx, const uint8_t *px, uint32_t w, uint32_t h,
                          const mars_encode_params *p, const mars_encode_options *o,
                          uint8_t **out, size_t *out_len, mars_encode_stats *stats)
{
    if (!ctx || !p || !o) return fail(ctx, MARS_HIP_E_ARG, "null argument");
    mars_hip_params hp{};
    hp.min_size = p->min_size;
    hp.max_size = p->max_size;
    hp.shift = p->shift;
    hp.bits_alfa = p->bits_alfa;
    hp.bits_beta = p->bits_beta;
    hp.max_alfa = p->max_alfa;
    // The walks request stride `shift`, and 2*shift for flat blocks when
    // adaptive density is on (encode.rs:432-437, RD path only).
    uint32_t shifts[2] = {p->shift, 2 * p->shift};
    const uint32_t ns = (o->adaptive_density && p->has_lambda) ? 2 : 1;
    int rc = mars_hip_prepare(ctx, px, w, h, &hp, shifts, ns);
    if (rc) return rc;
    rc = mars_encode_plane(px, w, h, p, o, table_search_cb, ctx, table_search_cb, ctx, out,
                           out_len, stats);
    if (rc && ctx->err.empty()) return fail(ctx, rc, "encode failed");
    return rc;
}

} // extern "C"
