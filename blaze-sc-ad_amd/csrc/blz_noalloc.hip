// blz_noalloc.hip -- blz_prelim_run_noalloc: the prelim stage of one query over a resident
// database with every device work buffer inside one fixed array the library holds per GPU, for
// callers that cannot allocate (code under HIP graph capture, tight serving loops, harnesses that
// forbid hipMalloc).  The kernels are the engine's own (the first seed kernel, the segment and
// wide gapped passes); the lists are put in subject order on the device and copied straight into
// the caller's sink, so the call also needs no pinned staging and no host heap.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <mutex>

#include "blz_host.h"

namespace blz {
namespace {

constexpr size_t kNaBytes = 64ull << 20;        // the scratch array: 64 MiB per GPU
constexpr int64_t kNaMaxSubjects = 1 << 17;     // largest database the entry point takes
constexpr int kNaStageCap = 256;                // init hits per subject staged by the seed pass
constexpr int kNaSeedWaves = 2048;              // seed waves at most (one stage each)
constexpr size_t kNaPool = 1 << 19;             // init hits of a run
constexpr size_t kNaHsps = 1 << 18;             // HSPs of a run
constexpr int kNaSaved = 32;                    // HSPs per subject in the segment pass
constexpr int kNaMaxHeavyBlocks = 2048;         // segment-pass workgroups at most
constexpr int kNaWideLanes = 64;                // wide pass: one workgroup
constexpr int kNaWideRing = 2048;               // wide pass: DP band ring per lane
constexpr int kNaWideSaved = 1024;              // wide pass: HSPs per subject
constexpr int kNaChunk = 1024;                  // subjects per ordering workgroup
constexpr int kNaMaxChunks = (int)(kNaMaxSubjects / kNaChunk);

// Every work buffer of the entry point.  In host code this name is only a stand-in for the symbol:
// each device's copy sits at its own address, looked up once per device (noalloc_setup) and used
// through g_scratch[device].
static __device__ __align__(256) uint8_t scratch[kNaBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
uint8_t *g_scratch[kMaxDevices] = {};
bool g_looked_up[kMaxDevices] = {};

// The run's buffers carved from one device's scratch (256-byte aligned).
struct Layout {
    DevCounters *ctr;
    uint32_t *totals;      // [2]: lists, HSPs
    QueryDev *qd;
    uint32_t *meta;        // the query's own member table
    uint32_t *idx;         // [nsubj]: result index + 1 of each subject with a list
    uint32_t *chunk_lists, *chunk_hsps;
    InitHit *stage, *pool;
    WorkItem *work;
    uint32_t *order, *redo;
    ResultItem *results;
    int64_t *ovf;
    Hsp *hsps, *saved, *saved_wide, *out_hsps;
    int2 *dp;
    int64_t *oids;
    int32_t *counts;
    size_t used;
};

bool carve(uint8_t *base, size_t seed_waves, size_t heavy_blocks, Layout &L) {
    size_t at = 0;
    auto take = [&](size_t bytes) {
        uint8_t *p = base + at;
        at = (at + bytes + 255) & ~size_t(255);
        return p;
    };
    L.ctr = reinterpret_cast<DevCounters *>(take(sizeof(DevCounters)));
    L.totals = reinterpret_cast<uint32_t *>(take(2 * sizeof(uint32_t)));
    L.qd = reinterpret_cast<QueryDev *>(take(sizeof(QueryDev)));
    L.meta = reinterpret_cast<uint32_t *>(take(3 * sizeof(uint32_t)));
    L.idx = reinterpret_cast<uint32_t *>(take(kNaMaxSubjects * sizeof(uint32_t)));
    L.chunk_lists = reinterpret_cast<uint32_t *>(take(kNaMaxChunks * sizeof(uint32_t)));
    L.chunk_hsps = reinterpret_cast<uint32_t *>(take(kNaMaxChunks * sizeof(uint32_t)));
    L.stage = reinterpret_cast<InitHit *>(take(seed_waves * kNaStageCap * sizeof(InitHit)));
    L.pool = reinterpret_cast<InitHit *>(take(kNaPool * sizeof(InitHit)));
    L.work = reinterpret_cast<WorkItem *>(take(kNaMaxSubjects * sizeof(WorkItem)));
    L.order = reinterpret_cast<uint32_t *>(take(kNaMaxSubjects * sizeof(uint32_t)));
    L.redo = reinterpret_cast<uint32_t *>(take(kNaMaxSubjects * sizeof(uint32_t)));
    L.results = reinterpret_cast<ResultItem *>(take(kNaMaxSubjects * sizeof(ResultItem)));
    L.ovf = reinterpret_cast<int64_t *>(take(kNaMaxSubjects * sizeof(int64_t)));
    L.hsps = reinterpret_cast<Hsp *>(take(kNaHsps * sizeof(Hsp)));
    L.saved = reinterpret_cast<Hsp *>(take(heavy_blocks * 4 * kNaSaved * sizeof(Hsp)));
    L.dp = reinterpret_cast<int2 *>(take((size_t)kNaWideLanes * kNaWideRing * sizeof(int2)));
    L.saved_wide = reinterpret_cast<Hsp *>(take((size_t)kNaWideLanes * kNaWideSaved * sizeof(Hsp)));
    L.oids = reinterpret_cast<int64_t *>(take(kNaMaxSubjects * sizeof(int64_t)));
    L.counts = reinterpret_cast<int32_t *>(take(kNaMaxSubjects * sizeof(int32_t)));
    L.out_hsps = reinterpret_cast<Hsp *>(take(kNaHsps * sizeof(Hsp)));
    L.used = at;
    return at <= kNaBytes;
}

// ---- lists in subject order: scatter by subject id, count per chunk, scan, place ----------

__global__ void __launch_bounds__(256) k_na_scatter(const ResultItem *res, const DevCounters *ctr, uint32_t *idx) {
    const uint32_t n = ctr->result_count < (uint32_t)kNaMaxSubjects ? ctr->result_count : (uint32_t)kNaMaxSubjects;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        idx[res[i].sid] = i + 1;
}

// Block-wide exclusive sums of two values per thread (1024 threads), and the block totals.
__device__ void block_exclusive2(uint32_t a, uint32_t b, uint32_t &ea, uint32_t &eb, uint32_t &ta, uint32_t &tb) {
    __shared__ uint32_t wa[16], wb[16];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t ia = a, ib = b;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t xa = __shfl_up(ia, o, 64), xb = __shfl_up(ib, o, 64);
        if (lane >= o) {
            ia += xa;
            ib += xb;
        }
    }
    if (lane == 63) {
        wa[wave] = ia;
        wb[wave] = ib;
    }
    __syncthreads();
    uint32_t pa = 0, pb = 0;
    ta = tb = 0;
    for (int w = 0; w < (int)(blockDim.x >> 6); w++) {
        if (w < wave) {
            pa += wa[w];
            pb += wb[w];
        }
        ta += wa[w];
        tb += wb[w];
    }
    ea = pa + ia - a;
    eb = pb + ib - b;
}

__global__ void __launch_bounds__(1024) k_na_count(const ResultItem *res, const uint32_t *idx, int64_t nsubj,
                                                   uint32_t *chunk_lists, uint32_t *chunk_hsps) {
    const int64_t s = (int64_t)blockIdx.x * kNaChunk + threadIdx.x;
    const uint32_t v = s < nsubj ? idx[s] : 0u;
    uint32_t ea, eb, ta, tb;
    block_exclusive2(v ? 1u : 0u, v ? (uint32_t)res[v - 1].n : 0u, ea, eb, ta, tb);
    if (threadIdx.x == 0) {
        chunk_lists[blockIdx.x] = ta;
        chunk_hsps[blockIdx.x] = tb;
    }
}

__global__ void __launch_bounds__(64) k_na_scan(uint32_t *chunk_lists, uint32_t *chunk_hsps, int nchunks,
                                                uint32_t *totals) {
    if (threadIdx.x != 0) return;
    uint32_t l = 0, h = 0;
    for (int c = 0; c < nchunks; c++) {
        const uint32_t cl = chunk_lists[c], ch = chunk_hsps[c];
        chunk_lists[c] = l;
        chunk_hsps[c] = h;
        l += cl;
        h += ch;
    }
    totals[0] = l;
    totals[1] = h;
}

__global__ void __launch_bounds__(1024) k_na_place(const ResultItem *res, const Hsp *pool, const uint32_t *idx,
                                                   int64_t nsubj, const uint32_t *chunk_lists,
                                                   const uint32_t *chunk_hsps, int64_t *oids, int32_t *counts,
                                                   Hsp *out) {
    const int64_t s = (int64_t)blockIdx.x * kNaChunk + threadIdx.x;
    const uint32_t v = s < nsubj ? idx[s] : 0u;
    const ResultItem r = v ? res[v - 1] : ResultItem{0, 0, 0, 0};
    uint32_t ea, eb, ta, tb;
    block_exclusive2(v ? 1u : 0u, v ? (uint32_t)r.n : 0u, ea, eb, ta, tb);
    if (!v) return;
    const uint32_t li = chunk_lists[blockIdx.x] + ea, hi = chunk_hsps[blockIdx.x] + eb;
    oids[li] = s;
    counts[li] = r.n;
    for (int k = 0; k < r.n; k++) out[hi + k] = pool[r.off + k];
}

// A failure message kept within the context's reserved error text (no allocation).
blz_status fail(blz_ctx *ctx, blz_status st, const char *what, hipError_t e = hipSuccess) {
    char buf[200];
    if (e != hipSuccess)
        std::snprintf(buf, sizeof(buf), "%s: %s", what, hipGetErrorString(e));
    else
        std::snprintf(buf, sizeof(buf), "%s", what);
    ctx->err.assign(buf);
    return st;
}

}  // namespace

// Looks up this device's copy of the scratch array (the device must be current); once per device.
bool noalloc_setup(int device) {
    if (device < 0 || device >= kMaxDevices) return false;
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_looked_up[device]) return g_scratch[device] != nullptr;
    g_looked_up[device] = true;
    void *p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return false;
    g_scratch[device] = static_cast<uint8_t *>(p);
    return true;
}

}  // namespace blz

extern "C" blz_status blz_prelim_run_noalloc(blz_ctx *ctx, const blz_query *q, const blz_db *db, blz_hsp_sink *sink,
                                             blz_run_stats *stats) {
    using namespace blz;
    if (!ctx || !q || !db || !sink || q->ctx->device != ctx->device || db->ctx->device != ctx->device)
        return BLZ_EINVAL;
    if (db->nsubj > kNaMaxSubjects)
        return fail(ctx, BLZ_EINVAL, "blz_prelim_run_noalloc takes databases of at most 131072 subjects");
    if (hipSetDevice(ctx->device) != hipSuccess) return fail(ctx, BLZ_EDEVICE, "hipSetDevice");
    uint8_t *base = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        base = ctx->device < kMaxDevices ? g_scratch[ctx->device] : nullptr;
    }
    if (!base) return fail(ctx, BLZ_EDEVICE, "the no-allocation scratch array is not available on this device");
    hipStream_t st = ctx->stream;

    // seed launch shape: the first seed kernel, its diagonal ring in LDS
    QueryDev d = query_dev(q);
    const uint32_t ring = d.ring_mask + 1;
    int waves = 4;
    while (waves > 1 && seed_lds_bytes(ring, waves) > ctx->lds_per_block) waves--;
    if (seed_lds_bytes(ring, waves) > ctx->lds_per_block)
        return fail(ctx, BLZ_EINVAL, "query too long for the diagonal ring in LDS");
    const int blocks = std::max(1, std::min(ctx->num_cus * 8 / waves, kNaSeedWaves / waves));
    const int heavy_blocks = std::min(kNaMaxHeavyBlocks, ctx->num_cus * gapped_seg_blocks_per_cu(ctx->tune.gapped_width == 16));
    Layout L;
    if (heavy_blocks < 1 || !carve(base, (size_t)blocks * waves, (size_t)heavy_blocks, L))
        return fail(ctx, BLZ_EINVAL, "launch shape does not fit the no-allocation scratch array");

    // the query's entry of the kernels' query table, with a member table of its own
    const uint32_t meta[3] = {0u, (uint32_t)q->qlen, 0u};
    d.member_start = L.meta;
    d.member_idx = L.meta + 2;
    hipError_t e = hipMemcpyAsync(L.meta, meta, sizeof(meta), hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(L.qd, &d, sizeof(d), hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = launch_fill_u32(reinterpret_cast<uint32_t *>(L.ctr), sizeof(DevCounters) / 4, 0, st);
    if (e == hipSuccess) e = launch_fill_u32(L.idx, (size_t)std::max<int64_t>(db->nsubj, 1), 0, st);
    if (e != hipSuccess) return fail(ctx, BLZ_EDEVICE, "no-allocation setup", e);

    SeedArgs sa{};
    sa.res = db->res;
    sa.offs = db->offs;
    sa.s_begin = 0;
    sa.s_end = db->nsubj;
    sa.query = d.query;
    sa.qlen = d.qlen;
    sa.prefix = d.prefix;
    sa.entries = d.entries;
    sa.window = d.window;
    sa.x_ungapped = d.x_ungapped;
    sa.ungapped_cutoff = d.ungapped_cutoff;
    sa.ring_mask = d.ring_mask;
    sa.stage = L.stage;
    sa.stage_cap = kNaStageCap;
    sa.pool = L.pool;
    sa.pool_cap = kNaPool;
    sa.work = L.work;
    sa.work_cap = (uint32_t)kNaMaxSubjects;
    sa.ovf = L.ovf;
    sa.ovf_cap = (uint32_t)kNaMaxSubjects;
    sa.ctr = L.ctr;

    GappedArgs g{};
    g.res = db->res;
    g.offs = db->offs;
    g.qd = L.qd;
    g.pool = L.pool;
    g.work = L.work;
    g.order = L.order;
    g.count = &L.ctr->work_count;
    g.count_cap = (uint32_t)kNaMaxSubjects;
    g.next = &L.ctr->next_work;
    g.next_heavy = &L.ctr->next_heavy;
    g.heavy_classes = kBuckets;  // every cost class on the segment pass
    g.redo = L.redo;
    g.redo_cap = (uint32_t)kNaMaxSubjects;
    g.dp_scratch = L.dp;
    g.dp_wmask = kNaWideRing - 1;
    g.hsp_scratch = L.saved;
    g.saved_cap = kNaSaved;
    g.hsps = L.hsps;
    g.hsp_cap = kNaHsps;
    g.results = L.results;
    g.ovf = L.ovf;
    g.ovf_cap = (uint32_t)kNaMaxSubjects;
    g.ctr = L.ctr;
    GappedArgs w = g;  // wide pass over the segment pass's redo list
    w.order = L.redo;
    w.count = &L.ctr->redo_count;
    w.next = &L.ctr->next_redo;
    w.heavy_classes = 0;
    w.redo = nullptr;
    w.redo_cap = 0;
    w.hsp_scratch = L.saved_wide;
    w.saved_cap = kNaWideSaved;

    const int nchunks = (int)((db->nsubj + kNaChunk - 1) / kNaChunk);
    const bool any = db->nsubj > 0;
    e = hipEventRecord(ctx->ev[0], st);
    if (e == hipSuccess && any) e = launch_seed(sa, blocks, waves, st);
    if (e == hipSuccess) e = hipEventRecord(ctx->ev[1], st);
    if (e == hipSuccess && any) e = launch_order(L.work, L.ctr, L.order, (uint32_t)kNaMaxSubjects, 128, st);
    if (e == hipSuccess && any) e = launch_gapped_seg(g, heavy_blocks, ctx->tune.gapped_width == 16, st);
    if (e == hipSuccess && any) e = launch_gapped_wide(w, 1, kNaWideLanes, st);
    if (e == hipSuccess) e = hipEventRecord(ctx->ev[2], st);
    if (e == hipSuccess && any) {
        hipLaunchKernelGGL(k_na_scatter, dim3(64), dim3(256), 0, st, L.results, L.ctr, L.idx);
        hipLaunchKernelGGL(k_na_count, dim3(nchunks), dim3(kNaChunk), 0, st, L.results, L.idx, db->nsubj,
                           L.chunk_lists, L.chunk_hsps);
        hipLaunchKernelGGL(k_na_scan, dim3(1), dim3(64), 0, st, L.chunk_lists, L.chunk_hsps, nchunks, L.totals);
        hipLaunchKernelGGL(k_na_place, dim3(nchunks), dim3(kNaChunk), 0, st, L.results, L.hsps, L.idx, db->nsubj,
                           L.chunk_lists, L.chunk_hsps, L.oids, L.counts, L.out_hsps);
        e = hipGetLastError();
    }
    if (e != hipSuccess) return fail(ctx, BLZ_EDEVICE, "no-allocation launch", e);
    DevCounters c{};
    uint32_t totals[2] = {0, 0};
    e = hipMemcpyAsync(&c, L.ctr, sizeof(c), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess && any) e = hipMemcpyAsync(totals, L.totals, sizeof(totals), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return fail(ctx, BLZ_EDEVICE, "no-allocation run", e);

    blz_status rc = BLZ_OK;
    sink->n_subjects = totals[0];
    sink->n_hsps = totals[1];
    if ((int64_t)totals[0] > sink->subject_cap || (int64_t)totals[1] > sink->hsp_cap) {
        rc = BLZ_ECAPACITY;
    } else if (totals[0] > 0) {
        e = hipMemcpyAsync(sink->oids, L.oids, sizeof(int64_t) * totals[0], hipMemcpyDeviceToHost, st);
        if (e == hipSuccess)
            e = hipMemcpyAsync(sink->counts, L.counts, sizeof(int32_t) * totals[0], hipMemcpyDeviceToHost, st);
        if (e == hipSuccess)
            e = hipMemcpyAsync(sink->hsps, L.out_hsps, sizeof(Hsp) * totals[1], hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        if (e != hipSuccess) return fail(ctx, BLZ_EDEVICE, "no-allocation lists back", e);
    }
    if (stats) {
        float ms_seed = 0, ms_gapped = 0;
        (void)hipEventElapsedTime(&ms_seed, ctx->ev[0], ctx->ev[1]);
        (void)hipEventElapsedTime(&ms_gapped, ctx->ev[1], ctx->ev[2]);
        stats->subjects += (uint64_t)db->nsubj;
        stats->residues += db->nres;
        stats->word_hits += c.word_hits;
        stats->ungapped_ext += c.ungapped_ext;
        stats->init_hits += c.init_hits;
        stats->gapped_subjects += std::min<uint64_t>(c.work_count, (uint64_t)kNaMaxSubjects);
        stats->gapped_ext += c.gapped_ext;
        stats->dp_cells += c.dp_cells;
        stats->hsps += totals[1];
        stats->overflow_subjects += c.overflow_count;
        stats->overflow_left += c.overflow_count;
        stats->ms_seed += ms_seed;
        stats->ms_gapped += ms_gapped;
        stats->ms_total += ms_seed + ms_gapped;
        stats->seed_launches += any ? 1 : 0;
        stats->gapped_launches += any ? 1 : 0;
    }
    if (c.overflow_count > 0 && rc == BLZ_OK)
        rc = fail(ctx, BLZ_EOVERFLOW, "subjects past the no-allocation caps have no list; run blz_prelim_run for them");
    return rc;
}
