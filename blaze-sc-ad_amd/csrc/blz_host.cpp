// blz_host.cpp -- host runtime behind the C ABI in include/blz_prelim.h.
//
// Owns the device (one context = one HIP device + one stream + reusable work buffers), the
// resident database, the per-query lookup tables, and the two-stage launch sequence.
// Nothing here falls back to a CPU implementation: if the device path cannot run, the call
// returns an error status.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "blz_common.h"
#include "blz_device.h"
#include "blz_prelim.h"
#include "blz_host.h"

namespace blz {
const int16_t kBlosum62[28][28] = {
#include "blz_blosum62.inc"
};
}  // namespace blz

using namespace blz;

namespace {

std::atomic<uint64_t> g_next_uid{1};

constexpr int kStageCap = 4096;   // init hits per subject kept on the device
constexpr int kDpRing = 4096;     // wide gapped pass: DP band ring per lane (power of two)
constexpr int kSavedCap = 128;    // LDS gapped pass: HSPs per subject before the purge
constexpr int kSavedCapWide = 4096;  // wide gapped pass: HSPs per subject before the purge
constexpr int kSavedCapWave = 1024;  // wave gapped pass: HSPs per subject before the purge
constexpr int kSeedWavesMax = 4;
constexpr int kGappedBlocks = 64;        // wide gapped pass (redo list): blocks x 64 lanes
constexpr size_t kRedoCap = 1 << 20;
constexpr size_t kWorkCapStart = 48u << 20; // work items a run starts with (grown when a run fills them)
constexpr int kRetryCap = 1 << 16;        // retry launch: init hits per subject
constexpr int kRetryBlocks = 16, kRetryWaves = 4;
constexpr size_t kOvfStart = 1 << 16;     // overflow-list entries a run starts with
// overflow queue: init hits staged per pair in its first round (x8 each further round), and the
// most device memory one round's per-pair staging and per-lane DP rings may take
constexpr size_t kQueueStage0 = 1 << 17;
constexpr size_t kQueueWaves = 1024;  // waves of one round's scan (one pair each at a time)
constexpr size_t kQueueBytes = 2ull << 30;
constexpr int kQueueRounds = 5;

}  // namespace

bool blz::check(blz_ctx *ctx, hipError_t e, const char *what) {
    if (e == hipSuccess) return true;
    ctx->err = std::string(what) + ": " + hipGetErrorString(e);
    return false;
}

namespace {

// Grows a device buffer to `need` elements (contents are not kept).  A buffer that a captured
// launch sequence may name bumps the context's generation, which retires the captured graphs;
// the overflow queue's own buffers (never captured) pass bump = false.
template <class T>
bool ensure(blz_ctx *ctx, T *&ptr, size_t &cap, size_t need, const char *what, bool bump = true) {
    if (need <= cap && ptr) return true;
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    cap = 0;
    if (bump) ctx->gen++;
    size_t n = std::max<size_t>(need, 1);
    if (!check(ctx, hipMalloc((void **)&ptr, n * sizeof(T)), what)) return false;
    cap = n;
    return true;
}

// Neighbourhoods of all 28^3 words at one threshold, computed once per process.
struct Neighbourhoods {
    int threshold = -1;
    std::vector<uint32_t> start;  // [kWords+1]
    std::vector<uint16_t> words;
};
std::mutex g_nb_mutex;
Neighbourhoods g_nb;

// {u : score(w,u) >= T} plus w itself (prelim_stage.cpp:135-155 and :378-386).
void build_neighbourhoods(int T, Neighbourhoods &nb) {
    nb.threshold = T;
    nb.start.assign(kWords + 1, 0);
    nb.words.clear();
    nb.words.reserve(800000);
    int rowmax[28];
    for (int x = 0; x < 28; x++) {
        rowmax[x] = -1000000;
        for (int i = 0; i < 28; i++) rowmax[x] = std::max<int>(rowmax[x], kBlosum62[x][i]);
    }
    for (int w = 0; w < kWords; w++) {
        const int x = w / 784, y = (w / 28) % 28, z = w % 28;
        nb.start[w] = (uint32_t)nb.words.size();
        for (int a = 0; a < 28; a++) {
            const int sa = kBlosum62[x][a];
            if (sa + rowmax[y] + rowmax[z] < T) continue;
            for (int b = 0; b < 28; b++) {
                const int sb = sa + kBlosum62[y][b];
                if (sb + rowmax[z] < T) continue;
                for (int c = 0; c < 28; c++)
                    if (sb + kBlosum62[z][c] >= T) nb.words.push_back((uint16_t)(a * 784 + b * 28 + c));
            }
        }
        if (kBlosum62[x][x] + kBlosum62[y][y] + kBlosum62[z][z] < T) nb.words.push_back((uint16_t)w);
    }
    nb.start[kWords] = (uint32_t)nb.words.size();
}

const Neighbourhoods &neighbourhoods(int T) {
    std::lock_guard<std::mutex> lk(g_nb_mutex);
    if (g_nb.threshold != T) build_neighbourhoods(T, g_nb);
    return g_nb;
}

// Seed kernel 2 ring: every diagonal a hit within window letters behind the current 64-letter
// step can touch, plus a step of slack; a multiple of 8 entries (16-byte clears).
constexpr uint32_t kRingLenMax = 30000;  // the multiply-high modulo is exact up to here
uint32_t ring_len_for(int qlen, int window) { return ((uint32_t)qlen + (uint32_t)window + 128u + 7u) & ~7u; }

void set_ring(QueryDev &d, int qlen, int window) {
    d.ring_len = ring_len_for(qlen, window);
    d.ring_base = ((65536u + d.ring_len - 1) / d.ring_len) * d.ring_len;
    d.ring_magic = (uint32_t)((((uint64_t)1 << 32) + d.ring_len - 1) / d.ring_len);
}

uint32_t ring_words_for(int qlen, int window) {
    uint32_t need = (uint32_t)qlen + 64u + (uint32_t)window + 64u;
    uint32_t m = 256;
    while (m < need) m <<= 1;
    return m;
}

// Lookup table: for every word u, the offsets q (ascending) whose word has u in its
// neighbourhood (prelim_stage.cpp:360-386; flattened like :534-571 but with a u32 prefix), plus
// the compact prefix the seed kernel keeps in LDS (u32 base per 64-word block, u16 per word).
blz_status build_tables(blz_ctx *ctx, const uint8_t *letters, int len, const std::vector<std::pair<int, int>> &segs,
                        int threshold, DevTables &t) {
    const Neighbourhoods &nb = neighbourhoods(threshold);
    std::vector<uint32_t> prefix(kWords + 1, 0);
    for (const auto &sg : segs)
        for (int q = sg.first; q + kWord <= sg.second; q++) {
            const int w = letters[q] * 784 + letters[q + 1] * 28 + letters[q + 2];
            for (uint32_t k = nb.start[w]; k < nb.start[w + 1]; k++) prefix[nb.words[k] + 1]++;
        }
    for (int u = 0; u < kWords; u++) prefix[u + 1] += prefix[u];
    // two spare entries: the seed kernel stages the table into LDS in 32-bit words
    std::vector<uint16_t> entries(prefix[kWords] + 2, 0);
    std::vector<uint32_t> cur(prefix.begin(), prefix.end() - 1);
    for (const auto &sg : segs)
        for (int q = sg.first; q + kWord <= sg.second; q++) {
            const int w = letters[q] * 784 + letters[q + 1] * 28 + letters[q + 2];
            for (uint32_t k = nb.start[w]; k < nb.start[w + 1]; k++) entries[cur[nb.words[k]]++] = (uint16_t)q;
        }
    std::vector<uint32_t> blockbase(kRelBlocks);
    std::vector<uint16_t> rel16(kWords + 1);
    bool compact_ok = true;
    for (int b = 0; b < kRelBlocks; b++) blockbase[b] = prefix[std::min(b * 64, kWords)];
    for (int u = 0; u <= kWords; u++) {
        const uint32_t r = prefix[u] - blockbase[u >> 6];
        if (r > 0xFFFFu) compact_ok = false;
        rel16[u] = (uint16_t)r;
    }
    // word records: count (saturating at 255: then read the prefix) | first entry << 8 in x,
    // the first kWordRecInline entries in y, z, w
    std::vector<uint4> rec(kWords);
    const bool rec_start_fits = prefix[kWords] < (1u << 24);
    for (int u = 0; u < kWords; u++) {
        const uint32_t a = prefix[u], n = prefix[u + 1] - a;
        uint16_t e[kWordRecInline] = {};
        for (uint32_t k = 0; k < n && k < (uint32_t)kWordRecInline; k++) e[k] = entries[a + k];
        const uint32_t cb = n < 255 && rec_start_fits ? n : 255u;
        rec[u] = uint4{cb | (rec_start_fits ? a << 8 : 0u), e[0] | (uint32_t)e[1] << 16, e[2] | (uint32_t)e[3] << 16,
                       e[4] | (uint32_t)e[5] << 16};
    }
    t.len = len;
    t.n_entries = prefix[kWords];
    t.compact_ok = compact_ok;
    // Stream-ordered allocations on the context's stream: releasing them later (hipFreeAsync)
    // does not synchronise the device, so a serving loop can drop queries while other
    // contexts' batches run.
    hipStream_t st = ctx->stream;
    t.st = ctx->async_alloc ? st : nullptr;
    auto alloc = [&](void **p, size_t n) { return ctx->async_alloc ? hipMallocAsync(p, n, st) : hipMalloc(p, n); };
    if (!check(ctx, alloc((void **)&t.d_query, len + 64), "query") ||
        !check(ctx, alloc((void **)&t.d_prefix, sizeof(uint32_t) * (kWords + 1)), "prefix") ||
        !check(ctx, alloc((void **)&t.d_entries, sizeof(uint16_t) * entries.size()), "entries") ||
        !check(ctx, alloc((void **)&t.d_rel16, sizeof(uint16_t) * (kWords + 1)), "rel16") ||
        !check(ctx, alloc((void **)&t.d_blockbase, sizeof(uint32_t) * kRelBlocks), "blockbase") ||
        !check(ctx, alloc((void **)&t.d_rec, sizeof(uint4) * kWords), "word records")) {
        t.release();
        return BLZ_ENOMEM;
    }
    std::vector<uint8_t> qpad(letters, letters + len);
    qpad.resize((size_t)len + 64, 0);
    if (!check(ctx, hipMemcpyAsync(t.d_query, qpad.data(), qpad.size(), hipMemcpyHostToDevice, st), "query copy") ||
        !check(ctx, hipMemcpyAsync(t.d_prefix, prefix.data(), sizeof(uint32_t) * (kWords + 1), hipMemcpyHostToDevice,
                                   st), "prefix copy") ||
        !check(ctx, hipMemcpyAsync(t.d_entries, entries.data(), sizeof(uint16_t) * entries.size(),
                                   hipMemcpyHostToDevice, st), "entries copy") ||
        !check(ctx, hipMemcpyAsync(t.d_rel16, rel16.data(), sizeof(uint16_t) * (kWords + 1), hipMemcpyHostToDevice,
                                   st), "rel16 copy") ||
        !check(ctx, hipMemcpyAsync(t.d_blockbase, blockbase.data(), sizeof(uint32_t) * kRelBlocks,
                                   hipMemcpyHostToDevice, st), "blockbase copy") ||
        !check(ctx, hipMemcpyAsync(t.d_rec, rec.data(), sizeof(uint4) * kWords, hipMemcpyHostToDevice, st),
               "word records copy") ||
        !check(ctx, hipStreamSynchronize(st), "table upload")) {
        t.release();
        return BLZ_EDEVICE;
    }
    return BLZ_OK;
}

// Residues to the device through two pinned staging buffers on a side stream: host threads
// check and copy chunk k + 1 into one buffer while the DMA of chunk k runs from the other.
// Replaces the per-chunk gather + pageable H2D of the reference (gpuBLAZE.cu:1587-1618).
constexpr size_t kUploadChunk = 64u << 20;
blz_status upload_residues(blz_ctx *ctx, uint8_t *dst, const uint8_t *src, uint64_t n) {
    if (n == 0) return BLZ_OK;
    if (!ctx->up_stream && !check(ctx, hipStreamCreateWithFlags(&ctx->up_stream, hipStreamNonBlocking), "upload stream"))
        return BLZ_EDEVICE;
    for (int b = 0; b < 2; b++) {
        if (!ctx->up_buf[b] && !check(ctx, hipHostMalloc((void **)&ctx->up_buf[b], kUploadChunk, hipHostMallocDefault),
                                      "pinned staging"))
            return BLZ_ENOMEM;
        if (!ctx->up_ev[b] && !check(ctx, hipEventCreateWithFlags(&ctx->up_ev[b], hipEventDisableTiming), "event"))
            return BLZ_EDEVICE;
    }
    const int nthreads = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    bool bad = false;
    for (uint64_t off = 0, k = 0; off < n; off += kUploadChunk, k++) {
        const size_t len = (size_t)std::min<uint64_t>(kUploadChunk, n - off);
        uint8_t *buf = ctx->up_buf[k & 1];
        if (k >= 2 && !check(ctx, hipEventSynchronize(ctx->up_ev[k & 1]), "upload"))  // buffer free again
            return BLZ_EDEVICE;
        std::vector<std::thread> th;
        std::vector<char> flag((size_t)nthreads, 0);
        const size_t part = (len + nthreads - 1) / nthreads;
        for (int t = 0; t < nthreads; t++)
            th.emplace_back([&, t] {
                const size_t a = std::min(len, t * part), e = std::min(len, a + part);
                uint8_t any = 0;
                for (size_t i = a; i < e; i++) {
                    const uint8_t v = src[off + i];
                    any |= (uint8_t)(v >= kAlpha);
                    buf[i] = v;
                }
                flag[t] = (char)(any != 0);
            });
        for (auto &x : th) x.join();
        for (char f : flag) bad |= f != 0;
        if (bad) break;
        if (!check(ctx, hipMemcpyAsync(dst + off, buf, len, hipMemcpyHostToDevice, ctx->up_stream), "db copy") ||
            !check(ctx, hipEventRecord(ctx->up_ev[k & 1], ctx->up_stream), "event"))
            return BLZ_EDEVICE;
    }
    if (!check(ctx, hipStreamSynchronize(ctx->up_stream), "db copy")) return BLZ_EDEVICE;
    if (bad) {
        ctx->err = "residue outside NCBIstdaa (>= 28)";
        return BLZ_EINVAL;
    }
    return BLZ_OK;
}

}  // namespace

extern "C" {

void blz_default_params(blz_params *p) {
    p->word_threshold = 11;
    p->window = 40;
    p->x_ungapped = 16;
    p->x_gapped = 38;
    p->gap_open = 11;
    p->gap_extend = 1;
    p->ungapped_cutoff = 41;
    p->gapped_cutoff = 67;
    p->reap_score = 0;
}

blz_status blz_ctx_create(int32_t device, blz_ctx **out) {
    if (!out) return BLZ_EINVAL;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0 || device < 0 || device >= n) return BLZ_EDEVICE;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) return BLZ_EDEVICE;
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) return BLZ_EDEVICE;
    blz_ctx *ctx = new blz_ctx();
    ctx->device = device;
    ctx->num_cus = prop.multiProcessorCount;
    ctx->lds_per_block = prop.sharedMemPerBlock;
    int pools = 0;
    ctx->async_alloc = hipDeviceGetAttribute(&pools, hipDeviceAttributeMemoryPoolsSupported, device) == hipSuccess &&
                       pools != 0;
    ctx->err.reserve(256);  // error text of the allocation-free entry point fits without growing
    if (!check(ctx, hipSetDevice(device), "hipSetDevice") ||
        !check(ctx, hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking), "stream")) {
        delete ctx;
        return BLZ_EDEVICE;
    }
    (void)noalloc_setup(device);  // this device's scratch array for blz_prelim_run_noalloc
    for (auto &e : ctx->ev)
        if (!check(ctx, hipEventCreate(&e), "event")) {
            delete ctx;
            return BLZ_EDEVICE;
        }
    if (!check(ctx, hipMalloc((void **)&ctx->ctr, sizeof(DevCounters)), "counters")) {
        delete ctx;
        return BLZ_ENOMEM;
    }
    *out = ctx;
    return BLZ_OK;
}

void blz_ctx_destroy(blz_ctx *ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    void *bufs[] = {ctx->retry_list, ctx->retry_stage, ctx->retry_ring, ctx->long_ring, ctx->items,
                    ctx->ctr,        ctx->stage,       ctx->pool,       ctx->work,      ctx->ovf,
                    ctx->dp,         ctx->saved,       ctx->hsps,       ctx->results,   ctx->qdev,
                    ctx->item_ctr,   ctx->order,       ctx->redo,       ctx->saved_wide, ctx->pack_meta,
                    ctx->q_ctr,      ctx->q_list,      ctx->q_stage,    ctx->q_ring,    ctx->q_pool,
                    ctx->q_work,     ctx->q_order,     ctx->q_ovf,      ctx->q_dp,      ctx->q_saved,
                    ctx->q_hsps,     ctx->q_results};
    ctx->drop_graph();
    tb_release(ctx->tb);
    for (auto &pk : ctx->packs) pk.t.release();
    for (void *b : bufs)
        if (b) (void)hipFree(b);
    for (auto &e : ctx->ev)
        if (e) (void)hipEventDestroy(e);
    if (ctx->pin) (void)hipHostFree(ctx->pin);
    for (int b = 0; b < 2; b++) {
        if (ctx->up_buf[b]) (void)hipHostFree(ctx->up_buf[b]);
        if (ctx->up_ev[b]) (void)hipEventDestroy(ctx->up_ev[b]);
    }
    if (ctx->up_stream) (void)hipStreamDestroy(ctx->up_stream);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

blz_status blz_ctx_set_tuning(blz_ctx *ctx, const blz_tuning *t) {
    if (!ctx || !t) return BLZ_EINVAL;
    if (t->seed_variant < 0 || t->seed_variant > 2 || t->seed_waves < 0 || t->seed_waves > 16 ||
        t->seed_blocks_per_cu < 0 || t->gapped_blocks < 0 || t->gapped_threads < 0 || t->gapped_threads > 1024 ||
        (t->gapped_threads % 64) != 0 || t->seed_table < 0 || t->seed_table > 3 || t->gapped_wave_classes < 0 ||
        t->gapped_wave_classes > kBuckets + 1 || t->query_packing < 0 || t->query_packing > 2 || t->pack_len < 0 ||
        t->pack_len > 16000 || t->bin_ring < 0 || t->graph < 0 || t->graph > 1 ||
        (t->gapped_lanes != 0 && t->gapped_lanes != 16 && t->gapped_lanes != 64) || t->subject_groups < 0 ||
        t->subject_groups > 1 || t->overflow_mode < 0 || t->overflow_mode > 1 || t->cap_stage < 0 ||
        t->cap_retry < 0 || t->cap_ovf < 0 || t->cap_wide_ring < 0 || t->cap_wide_saved < 0 || t->cap_work < 0 ||
        t->cap_work > 0xFFFFFFFFll || t->cap_pool < 0 || t->cap_hsps < 0 ||
        t->cap_tb_bytes < 0 || t->cap_tb_runs < 0 || (t->tb_lanes != 0 && t->tb_lanes != 16 && t->tb_lanes != 64) ||
        (t->gapped_width != 0 && t->gapped_width != 16 && t->gapped_width != 32))
        return BLZ_EINVAL;
    if (t->query_packing != ctx->tune.query_packing || t->pack_len != ctx->tune.pack_len) {
        // cached packs were made under the old choice
        for (auto &pk : ctx->packs) pk.t.release();
        ctx->packs.clear();
        ctx->pack_key.clear();
    }
    ctx->tune = *t;
    return BLZ_OK;
}

blz_status blz_ctx_get_tuning(const blz_ctx *ctx, blz_tuning *t) {
    if (!ctx || !t) return BLZ_EINVAL;
    *t = ctx->tune;
    return BLZ_OK;
}

const char *blz_last_error(const blz_ctx *ctx) { return ctx ? ctx->err.c_str() : "no context"; }

void *blz_ctx_stream(blz_ctx *ctx) { return ctx ? (void *)ctx->stream : nullptr; }

blz_status blz_db_upload(blz_ctx *ctx, const uint8_t *residues, const uint64_t *offsets, int64_t nsubj,
                         blz_db **out) {
    if (!ctx || !out || !offsets || nsubj < 0 || (nsubj > 0 && !residues)) return BLZ_EINVAL;
    *out = nullptr;
    if (offsets[0] != 0) {
        ctx->err = "offsets[0] must be 0";
        return BLZ_EINVAL;
    }
    uint64_t max_len = 0;
    for (int64_t i = 0; i < nsubj; i++) {
        if (offsets[i + 1] < offsets[i] || offsets[i + 1] - offsets[i] > kMaxSubjectLen) {
            ctx->err = "subject offsets must be non-decreasing and subjects at most 16,777,215 letters long";
            return BLZ_EINVAL;
        }
        max_len = std::max<uint64_t>(max_len, offsets[i + 1] - offsets[i]);
    }
    const uint64_t nres = offsets[nsubj];
    (void)hipSetDevice(ctx->device);
    blz_db *db = new blz_db();
    db->ctx = ctx;
    db->uid = g_next_uid++;
    db->nsubj = nsubj;
    db->nres = nres;
    db->max_len = max_len;
    // 64 padding bytes so word reads near the end stay inside the allocation
    if (!check(ctx, hipMalloc((void **)&db->res, nres + 64), "db residues") ||
        !check(ctx, hipMalloc((void **)&db->offs, sizeof(uint64_t) * (nsubj + 1)), "db offsets")) {
        blz_db_free(db);
        return BLZ_ENOMEM;
    }
    const blz_status up = upload_residues(ctx, db->res, residues, nres);
    if (up != BLZ_OK) {
        blz_db_free(db);
        return up;
    }
    if (!check(ctx, hipMemset(db->res + nres, 0, 64), "db pad") ||
        !check(ctx, hipMemcpy(db->offs, offsets, sizeof(uint64_t) * (nsubj + 1), hipMemcpyHostToDevice),
               "db copy")) {
        blz_db_free(db);
        return BLZ_EDEVICE;
    }
    db->h_offs.assign(offsets, offsets + nsubj + 1);
    std::vector<int64_t> long_ids;
    for (int64_t i = 0; i < nsubj; i++)
        if (offsets[i + 1] - offsets[i] > (uint64_t)kSeedNarrowMax) long_ids.push_back(i);
    db->n_long = (int64_t)long_ids.size();
    if (!long_ids.empty() &&
        (!check(ctx, hipMalloc((void **)&db->long_ids, sizeof(int64_t) * long_ids.size()), "long subjects") ||
         !check(ctx, hipMemcpy(db->long_ids, long_ids.data(), sizeof(int64_t) * long_ids.size(), hipMemcpyHostToDevice),
                "long subjects"))) {
        blz_db_free(db);
        return BLZ_EDEVICE;
    }
    *out = db;
    return BLZ_OK;
}

void blz_db_free(blz_db *db) {
    if (!db) return;
    if (db->ctx) db->ctx->drop_graph();  // a captured batch may read this database
    if (db->res) (void)hipFree(db->res);
    if (db->offs) (void)hipFree(db->offs);
    if (db->long_ids) (void)hipFree(db->long_ids);
    delete db;
}

int64_t blz_db_num_subjects(const blz_db *db) { return db ? db->nsubj : 0; }
uint64_t blz_db_num_residues(const blz_db *db) { return db ? db->nres : 0; }

blz_status blz_query_prepare(blz_ctx *ctx, const uint8_t *query, int32_t qlen, const blz_params *params,
                             blz_query **out) {
    if (!ctx || !out || !query || qlen < 1 || qlen > 16000) {
        if (ctx) ctx->err = "query length must be 1..16000";
        return BLZ_EINVAL;
    }
    *out = nullptr;
    blz_params p;
    if (params)
        p = *params;
    else
        blz_default_params(&p);
    // walks stop at sequence ends and pack separators on letter 0's score of -32768, which
    // must exceed any ungapped X-drop
    if (p.window < 1 || p.x_ungapped < 1 || p.x_ungapped > 16384 || p.x_gapped < 0 || p.gap_open < 0 ||
        p.gap_extend < 1 || p.reap_score < 0) {
        ctx->err = "invalid parameters";
        return BLZ_EINVAL;
    }
    for (int i = 0; i < qlen; i++)
        if (query[i] >= kAlpha) {
            ctx->err = "query residue outside NCBIstdaa (>= 28)";
            return BLZ_EINVAL;
        }
    (void)hipSetDevice(ctx->device);
    blz_query *qq = new blz_query();
    qq->ctx = ctx;
    qq->uid = g_next_uid++;
    qq->qlen = qlen;
    qq->params = p;
    qq->letters.assign(query, query + qlen);
    const blz_status st = build_tables(ctx, query, qlen, {{0, qlen}}, p.word_threshold, qq->t);
    if (st != BLZ_OK) {
        delete qq;
        return st;
    }
    *out = qq;
    return BLZ_OK;
}

void blz_query_free(blz_query *q) {
    if (!q) return;
    if (q->ctx) {
        q->ctx->drop_graph();  // a captured batch may read this query's tables
        // a cached pack may hold this query; drop the cache
        for (uint64_t k : q->ctx->pack_key)
            if (k == q->uid) {
                for (auto &pk : q->ctx->packs) pk.t.release();
                q->ctx->packs.clear();
                q->ctx->pack_key.clear();
                break;
            }
    }
    q->t.release();
    delete q;
}

}  // extern "C"

QueryDev blz::query_dev(const blz_query *q) {
    QueryDev d{};
    d.query = q->t.d_query;
    d.prefix = q->t.d_prefix;
    d.entries = q->t.d_entries;
    d.rec = q->t.d_rec;
    d.rel16 = q->t.d_rel16;
    d.blockbase = q->t.d_blockbase;
    d.n_entries = (uint32_t)q->t.n_entries;
    d.qlen = q->qlen;
    const uint32_t ring = ring_words_for(q->qlen, q->params.window);
    d.ring_mask = ring - 1;
    set_ring(d, q->qlen, q->params.window);
    d.claim_mask = std::min<uint32_t>(ring, 128) - 1;
    d.window = q->params.window;
    d.x_ungapped = q->params.x_ungapped;
    d.ungapped_cutoff = q->params.ungapped_cutoff;
    d.x_gapped = q->params.x_gapped;
    d.gap_open = q->params.gap_open;
    d.gap_extend = q->params.gap_extend;
    d.gapped_cutoff = q->params.gapped_cutoff;
    d.reap_score = q->params.reap_score;
    d.n_members = 1;
    return d;
}

namespace {

// Queries may share a seed scan when the seed stage treats them alike; the gapped stage reads each
// query's own parameters.
bool same_params(const blz_params &a, const blz_params &b) {
    return a.word_threshold == b.word_threshold && a.window == b.window && a.x_ungapped == b.x_ungapped &&
           a.ungapped_cutoff == b.ungapped_cutoff;
}

// Concatenated letters of a pack: at most what keeps the diagonal ring at 1,024 words.
int pack_len(int window) { return std::max(64, 1024 - 128 - window); }

// Groups the batch's short queries (by length) into packs; a query longer than half a pack
// forms its own.  Packs of several queries get merged lookup tables over the concatenation,
// with window + 3 letter-0 separators so no two-hit pair or X-drop walk crosses queries.
blz_status make_packs(blz_ctx *ctx, const blz_query *const *queries, int nq, std::vector<Pack> &packs) {
    std::vector<int> idx(nq);
    for (int i = 0; i < nq; i++) idx[i] = i;
    std::stable_sort(idx.begin(), idx.end(), [&](int x, int y) { return queries[x]->qlen < queries[y]->qlen; });
    packs.clear();
    const bool packing = nq > 1 && ctx->tune.query_packing != 1;
    for (size_t k = 0; k < idx.size();) {
        const blz_query *q0 = queries[idx[k]];
        Pack pk;
        const int gap = q0->params.window + 3;
        const int max_len = ctx->tune.pack_len ? ctx->tune.pack_len : pack_len(q0->params.window);
        int len = 0;
        while (k < idx.size()) {
            const blz_query *q = queries[idx[k]];
            const int add = (pk.members.empty() ? 0 : gap) + q->qlen;
            if (!pk.members.empty() &&
                (!packing || q->qlen > max_len / 2 || len + add > max_len || (int)pk.members.size() >= kMaxMembers ||
                 !same_params(q->params, q0->params)))
                break;
            pk.starts.push_back((uint32_t)(len + (pk.members.empty() ? 0 : gap)));
            pk.members.push_back(idx[k]);
            len += add;
            k++;
            if (!packing || q->qlen > max_len / 2) break;
        }
        if (pk.members.size() > 1) {
            std::vector<uint8_t> cat((size_t)len, 0);
            std::vector<std::pair<int, int>> segs;
            for (size_t m = 0; m < pk.members.size(); m++) {
                const blz_query *q = queries[pk.members[m]];
                std::memcpy(cat.data() + pk.starts[m], q->letters.data(), q->qlen);
                segs.push_back({(int)pk.starts[m], (int)pk.starts[m] + q->qlen});
            }
            const blz_status st = build_tables(ctx, cat.data(), len, segs, q0->params.word_threshold, pk.t);
            if (st != BLZ_OK) {
                for (auto &p : packs) p.t.release();
                packs.clear();
                return st;
            }
        }
        pk.starts.push_back((uint32_t)len);
        packs.push_back(std::move(pk));
    }
    return BLZ_OK;
}

// Launch shape of the seed kernel for a set of queries (sizes are maxima).  table: where the
// lookup tables live (2: prefix and entries in LDS, 1: entries in LDS, 0: global memory).
struct SeedPlan {
    SeedArgs sa{};
    int table = 0;
    int waves = 0;
    int blocks = 0;
    int first_pack = 0;     // index of the launch's first pack in the query table
    size_t item_off = 0;    // its work items in the run's item table (three values each)
    bool long_subjects = false;  // the launch over the database's long subjects
};

// Work items of a seed launch over the packs of `qlens` (letters per pack): (pack, subject range),
// pack-major, each pack's ranges from the database's end to its start (a length-sorted database
// ends with its longest subjects, which go first).  Guided sizes: a range holds about 1/(1.5 blocks)
// of the work not yet handed out (work = residues x (pack letters + 200)), and at least
// R / (64 blocks) residues, so the launch starts on large items and drains on small ones (equal
// slices of eight per workgroup left the last workgroups 1.2 ms behind the median at the scan
// workload).  Appends three values per item to out.
void seed_items(const blz_db *db, const std::vector<int> &qlens, int blocks, std::vector<int64_t> &out) {
    const int64_t ns = db->nsubj;
    const uint64_t R = db->nres;
    if (ns == 0) return;
    const double share = 1.0 / (1.5 * std::max(1, blocks));
    const uint64_t min_res = std::max<uint64_t>(4096, R / (64 * (uint64_t)std::max(1, blocks)));
    double later = 0;  // work of the packs after the current one
    for (int l : qlens) later += (double)R * (l + 200);
    for (size_t p = 0; p < qlens.size(); p++) {
        const double w = qlens[p] + 200.0;
        later -= (double)R * w;
        int64_t hi = ns;
        while (hi > 0) {
            const uint64_t left = db->h_offs[hi];  // residues of this pack not yet handed out
            const uint64_t size = std::max<uint64_t>(min_res, (uint64_t)(((double)left * w + later) * share / w));
            const uint64_t target = left > size ? left - size : 0;
            // first subject at or past target, searched backwards from hi (galloping: the range
            // is near hi, and a plain binary search over the whole offset table costs 0.5 ms)
            int64_t step = 1;
            while (step < hi && db->h_offs[hi - step] >= target) step *= 2;
            const int64_t from = step < hi ? hi - step : 0;
            int64_t lo = (int64_t)(std::lower_bound(db->h_offs.begin() + from, db->h_offs.begin() + hi, target) -
                                   db->h_offs.begin());
            if (lo >= hi) lo = hi - 1;  // at least one subject
            out.push_back((int64_t)p);
            out.push_back(lo);
            out.push_back(hi);
            hi = lo;
        }
    }
}

bool plan_seed2(blz_ctx *ctx, const std::vector<QueryDev> &qs, const std::vector<int> &idx, int table,
                SeedPlan &pl, bool long_subjects = false) {
    SeedArgs &sa = pl.sa;
    for (int i : idx) {
        sa.qlen = std::max(sa.qlen, qs[i].qlen);
        sa.ring_mask = std::max(sa.ring_mask, qs[i].ring_mask);
        sa.ring_len = std::max(sa.ring_len, qs[i].ring_len);
        sa.claim_mask = std::max(sa.claim_mask, qs[i].claim_mask);
        sa.n_entries = std::max(sa.n_entries, qs[i].n_entries);
    }
    const size_t cap = ctx->lds_per_block;
    // most resident waves per CU (registers and LDS); more waves per workgroup on ties (fewer
    // workgroups stage the query and wait at item boundaries).  A tuned wave count that does not
    // fit falls back to this choice.
    int waves = 0, per_cu = 0;
    for (int pass = ctx->tune.seed_waves ? 0 : 1; pass < 2 && !waves; pass++)
        for (int w = 1; w <= 16; w++) {
            if (pass == 0 && w != ctx->tune.seed_waves) continue;
            if (seed2_lds_bytes(sa, table, w) > cap) break;
            const int n = seed2_blocks_per_cu(sa, table, w, long_subjects);
            if (n > 0 && w * n >= waves * per_cu) {
                waves = w;
                per_cu = n;
            }
        }
    if (!waves) return false;
    if (ctx->tune.seed_blocks_per_cu) per_cu = ctx->tune.seed_blocks_per_cu;
    pl.table = table;
    pl.waves = waves;
    pl.blocks = std::max(1, ctx->num_cus * per_cu);
    pl.long_subjects = long_subjects;
    return true;
}

// Table placement for one pack: the tuned one, else the one with the most resident waves.
int pick_table(blz_ctx *ctx, const std::vector<QueryDev> &qs, int i, bool compact_ok) {
    const int want = ctx->tune.seed_table;  // 0 auto, else table + 1
    if (want) {
        if (want - 1 == 2 && !compact_ok) return 1;
        return want - 1;
    }
    // built-in choice: the placement with the most resident waves per CU (the kernel is bound by
    // LDS latency, which waves hide); entries in LDS only when that costs no waves
    SeedPlan g, l;
    if (!plan_seed2(ctx, qs, {i}, 0, g)) return 0;
    if (plan_seed2(ctx, qs, {i}, 1, l) && l.waves * l.blocks >= g.waves * g.blocks) return 1;
    return 0;
}

// Host-side phase times of a run on stderr when BLZ_LOG_TIME is set (tuning by hand).
bool log_time() {
    static const bool on = std::getenv("BLZ_LOG_TIME") != nullptr;
    return on;
}
double now_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// Query qi's lists of the last run into a sink; ECAPACITY (sizes still reported) if it is too small.
blz_status copy_result(const blz_ctx *ctx, int qi, blz_hsp_sink *sink) {
    const size_t r0 = ctx->last_first[qi], r1 = ctx->last_first[qi + 1];
    sink->n_subjects = (int64_t)(r1 - r0);
    sink->n_hsps = ctx->last_nh[qi];
    if (sink->n_subjects > sink->subject_cap || sink->n_hsps > sink->hsp_cap) return BLZ_ECAPACITY;
    int64_t o = 0;
    for (size_t i = r0; i < r1; i++) {
        const ResultItem &r = ctx->last_res[i];
        sink->oids[i - r0] = r.sid;
        sink->counts[i - r0] = r.n;
        std::memcpy(sink->hsps + o, ctx->last_hs.data() + r.off, sizeof(Hsp) * r.n);
        o += r.n;
    }
    return BLZ_OK;
}

blz_status plan_and_launch(blz_ctx *ctx, const blz_query *const *queries, int nq, const blz_db *db, int variant,
                           bool use_graph, const std::vector<uint64_t> &key, int &n_launches);

// The captured stages, bracketed by the stage events.
blz_status replay_graph(blz_ctx *ctx) {
    hipStream_t st = ctx->stream;
    return check(ctx, hipEventRecord(ctx->ev[0], st), "event") &&
                   check(ctx, hipGraphLaunch(ctx->graph_exec[0], st), "graph launch") &&
                   check(ctx, hipEventRecord(ctx->ev[1], st), "event") &&
                   check(ctx, hipGraphLaunch(ctx->graph_exec[1], st), "graph launch") &&
                   check(ctx, hipEventRecord(ctx->ev[2], st), "event")
               ? BLZ_OK
               : BLZ_EDEVICE;
}

int collect(blz_ctx *ctx, int nq, const blz_db *db, blz_hsp_sink *sinks, blz_run_stats *stats, int n_launches,
            int reruns);
constexpr int kMaxReruns = 3;
// Internal outcome of collect: a work buffer of the run filled up; the caller runs it again.
constexpr int kRunAgain = -1;

// Prelim stage of queries[0..nq) over the database: seed launches (one per bin), the gapped
// stage, results copied into the sinks.  The launch sequence is captured as two hipGraphs (one
// per stage) and replayed while the same queries and database come again.
blz_status run_queries(blz_ctx *ctx, const blz_query *const *queries, int nq, const blz_db *db,
                       blz_hsp_sink *sinks, blz_run_stats *stats) {
    for (int i = 0; i < nq; i++)
        if (!queries[i] || queries[i]->ctx->device != ctx->device) return BLZ_EINVAL;
    const int variant = ctx->tune.seed_variant ? ctx->tune.seed_variant : 2;
    if (variant == 1 && db->max_len > kSeed1MaxSubjectLen) {
        ctx->err = "seed variant 1 takes subjects of at most 65,535 letters";
        return BLZ_EINVAL;
    }
    const size_t nsubj = (size_t)db->nsubj;
    hipStream_t st = ctx->stream;
    DevCounters *ctr = ctx->ctr;  // one counter block for the run; one item counter per seed launch
    const bool use_graph = variant != 1 && ctx->tune.graph != 1 && nsubj > 0;
    std::vector<uint64_t> key;
    if (use_graph) {
        for (int i = 0; i < nq; i++) key.push_back(queries[i]->uid);
        key.push_back(db->uid);
    }
    for (int reruns = 0;; reruns++) {
        int n_launches = 0;
        if (use_graph && ctx->graph_exec[0] && key == ctx->graph_key && ctx->graph_gen == ctx->gen &&
            std::memcmp(&ctx->graph_tune, &ctx->tune, sizeof(blz_tuning)) == 0) {
            // the same queries again: the tables on the device are the ones the graphs were captured with
            n_launches = ctx->graph_launches;
            const blz_status rc = replay_graph(ctx);
            if (rc != BLZ_OK) return rc;
        } else {
            const double t0 = log_time() ? now_ms() : 0;
            const blz_status rc = plan_and_launch(ctx, queries, nq, db, variant, use_graph, key, n_launches);
            if (rc != BLZ_OK) return rc;
            if (log_time()) std::fprintf(stderr, "[blz] plan + enqueue %.3f ms\n", now_ms() - t0);
        }
        const int rc = collect(ctx, nq, db, sinks, stats, n_launches, reruns);
        if (rc != kRunAgain) return (blz_status)rc;
        // a work buffer filled up: its size is raised to the run's count; the captured sequence
        // holds the old caps
        ctx->drop_graph();
        ctx->gen++;
        if (reruns + 1 >= kMaxReruns) {
            ctx->err = "work buffers still full after growing them";
            return BLZ_ENOMEM;
        }
    }
}

// Plans the run (packs, bins, launch shapes, work items), uploads the run's tables and enqueues the
// launch sequence; with use_graph the sequence is captured into ctx->graph first.
blz_status plan_and_launch(blz_ctx *ctx, const blz_query *const *queries, int nq, const blz_db *db, int variant,
                           bool use_graph, const std::vector<uint64_t> &key, int &n_launches) {
    ctx->drop_graph();
    const size_t nsubj = (size_t)db->nsubj;
    hipStream_t st = ctx->stream;
    DevCounters *ctr = ctx->ctr;
    // query table: [0, nq) the caller's queries (gapped stage), then one entry per pack (seed)
    std::vector<QueryDev> qs(nq);
    for (int i = 0; i < nq; i++) qs[i] = query_dev(queries[i]);

    std::vector<SeedPlan> plans;
    std::vector<uint32_t> meta;  // per pack: member starts (n+1) then member indices (n)
    std::vector<size_t> meta_off;
    int max_waves = 1;
    size_t max_blocks = 1;
    // Every query is also a pack of one of its own: the retry scan and the overflow queue scan
    // (query, subject) pairs with the query's own tables and emit through its member table.
    // Appended after the packs' entries; the table is uploaded with them.
    auto own_packs = [&]() -> bool {
        std::vector<size_t> own_off(nq);
        for (int i = 0; i < nq; i++) {
            own_off[i] = meta.size();
            meta.push_back(0);
            meta.push_back((uint32_t)qs[i].qlen);
            meta.push_back((uint32_t)i);
        }
        if (!ensure(ctx, ctx->pack_meta, ctx->pack_meta_cap, std::max<size_t>(meta.size(), 1), "pack table"))
            return false;
        for (int i = 0; i < nq; i++) {
            qs[i].member_start = ctx->pack_meta + own_off[i];
            qs[i].member_idx = qs[i].member_start + 2;
        }
        return true;
    };
    if (variant == 1) {
        if (nq != 1) {
            ctx->err = "seed variant 1 runs one query at a time";
            return BLZ_EINVAL;
        }
        if (!own_packs()) return BLZ_ENOMEM;
        SeedPlan pl;
        const uint32_t ring = qs[0].ring_mask + 1;
        int waves = ctx->tune.seed_waves ? std::min(ctx->tune.seed_waves, kSeedWavesMax) : kSeedWavesMax;
        while (waves > 1 && seed_lds_bytes(ring, waves) > ctx->lds_per_block) waves--;
        if (seed_lds_bytes(ring, waves) > ctx->lds_per_block) {
            ctx->err = "query too long for the diagonal ring in LDS";
            return BLZ_EINVAL;
        }
        pl.waves = waves;
        pl.blocks = std::max(1, ctx->num_cus * 8 / waves);
        plans.push_back(pl);
    } else {
        std::vector<uint64_t> key(nq);
        for (int i = 0; i < nq; i++) key[i] = queries[i]->uid;
        if (key != ctx->pack_key) {
            for (auto &pk : ctx->packs) pk.t.release();
            ctx->packs.clear();
            ctx->pack_key.clear();
            const blz_status st = make_packs(ctx, queries, nq, ctx->packs);
            if (st != BLZ_OK) return st;
            ctx->pack_key = key;
        }
        const std::vector<Pack> &packs = ctx->packs;
        std::vector<QueryDev> pq(packs.size());
        std::vector<bool> compact(packs.size());
        for (size_t k = 0; k < packs.size(); k++) {
            const Pack &pk = packs[k];
            const blz_query *q0 = queries[pk.members[0]];
            QueryDev d = query_dev(q0);
            const DevTables &t = pk.members.size() > 1 ? pk.t : q0->t;
            d.query = t.d_query;
            d.prefix = t.d_prefix;
            d.entries = t.d_entries;
            d.rec = t.d_rec;
            d.rel16 = t.d_rel16;
            d.blockbase = t.d_blockbase;
            d.n_entries = (uint32_t)t.n_entries;
            d.qlen = t.len;
            const uint32_t ring = ring_words_for(t.len, q0->params.window);
            d.ring_mask = ring - 1;
            set_ring(d, t.len, q0->params.window);
            if (d.ring_len > kRingLenMax) {
                ctx->err = "query too long for the seed kernel's diagonal ring";
                return BLZ_EINVAL;
            }
            d.claim_mask = std::min<uint32_t>(ring, 128) - 1;
            d.n_members = (int32_t)pk.members.size();
            // short subjects share a scan while the virtual positions stay within the 14-bit
            // ring entries (no word spanning two subjects is looked up, so the letter-0
            // separators never hit)
            d.group_gap = ctx->tune.subject_groups != 1 && q0->params.window <= 1024 ? q0->params.window : 0;
            meta_off.push_back(meta.size());
            for (uint32_t v : pk.starts) meta.push_back(v);
            for (int m : pk.members) meta.push_back((uint32_t)m);
            pq[k] = d;
            compact[k] = t.compact_ok;
        }
        // bins: table placement x query size (diagonal ring length: up to E, 2E, 4E, beyond)
        const uint32_t E = ctx->tune.bin_ring ? (uint32_t)ctx->tune.bin_ring : 1152u;
        std::vector<int> bins[3 * 4];
        for (size_t k = 0; k < packs.size(); k++) {
            const uint32_t r = pq[k].ring_len;
            const int rc = r <= E ? 0 : r <= 2 * E ? 1 : r <= 4 * E ? 2 : 3;
            bins[pick_table(ctx, pq, (int)k, compact[k]) * 4 + rc].push_back((int)k);
        }
        // each bin's packs are made contiguous after the queries; the plan records its first
        std::vector<int> order;
        for (int bin = 3 * 4 - 1; bin >= 0; bin--) {
            const int table = bin / 4;
            const std::vector<int> &idx = bins[bin];
            if (idx.empty()) continue;
            SeedPlan pl;
            if (!plan_seed2(ctx, pq, idx, table, pl)) {
                ctx->err = "query too long for the diagonal ring in LDS";
                return BLZ_EINVAL;
            }
            pl.sa.nq = (int32_t)idx.size();
            pl.first_pack = nq + (int)order.size();
            for (int k : idx) order.push_back(k);
            plans.push_back(pl);
            if (db->n_long > 0) {  // the same packs over the long subjects, tables in global memory
                SeedPlan lp;
                if (!plan_seed2(ctx, pq, idx, 0, lp, true)) {
                    ctx->err = "query too long for the diagonal ring in LDS";
                    return BLZ_EINVAL;
                }
                lp.sa.nq = pl.sa.nq;
                lp.first_pack = pl.first_pack;
                const int64_t groups = (db->n_long + lp.waves - 1) / lp.waves * lp.sa.nq;
                lp.blocks = (int)std::max<int64_t>(1, std::min<int64_t>(lp.blocks, groups));
                plans.push_back(lp);
            }
        }
        if (!own_packs()) return BLZ_ENOMEM;
        for (int k : order) {
            QueryDev d = pq[k];
            d.member_start = ctx->pack_meta + meta_off[k];
            d.member_idx = d.member_start + d.n_members + 1;
            qs.push_back(d);
        }
    }
    std::vector<int64_t> items;  // the run's work items, three values each
    size_t long_ring = 0;
    for (auto &pl : plans) {
        max_waves = std::max(max_waves, pl.waves);
        max_blocks = std::max<size_t>(max_blocks, (size_t)pl.blocks * pl.waves);
        if (variant == 1) continue;
        pl.item_off = items.size();
        if (pl.long_subjects) {  // one subject per wave and item, the longest (last) first
            long_ring = std::max(long_ring, (size_t)pl.blocks * pl.waves * pl.sa.ring_len);
            for (int p = 0; p < pl.sa.nq; p++)
                for (int64_t k = ((db->n_long - 1) / pl.waves) * pl.waves; k >= 0; k -= pl.waves) {
                    items.push_back(p);
                    items.push_back(k);
                    items.push_back(std::min<int64_t>(k + pl.waves, db->n_long));
                }
        } else {
            std::vector<int> qlens;
            for (int p = 0; p < pl.sa.nq; p++) qlens.push_back(qs[pl.first_pack + p].qlen);
            seed_items(db, qlens, pl.blocks, items);
        }
        pl.sa.n_items = (uint32_t)((items.size() - pl.item_off) / 3);
        if (std::getenv("BLZ_LOG_PLAN"))  // launch shapes, for tuning by hand
            std::fprintf(stderr,
                         "[blz] seed launch: table %d%s, %d packs, qlen <= %d, ring %u, entries <= %u, %d waves x %d "
                         "blocks, %u items, lds %zu\n",
                         pl.table, pl.long_subjects ? " (long subjects)" : "", pl.sa.nq, pl.sa.qlen, pl.sa.ring_len,
                         pl.sa.n_entries, pl.waves, pl.blocks, pl.sa.n_items, seed2_lds_bytes(pl.sa, pl.table, pl.waves));
    }
    const size_t lds_lanes = (size_t)(ctx->tune.gapped_blocks ? ctx->tune.gapped_blocks : ctx->num_cus * 4) * 64;
    const bool seg16 = ctx->tune.gapped_lanes != 64;  // four subjects per wave
    const bool packed = ctx->tune.gapped_width == 16;  // 16-bit DP scores, two columns per lane
    const int heavy_blocks = ctx->num_cus * (seg16 ? gapped_seg_blocks_per_cu(packed) : gapped_wave_blocks_per_cu());
    const size_t saved_records =
        std::max(lds_lanes * kSavedCap, (size_t)heavy_blocks * (seg16 ? 4 : 1) * kSavedCapWave);
    const size_t wide_lanes = (size_t)kGappedBlocks * 64;
    const size_t per_q = std::max<size_t>(nsubj, 1);
    // the run's caps: built-in sizes (or a tuned limit), raised to what an earlier run of this
    // context needed when it filled a buffer
    const blz_tuning &tn = ctx->tune;
    auto cap_of = [](int64_t limit, size_t builtin, size_t grown) {
        return std::max<size_t>(limit > 0 ? (size_t)limit : builtin, grown);
    };
    const size_t work_cap = cap_of(tn.cap_work, std::min(per_q * (size_t)nq, kWorkCapStart), ctx->grow_work);
    const size_t pool_cap = cap_of(tn.cap_pool, std::max<size_t>(1 << 22, 2 * work_cap), ctx->grow_pool);
    const size_t hsp_cap = cap_of(tn.cap_hsps, std::max<size_t>(1 << 22, per_q), ctx->grow_hsps);
    const size_t ovf_cap = cap_of(tn.cap_ovf, kOvfStart, ctx->grow_ovf);
    const int stage_cap = tn.cap_stage > 0 ? std::min(tn.cap_stage, kStageCap) : kStageCap;
    const int retry_cap = tn.cap_retry > 0 ? std::min(tn.cap_retry, kRetryCap) : kRetryCap;
    int wide_ring = kDpRing;
    if (tn.cap_wide_ring > 0)
        while (wide_ring > 16 && wide_ring / 2 >= tn.cap_wide_ring) wide_ring /= 2;
    const int wide_saved = tn.cap_wide_saved > 0 ? std::min(tn.cap_wide_saved, kSavedCapWide) : kSavedCapWide;
    uint32_t max_ring_len = 8;  // the retry launch's rings: any query of the run
    for (int i = 0; i < nq; i++) max_ring_len = std::max(max_ring_len, qs[i].ring_len);
    if (!ensure(ctx, ctx->stage, ctx->stage_records, max_blocks * stage_cap, "stage") ||
        !ensure(ctx, ctx->pool, ctx->pool_cap, pool_cap, "pool") ||
        !ensure(ctx, ctx->work, ctx->work_cap, work_cap, "work") ||
        !ensure(ctx, ctx->order, ctx->order_cap, work_cap, "order") ||
        !ensure(ctx, ctx->redo, ctx->redo_cap, kRedoCap, "redo") ||
        !ensure(ctx, ctx->ovf, ctx->ovf_cap, ovf_cap, "ovf") ||
        !ensure(ctx, ctx->dp, ctx->dp_entries, wide_lanes * wide_ring, "dp") ||
        !ensure(ctx, ctx->saved, ctx->saved_records, saved_records, "saved") ||
        !ensure(ctx, ctx->saved_wide, ctx->saved_wide_records, wide_lanes * wide_saved, "saved wide") ||
        !ensure(ctx, ctx->hsps, ctx->hsp_cap, hsp_cap, "hsps") ||
        !ensure(ctx, ctx->results, ctx->results_cap, work_cap, "results") ||
        !ensure(ctx, ctx->qdev, ctx->qdev_cap, qs.size(), "query table") ||
        !ensure(ctx, ctx->item_ctr, ctx->item_ctr_cap, plans.size() + 1, "item counters") ||
        !ensure(ctx, ctx->items, ctx->items_cap, std::max<size_t>(items.size(), 1), "work items") ||
        (long_ring && !ensure(ctx, ctx->long_ring, ctx->long_ring_cap, long_ring, "long-subject rings")) ||
        (variant != 1 && (!ensure(ctx, ctx->retry_list, ctx->retry_list_cap, ovf_cap, "retry list") ||
                          !ensure(ctx, ctx->retry_stage, ctx->retry_stage_cap,
                                  (size_t)kRetryBlocks * kRetryWaves * retry_cap, "retry stage") ||
                          !ensure(ctx, ctx->retry_ring, ctx->retry_ring_cap,
                                  (size_t)kRetryBlocks * kRetryWaves * max_ring_len, "retry rings"))))
        return BLZ_ENOMEM;
    ctx->run_work_cap = work_cap;
    ctx->run_pool_cap = pool_cap;
    ctx->run_hsp_cap = hsp_cap;
    ctx->run_ovf_cap = ovf_cap;
    ctx->run_ring_max = max_ring_len;

    if (!check(ctx, hipMemcpyAsync(ctx->qdev, qs.data(), sizeof(QueryDev) * qs.size(), hipMemcpyHostToDevice, st),
               "query table") ||
        (!meta.empty() && !check(ctx, hipMemcpyAsync(ctx->pack_meta, meta.data(), sizeof(uint32_t) * meta.size(),
                                                      hipMemcpyHostToDevice, st), "pack table")) ||
        (!items.empty() && !check(ctx, hipMemcpyAsync(ctx->items, items.data(), sizeof(int64_t) * items.size(),
                                                       hipMemcpyHostToDevice, st), "work items")))
        return BLZ_EDEVICE;

    for (size_t pi = 0; pi < plans.size() && nsubj > 0; pi++) {
        SeedPlan &pl = plans[pi];
        SeedArgs &sa = pl.sa;
        sa.res = db->res;
        sa.offs = db->offs;
        sa.stage = ctx->stage;
        sa.stage_cap = stage_cap;
        sa.pool = ctx->pool;
        sa.pool_cap = pool_cap;
        sa.work = ctx->work;
        sa.work_cap = (uint32_t)work_cap;
        sa.ovf = ctx->ovf;
        sa.ovf_cap = (uint32_t)ovf_cap;
        sa.ctr = ctr;
        if (variant == 1) {
            const QueryDev &d = qs[0];
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
        } else {
            sa.qd = ctx->qdev + pl.first_pack;
            sa.s_begin = 0;
            sa.s_end = db->nsubj;
            sa.items = ctx->items + pl.item_off;
            sa.long_ring = ctx->long_ring;
            sa.long_ids = db->long_ids;
            sa.item_ctr = ctx->item_ctr + pi;
        }
    }
    GappedArgs ga{};
    ga.res = db->res;
    ga.offs = db->offs;
    ga.qd = ctx->qdev;
    ga.pool = ctx->pool;
    ga.work = ctx->work;
    ga.order = ctx->order;
    ga.count = &ctr->work_count;
    ga.count_cap = (uint32_t)work_cap;
    ga.next = &ctr->next_work;
    ga.next_heavy = &ctr->next_heavy;
    ga.heavy_classes = ctx->tune.gapped_wave_classes > 0 ? ctx->tune.gapped_wave_classes - 1 : kBuckets;
    ga.redo = ctx->redo;
    ga.redo_cap = (uint32_t)ctx->redo_cap;
    ga.dp_scratch = ctx->dp;
    ga.dp_wmask = wide_ring - 1;
    ga.hsp_scratch = ctx->saved;
    ga.saved_cap = kSavedCap;
    ga.hsps = ctx->hsps;
    ga.hsp_cap = hsp_cap;
    ga.results = ctx->results;
    ga.ovf = ctx->ovf;
    ga.ovf_cap = (uint32_t)ovf_cap;
    ga.ctr = ctr;
    GappedArgs gw = ga;  // wide pass over the redo list
    gw.order = ctx->redo;
    gw.count = &ctr->redo_count;
    gw.count_cap = (uint32_t)ctx->redo_cap;
    gw.next = &ctr->next_redo;
    gw.heavy_classes = 0;
    gw.redo = nullptr;
    gw.redo_cap = 0;
    gw.hsp_scratch = ctx->saved_wide;
    gw.saved_cap = wide_saved;
    const int lds_blocks = ctx->tune.gapped_blocks ? ctx->tune.gapped_blocks : ctx->num_cus * 4;
    GappedArgs gh = ga;  // wave pass: per-block saved list of kSavedCapWave
    gh.saved_cap = kSavedCapWave;
    const size_t n_item_ctr = plans.size() + 1;
    // the two stages' launch sequences
    auto enqueue_seed = [&]() -> bool {
        if (!check(ctx, launch_fill_u32(reinterpret_cast<uint32_t *>(ctr), sizeof(DevCounters) / 4, 0, st), "clear") ||
            !check(ctx, launch_fill_u32(ctx->item_ctr, n_item_ctr, 0, st), "clear"))
            return false;
        for (size_t pi = 0; pi < plans.size() && nsubj > 0; pi++) {
            const SeedPlan &pl = plans[pi];
            const hipError_t e = variant == 1 ? launch_seed(pl.sa, pl.blocks, pl.waves, st)
                                              : launch_seed2(pl.sa, pl.table, pl.blocks, pl.waves, pl.long_subjects, st);
            if (!check(ctx, e, "k_seed launch")) return false;
        }
        if (variant != 1 && nsubj > 0) {  // second chance for the subjects past a cap
            SeedArgs r{};
            r.res = db->res;
            r.offs = db->offs;
            r.qd = ctx->qdev;  // the run's own queries
            r.claim_mask = 127;
            r.ring_len = max_ring_len;
            r.stage = ctx->retry_stage;
            r.stage_cap = retry_cap;
            r.pool = ctx->pool;
            r.pool_cap = pool_cap;
            r.work = ctx->work;
            r.work_cap = (uint32_t)work_cap;
            r.ovf = ctx->ovf;
            r.ovf_cap = (uint32_t)ovf_cap;
            r.ctr = ctr;
            r.long_ring = ctx->retry_ring;
            if (!check(ctx, launch_seed_retry(r, ctx->retry_list, kRetryBlocks, kRetryWaves, st), "retry launch"))
                return false;
        }
        return true;
    };
    auto enqueue_gapped = [&]() -> bool {
        return (nsubj == 0 || check(ctx, launch_order(ctx->work, ctr, ctx->order, (uint32_t)work_cap, 512, st),
                                    "k_order launch")) &&
               (nsubj == 0 || ga.heavy_classes == 0 ||
                check(ctx, seg16 ? launch_gapped_seg(gh, heavy_blocks, packed, st) : launch_gapped_wave(gh, heavy_blocks, st),
                      "k_gapped_wave launch")) &&
               (nsubj == 0 || ga.heavy_classes >= kBuckets ||
                check(ctx, launch_gapped_lds(ga, lds_blocks, st), "k_gapped_lds launch")) &&
               (nsubj == 0 || check(ctx, launch_gapped_wide(gw, kGappedBlocks, 64, st), "k_gapped_wide launch"));
    };
    n_launches = (int)plans.size();
    if (!use_graph)
        return check(ctx, hipEventRecord(ctx->ev[0], st), "event") && enqueue_seed() &&
                       check(ctx, hipEventRecord(ctx->ev[1], st), "event") && enqueue_gapped() &&
                       check(ctx, hipEventRecord(ctx->ev[2], st), "event")
                   ? BLZ_OK
                   : BLZ_EDEVICE;
    // one graph per stage, so the stage events can sit between them on the stream
    for (int k = 0; k < 2; k++) {
        if (!check(ctx, hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed), "capture")) return BLZ_EDEVICE;
        const bool ok = k == 0 ? enqueue_seed() : enqueue_gapped();
        hipGraph_t g = nullptr;
        const hipError_t ec = hipStreamEndCapture(st, &g);
        if (!ok || !check(ctx, ec, "capture")) {
            if (g) (void)hipGraphDestroy(g);
            ctx->drop_graph();
            return BLZ_EDEVICE;
        }
        ctx->graph[k] = g;
        if (!check(ctx, hipGraphInstantiate(&ctx->graph_exec[k], g, nullptr, nullptr, 0), "graph instantiate")) {
            ctx->drop_graph();
            return BLZ_EDEVICE;
        }
    }
    ctx->graph_key = key;
    ctx->graph_tune = ctx->tune;
    ctx->graph_gen = ctx->gen;
    ctx->graph_launches = n_launches;
    return replay_graph(ctx);
}

// Result of the overflow queue: lists of the pairs it finished, its counters, what is left.
struct QueueOut {
    std::vector<ResultItem> res;  // HSP offsets into hs
    std::vector<Hsp> hs;
    unsigned long long init_hits = 0, gapped_ext = 0, dp_cells = 0, work_items = 0;
    std::vector<int64_t> left;    // pairs it could not size buffers for
    int rounds = 0;
};

// Overflow queue.  Pairs past a per-subject cap of the main passes (more init hits than the
// retry scan stages, a DP band wider than the wide pass's ring, more saved HSPs than it holds)
// are run again from their letters with buffers sized to them: the retry scan with a stage of
// kQueueStage0 init hits per pair (x8 per further round for the pairs that still do not fit, up to
// kQueueWaves waves within kQueueBytes of staging),
// then the wide gapped pass with a DP ring as wide as the longest of their subjects (a band
// never has more than subject length + 1 cells) and room for every init hit as a saved HSP.
// BLAZE flags such subjects for re-evaluation on the CPU (README.md:36-37; flag 255 at
// gpuBLAZE.cu:1088,1243-1247); here they stay on the device.  The run's query table (ctx->qdev)
// is still in place, so a pair's query index selects its tables.
blz_status run_overflow_queue(blz_ctx *ctx, const blz_db *db, std::vector<int64_t> pairs, QueueOut &out) {
    hipStream_t st = ctx->stream;
    size_t stage_cap = kQueueStage0;
    size_t pool_cap = kQueueStage0;
    if (!ensure(ctx, ctx->q_ctr, ctx->q_ctr_cap, 1, "queue counters", false)) return BLZ_ENOMEM;
    for (int round = 0; round < kQueueRounds && !pairs.empty(); round++) {
        out.rounds++;
        const size_t n = pairs.size();
        // one wave per pair at a time, as many waves as the staging budget allows
        size_t waves = std::min<size_t>({n, kQueueWaves, std::max<size_t>(1, kQueueBytes / (stage_cap * sizeof(InitHit)))});
        const int wpb = (int)std::min<size_t>(waves, 4);
        const int blocks = (int)((waves + wpb - 1) / wpb);
        waves = (size_t)blocks * wpb;
        if (!ensure(ctx, ctx->q_list, ctx->q_list_cap, n, "queue list", false) ||
            !ensure(ctx, ctx->q_stage, ctx->q_stage_cap, waves * stage_cap, "queue stage", false) ||
            !ensure(ctx, ctx->q_ring, ctx->q_ring_cap, waves * ctx->run_ring_max, "queue rings", false) ||
            !ensure(ctx, ctx->q_pool, ctx->q_pool_cap, pool_cap, "queue pool", false) ||
            !ensure(ctx, ctx->q_work, ctx->q_work_cap, n, "queue work", false) ||
            !ensure(ctx, ctx->q_order, ctx->q_order_cap, n, "queue order", false) ||
            !ensure(ctx, ctx->q_ovf, ctx->q_ovf_cap, n, "queue overflow", false) ||
            !ensure(ctx, ctx->q_results, ctx->q_results_cap, n, "queue results", false))
            return BLZ_ENOMEM;
        DevCounters c0{};
        c0.retry_count = (unsigned)n;
        if (!check(ctx, hipMemcpyAsync(ctx->q_list, pairs.data(), sizeof(int64_t) * n, hipMemcpyHostToDevice, st),
                   "queue list") ||
            !check(ctx, hipMemcpyAsync(ctx->q_ctr, &c0, sizeof(c0), hipMemcpyHostToDevice, st), "queue counters"))
            return BLZ_EDEVICE;
        SeedArgs r{};
        r.res = db->res;
        r.offs = db->offs;
        r.qd = ctx->qdev;
        r.claim_mask = 127;
        r.ring_len = ctx->run_ring_max;
        r.stage = ctx->q_stage;
        r.stage_cap = (int32_t)stage_cap;
        r.pool = ctx->q_pool;
        r.pool_cap = pool_cap;
        r.work = ctx->q_work;
        r.work_cap = (uint32_t)n;
        r.ovf = ctx->q_ovf;
        r.ovf_cap = (uint32_t)n;
        r.ctr = ctx->q_ctr;
        r.long_ring = ctx->q_ring;
        r.retry = ctx->q_list;
        DevCounters c{};
        if (!check(ctx, launch_seed_pairs(r, blocks, wpb, st), "queue scan") ||
            !check(ctx, launch_order(ctx->q_work, ctx->q_ctr, ctx->q_order, (uint32_t)n, 64, st), "queue order") ||
            !check(ctx, hipMemcpyAsync(&c, ctx->q_ctr, sizeof(c), hipMemcpyDeviceToHost, st), "queue counters") ||
            !check(ctx, hipStreamSynchronize(st), "queue sync"))
            return BLZ_EDEVICE;
        if (c.pool_used > pool_cap) {  // the round's init hits did not fit: the same round again
            pool_cap = (size_t)c.pool_used + c.pool_used / 4;
            round--;
            if (out.rounds > 2 * kQueueRounds) break;
            continue;
        }
        const size_t nw = c.work_count;
        // sizes of the gapped pass: the longest subject's band, every init hit saved
        std::vector<WorkItem> wk(nw);
        if (nw && !check(ctx, hipMemcpy(wk.data(), ctx->q_work, sizeof(WorkItem) * nw, hipMemcpyDeviceToHost),
                         "queue work"))
            return BLZ_EDEVICE;
        size_t max_n = 1, sum_n = 0;
        uint64_t max_len = 1;
        for (const WorkItem &w : wk) {
            max_n = std::max<size_t>(max_n, (size_t)w.n);
            sum_n += (size_t)w.n;
            max_len = std::max<uint64_t>(max_len, db->h_offs[w.sid + 1] - db->h_offs[w.sid]);
        }
        size_t ring = 16;
        while (ring < max_len + 2) ring <<= 1;
        const size_t per_lane = ring * sizeof(int2) + max_n * sizeof(Hsp);
        size_t lanes = std::min<size_t>({std::max<size_t>(nw, 1), (size_t)256, std::max<size_t>(1, kQueueBytes / per_lane)});
        lanes = (lanes + 63) / 64 * 64;
        if (!ensure(ctx, ctx->q_dp, ctx->q_dp_cap, lanes * ring, "queue DP rings", false) ||
            !ensure(ctx, ctx->q_saved, ctx->q_saved_cap, lanes * max_n, "queue saved HSPs", false) ||
            !ensure(ctx, ctx->q_hsps, ctx->q_hsps_cap, std::max<size_t>(sum_n, 1), "queue HSPs", false))
            return BLZ_ENOMEM;
        GappedArgs g{};
        g.res = db->res;
        g.offs = db->offs;
        g.qd = ctx->qdev;
        g.pool = ctx->q_pool;
        g.work = ctx->q_work;
        g.order = ctx->q_order;
        g.count = &ctx->q_ctr->work_count;
        g.count_cap = (uint32_t)n;
        g.next = &ctx->q_ctr->next_redo;
        g.next_heavy = &ctx->q_ctr->next_heavy;
        g.heavy_classes = 0;
        g.dp_scratch = ctx->q_dp;
        g.dp_wmask = (int32_t)ring - 1;
        g.hsp_scratch = ctx->q_saved;
        g.saved_cap = (int32_t)max_n;
        g.hsps = ctx->q_hsps;
        g.hsp_cap = std::max<size_t>(sum_n, 1);
        g.results = ctx->q_results;
        g.ovf = ctx->q_ovf;
        g.ovf_cap = (uint32_t)n;
        g.ctr = ctx->q_ctr;
        if (nw && (!check(ctx, launch_gapped_wide(g, (int)(lanes / 64), 64, st), "queue gapped") ||
                   !check(ctx, hipMemcpyAsync(&c, ctx->q_ctr, sizeof(c), hipMemcpyDeviceToHost, st), "queue counters") ||
                   !check(ctx, hipStreamSynchronize(st), "queue sync")))
            return BLZ_EDEVICE;
        const size_t nr = c.result_count, nh = std::min<size_t>(c.hsp_used, g.hsp_cap);
        std::vector<ResultItem> rs(nr);
        std::vector<Hsp> hs(nh);
        if ((nr && !check(ctx, hipMemcpy(rs.data(), ctx->q_results, sizeof(ResultItem) * nr, hipMemcpyDeviceToHost),
                          "queue results")) ||
            (nh && !check(ctx, hipMemcpy(hs.data(), ctx->q_hsps, sizeof(Hsp) * nh, hipMemcpyDeviceToHost),
                          "queue HSPs")))
            return BLZ_EDEVICE;
        for (ResultItem &x : rs) {
            x.off += out.hs.size();
            out.res.push_back(x);
        }
        out.hs.insert(out.hs.end(), hs.begin(), hs.end());
        out.init_hits += c.init_hits;
        out.gapped_ext += c.gapped_ext;
        out.dp_cells += c.dp_cells;
        out.work_items += nw;
        // what is still over a cap goes round again with a larger stage
        const size_t nleft = std::min<size_t>(c.overflow_count, n);
        pairs.assign(nleft, 0);
        if (nleft && !check(ctx, hipMemcpy(pairs.data(), ctx->q_ovf, sizeof(int64_t) * nleft, hipMemcpyDeviceToHost),
                            "queue overflow"))
            return BLZ_EDEVICE;
        stage_cap *= 8;
        pool_cap = std::max(pool_cap, stage_cap);
    }
    out.left = std::move(pairs);
    return BLZ_OK;
}

// Counters and HSP lists of the run just enqueued, into the caller's sinks (caller order,
// subjects ascending).  Returns kRunAgain, having raised the context's buffer sizes to what the
// run counted, when a run-wide buffer (init-hit pool, work list, HSP pool, overflow list) filled
// up: the counters keep counting past a full buffer, so one more run fits.  Pairs past a
// per-subject cap go to the overflow queue (tuning.overflow_mode 0) or are handed back.
int collect(blz_ctx *ctx, int nq, const blz_db *db, blz_hsp_sink *sinks, blz_run_stats *stats, int n_launches,
            int reruns) {
    hipStream_t st = ctx->stream;
    DevCounters *ctr = ctx->ctr;

    DevCounters c{};
    const double t0 = log_time() ? now_ms() : 0;
    if (!check(ctx, hipMemcpyAsync(&c, ctr, sizeof(c), hipMemcpyDeviceToHost, st), "counters") ||
        !check(ctx, hipStreamSynchronize(st), "sync"))
        return BLZ_EDEVICE;
    const double t1 = log_time() ? now_ms() : 0;
    const bool hand_back = ctx->tune.overflow_mode == 1;
    {
        // overflow records: those the retry prep could not move, plus those past the list now
        const size_t lost = (size_t)c.ovf_lost + (c.overflow_count > ctx->run_ovf_cap ? c.overflow_count - ctx->run_ovf_cap : 0);
        const size_t ovf_need = std::max<size_t>((size_t)c.retry_count + c.ovf_lost, c.overflow_count);
        auto grow = [](size_t need) { return need + need / 4 + 1024; };
        bool again = false;
        if (lost > 0) {
            ctx->grow_ovf = std::max(ctx->grow_ovf, grow(ovf_need));
            again = true;
        }
        if (!hand_back) {
            if (c.pool_used > ctx->run_pool_cap) {
                ctx->grow_pool = std::max(ctx->grow_pool, grow((size_t)c.pool_used));
                again = true;
            }
            if (c.work_count > ctx->run_work_cap) {
                ctx->grow_work = std::max(ctx->grow_work, grow((size_t)c.work_count));
                again = true;
            }
            if (c.hsp_used > ctx->run_hsp_cap) {
                ctx->grow_hsps = std::max(ctx->grow_hsps, grow((size_t)c.hsp_used));
                again = true;
            }
        }
        if (again) return kRunAgain;
    }
    // results written before a pool filled up stay valid: each points below the pool's end
    const size_t n_res = c.result_count, n_hs = std::min<size_t>(c.hsp_used, ctx->run_hsp_cap);
    const size_t res_bytes = (sizeof(ResultItem) * n_res + 255) & ~size_t(255);
    const size_t need = res_bytes + sizeof(Hsp) * n_hs;
    if (need > ctx->pin_cap) {  // grow the pinned staging (rarely: it keeps its size)
        if (ctx->pin) (void)hipHostFree(ctx->pin);
        ctx->pin = nullptr;
        ctx->pin_cap = 0;
        const size_t cap = std::max<size_t>(need + need / 2, 1 << 20);
        if (!check(ctx, hipHostMalloc(&ctx->pin, cap, hipHostMallocDefault), "pinned staging")) return BLZ_ENOMEM;
        ctx->pin_cap = cap;
    }
    ResultItem *res_pin = static_cast<ResultItem *>(ctx->pin);
    Hsp *hs_pin = reinterpret_cast<Hsp *>(static_cast<uint8_t *>(ctx->pin) + res_bytes);
    if ((n_res && !check(ctx, hipMemcpyAsync(res_pin, ctx->results, sizeof(ResultItem) * n_res, hipMemcpyDeviceToHost,
                                             st), "results")) ||
        (n_hs && !check(ctx, hipMemcpyAsync(hs_pin, ctx->hsps, sizeof(Hsp) * n_hs, hipMemcpyDeviceToHost, st), "hsps")) ||
        !check(ctx, hipStreamSynchronize(st), "sync"))
        return BLZ_EDEVICE;
    const double t2 = log_time() ? now_ms() : 0;
    float ms_seed = 0, ms_gapped = 0;
    (void)hipEventElapsedTime(&ms_seed, ctx->ev[0], ctx->ev[1]);
    (void)hipEventElapsedTime(&ms_gapped, ctx->ev[1], ctx->ev[2]);

    // pairs past a per-subject cap (all of them are on the list: a run that lost some ran again)
    std::vector<int64_t> ovf_pairs(std::min<size_t>(c.overflow_count, ctx->run_ovf_cap));
    if (!ovf_pairs.empty() && !check(ctx, hipMemcpy(ovf_pairs.data(), ctx->ovf, sizeof(int64_t) * ovf_pairs.size(),
                                                    hipMemcpyDeviceToHost), "overflow list"))
        return BLZ_EDEVICE;
    QueueOut qo;
    double ms_queue = 0;
    if (!ovf_pairs.empty() && !hand_back) {
        const double q0 = now_ms();
        const blz_status qs = run_overflow_queue(ctx, db, ovf_pairs, qo);
        if (qs != BLZ_OK) return qs;
        ms_queue = now_ms() - q0;
    } else {
        qo.left = ovf_pairs;
    }

    // per query (in the caller's order), subjects ascending: the (query, subject) keys
    // query * nsubj + subject are distinct, so an LSD radix sort of key << 24 | item index gives
    // the order (a comparison sort when that does not fit 64 bits); then gather.  The overflow
    // queue's lists follow the run's, their HSP offsets moved past the run's HSPs.
    const size_t n_all = n_res + qo.res.size();
    auto item = [&](size_t i) -> ResultItem {
        if (i < n_res) return res_pin[i];
        ResultItem r = qo.res[i - n_res];
        r.off += n_hs;
        return r;
    };
    std::vector<ResultItem> res(n_all);
    {
        std::vector<uint64_t> key(n_all), tmp(n_all);
        uint64_t kmax = 0;
        for (size_t i = 0; i < n_all; i++) {
            const ResultItem r = item(i);
            key[i] = (uint64_t)(uint32_t)r.qi * (uint64_t)db->nsubj + (uint64_t)r.sid;
            kmax = std::max(kmax, key[i]);
        }
        const int kbits = kmax ? 64 - __builtin_clzll(kmax) : 1;
        if (n_all < (1u << 24) && kbits + 24 <= 64) {
            for (size_t i = 0; i < n_all; i++) key[i] = (key[i] << 24) | i;
            constexpr int kDigit = 11;
            std::vector<uint32_t> cnt(1u << kDigit);
            for (int sh = 24; sh < 24 + kbits; sh += kDigit) {
                std::fill(cnt.begin(), cnt.end(), 0u);
                for (size_t i = 0; i < n_all; i++) cnt[(key[i] >> sh) & ((1u << kDigit) - 1)]++;
                uint32_t acc = 0;
                for (auto &c0 : cnt) {
                    const uint32_t v = c0;
                    c0 = acc;
                    acc += v;
                }
                for (size_t i = 0; i < n_all; i++) tmp[cnt[(key[i] >> sh) & ((1u << kDigit) - 1)]++] = key[i];
                key.swap(tmp);
            }
            for (size_t i = 0; i < n_all; i++) res[i] = item(key[i] & 0xFFFFFFu);
        } else {
            std::vector<std::pair<uint64_t, uint32_t>> kv(n_all);
            for (size_t i = 0; i < n_all; i++) kv[i] = {key[i], (uint32_t)i};
            std::sort(kv.begin(), kv.end());
            for (size_t i = 0; i < n_all; i++) res[i] = item(kv[i].second);
        }
    }
    std::vector<Hsp> hs(hs_pin, hs_pin + n_hs);
    hs.insert(hs.end(), qo.hs.begin(), qo.hs.end());
    ctx->last_first.assign((size_t)nq + 1, 0);
    ctx->last_nh.assign((size_t)nq, 0);
    int64_t total = 0;
    size_t r0 = 0;
    for (int qcaller = 0; qcaller < nq; qcaller++) {
        size_t r1 = r0;
        int64_t nh = 0;
        while (r1 < res.size() && res[r1].qi == qcaller) nh += res[r1++].n;
        ctx->last_first[qcaller] = r0;
        ctx->last_nh[qcaller] = nh;
        total += nh;
        r0 = r1;
    }
    ctx->last_first[nq] = r0;
    ctx->last_res.swap(res);
    ctx->last_hs.swap(hs);
    ctx->last_oid_of = db->oid_of;
    ctx->last_ovf.clear();
    for (int64_t e : qo.left) ctx->last_ovf.push_back(e & ~(int64_t)kOvfGapped);
    blz_status rc = BLZ_OK;
    if (sinks)
        for (int qcaller = 0; qcaller < nq; qcaller++) {
            const blz_status r = copy_result(ctx, qcaller, &sinks[qcaller]);
            if (r != BLZ_OK && rc == BLZ_OK) rc = r;
        }
    if (stats) {
        stats->subjects += (uint64_t)db->nsubj * (uint64_t)nq;
        stats->residues += db->nres * (uint64_t)nq;
        stats->word_hits += c.word_hits;
        stats->ungapped_ext += c.ungapped_ext;
        stats->init_hits += c.init_hits + qo.init_hits;
        stats->gapped_subjects += std::min<uint64_t>(c.work_count, ctx->run_work_cap) + qo.work_items;
        stats->gapped_ext += c.gapped_ext + qo.gapped_ext;
        stats->dp_cells += c.dp_cells + qo.dp_cells;
        stats->hsps += (uint64_t)total;
        stats->overflow_subjects += ovf_pairs.size();
        stats->overflow_left += qo.left.size();
        stats->ms_seed += ms_seed;
        stats->ms_gapped += ms_gapped;
        stats->ms_total += ms_seed + ms_gapped;
        stats->seed_launches += (int32_t)n_launches;
        stats->gapped_launches += 1;
        stats->reruns += reruns;
        stats->drain_rounds += qo.rounds;
        stats->ms_drain += ms_queue;
    }
    if (!qo.left.empty() && rc == BLZ_OK) {
        ctx->err = "pairs handed back past the device caps: " + std::to_string(qo.left.size());
        rc = BLZ_EOVERFLOW;
    }
    if (log_time())
        std::fprintf(stderr, "[blz] wait for device %.3f ms, lists back %.3f ms, host lists %.3f ms\n", t1 - t0,
                     t2 - t1, now_ms() - t2);
    return rc;
}

}  // namespace

extern "C" {

blz_status blz_prelim_run(blz_ctx *ctx, const blz_query *q, const blz_db *db, blz_hsp_sink *sink,
                          blz_run_stats *stats) {
    if (!ctx || !q || !db || q->ctx->device != ctx->device || db->ctx->device != ctx->device) return BLZ_EINVAL;
    (void)hipSetDevice(ctx->device);
    return run_queries(ctx, &q, 1, db, sink, stats);
}

blz_status blz_prelim_run_batch(blz_ctx *ctx, const blz_query *const *queries, int32_t nq, const blz_db *db,
                                blz_hsp_sink *sinks, blz_run_stats *stats) {
    if (!ctx || !queries || nq < 0 || !db || db->ctx->device != ctx->device) return BLZ_EINVAL;
    if (nq == 0) return BLZ_OK;
    (void)hipSetDevice(ctx->device);
    if ((ctx->tune.seed_variant ? ctx->tune.seed_variant : 2) == 1) {
        blz_status worst = BLZ_OK;
        for (int32_t i = 0; i < nq; i++) {
            blz_status rc = run_queries(ctx, &queries[i], 1, db, sinks ? &sinks[i] : nullptr, stats);
            if (rc == BLZ_EDEVICE || rc == BLZ_ENOMEM || rc == BLZ_EINVAL) return rc;
            if (rc != BLZ_OK && worst == BLZ_OK) worst = rc;
        }
        return worst;
    }
    return run_queries(ctx, queries, nq, db, sinks, stats);
}

blz_status blz_result_sizes(const blz_ctx *ctx, int32_t qi, int64_t *n_subjects, int64_t *n_hsps) {
    if (!ctx || qi < 0 || (size_t)qi + 1 >= ctx->last_first.size()) return BLZ_EINVAL;
    if (n_subjects) *n_subjects = (int64_t)(ctx->last_first[qi + 1] - ctx->last_first[qi]);
    if (n_hsps) *n_hsps = ctx->last_nh[qi];
    return BLZ_OK;
}

blz_status blz_result_copy(const blz_ctx *ctx, int32_t qi, blz_hsp_sink *sink) {
    if (!ctx || !sink || qi < 0 || (size_t)qi + 1 >= ctx->last_first.size()) return BLZ_EINVAL;
    return copy_result(ctx, qi, sink);
}

blz_status blz_result_hitlist(const blz_ctx *ctx, int32_t qi, const blz_search_stats *s, blz_hsp_sink *sink,
                              double *best_evalue) {
    if (!ctx || !s || !sink || qi < 0 || (size_t)qi + 1 >= ctx->last_first.size() || s->hitlist_size < 1)
        return BLZ_EINVAL;
    // one entry per subject list: its best e-value is its first HSP's (lists are in score order)
    struct Entry {
        double ev;
        int32_t score;
        int64_t sid;
        size_t item;
    };
    const size_t r0 = ctx->last_first[qi], r1 = ctx->last_first[qi + 1];
    std::vector<Entry> e;
    e.reserve(r1 - r0);
    for (size_t i = r0; i < r1; i++) {
        const ResultItem &r = ctx->last_res[i];
        const int32_t best = ctx->last_hs[r.off].score;
        e.push_back({blz_evalue(s, best), best, r.sid, i});
    }
    // NCBI's list order (s_EvalueCompareHSPLists): best e-value ascending, e-values below 1e-180
    // counting as equal; then best score descending; then the higher ordinal id first (the
    // database's own id where the upload reordered it)
    const std::vector<int64_t> *oid_of = ctx->last_oid_of.get();
    auto before = [oid_of](const Entry &a, const Entry &b) {
        constexpr double kTiny = 1.0e-180;
        if (!(a.ev < kTiny && b.ev < kTiny) && a.ev != b.ev) return a.ev < b.ev;
        if (a.score != b.score) return a.score > b.score;
        return oid_of ? (*oid_of)[a.sid] > (*oid_of)[b.sid] : a.sid > b.sid;
    };
    const size_t k = std::min<size_t>(e.size(), (size_t)s->hitlist_size);
    std::partial_sort(e.begin(), e.begin() + k, e.end(), before);
    int64_t nh = 0;
    for (size_t i = 0; i < k; i++) nh += ctx->last_res[e[i].item].n;
    sink->n_subjects = (int64_t)k;
    sink->n_hsps = nh;
    if ((int64_t)k > sink->subject_cap || nh > sink->hsp_cap) return BLZ_ECAPACITY;
    int64_t o = 0;
    for (size_t i = 0; i < k; i++) {
        const ResultItem &r = ctx->last_res[e[i].item];
        sink->oids[i] = r.sid;
        sink->counts[i] = r.n;
        std::memcpy(sink->hsps + o, ctx->last_hs.data() + r.off, sizeof(Hsp) * r.n);
        o += r.n;
        if (best_evalue) best_evalue[i] = e[i].ev;
    }
    return BLZ_OK;
}

blz_status blz_overflow_drain(const blz_ctx *ctx, int64_t *pairs, int64_t cap, int64_t *n) {
    if (!ctx || !n || cap < 0 || (cap > 0 && !pairs)) return BLZ_EINVAL;
    *n = (int64_t)ctx->last_ovf.size();
    if (*n > cap) return BLZ_ECAPACITY;
    std::copy(ctx->last_ovf.begin(), ctx->last_ovf.end(), pairs);
    return BLZ_OK;
}

}  // extern "C"
