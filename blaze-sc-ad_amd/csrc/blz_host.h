// blz_host.h -- the host runtime's objects behind the C ABI's opaque handles (blz_ctx, blz_db,
// blz_query), shared by the source files that implement entry points.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "blz_common.h"
#include "blz_device.h"
#include "blz_prelim.h"

namespace blz {

// Device copy of a neighbourhood lookup table over a letter string.  Words may only start inside
// the segments [start, end) -- one segment per query of a pack, so no word spans two queries.
struct DevTables {
    uint8_t *d_query = nullptr;
    int32_t len = 0;
    uint32_t *d_prefix = nullptr;
    uint16_t *d_entries = nullptr;
    uint64_t n_entries = 0;
    uint16_t *d_rel16 = nullptr;
    uint32_t *d_blockbase = nullptr;
    uint4 *d_rec = nullptr;  // word records (kWordRecInline entries inline)
    bool compact_ok = false;
    hipStream_t st = nullptr;  // stream-ordered allocations: freed without a device-wide sync
    void release() {
        void *bufs[] = {d_query, d_prefix, d_entries, d_rel16, d_blockbase, d_rec};
        for (void *b : bufs)
            if (b) (void)(st ? hipFreeAsync(b, st) : hipFree(b));
        *this = DevTables{};
    }
};

// Device buffers and the last result of the traceback stage (blz_traceback.cpp).
struct TbState;
void tb_release(TbState *s);

// Several short queries scanned as one concatenation (their lookup tables merged); the
// gapped stage still runs per query.
struct Pack {
    std::vector<int> members;       // caller's query indices
    std::vector<uint32_t> starts;   // member m occupies [starts[m], starts[m] + len_m) of the concatenation
    DevTables t;                    // built for packs of two or more
};

}  // namespace blz

// the handles below are the C ABI's global types; their members use the engine's record types
using namespace blz;

struct blz_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    int num_cus = 0;
    size_t lds_per_block = 0;
    bool async_alloc = false;  // stream-ordered allocator available
    blz_tuning tune{};
    std::string err;
    // reusable device buffers
    DevCounters *ctr = nullptr;
    InitHit *stage = nullptr;
    size_t stage_records = 0;
    InitHit *pool = nullptr;
    size_t pool_cap = 0;
    WorkItem *work = nullptr;
    size_t work_cap = 0;
    int64_t *ovf = nullptr;
    size_t ovf_cap = 0;
    int2 *dp = nullptr;
    size_t dp_entries = 0;
    Hsp *saved = nullptr;
    size_t saved_records = 0;
    Hsp *saved_wide = nullptr;
    size_t saved_wide_records = 0;
    Hsp *hsps = nullptr;
    size_t hsp_cap = 0;
    ResultItem *results = nullptr;
    size_t results_cap = 0;
    QueryDev *qdev = nullptr;      // query table of the current run
    size_t qdev_cap = 0;
    uint32_t *item_ctr = nullptr;  // work-item counter per seed launch
    size_t item_ctr_cap = 0;
    uint32_t *order = nullptr;     // work list in cost order
    size_t order_cap = 0;
    uint32_t *redo = nullptr;      // work items for the wide gapped pass
    size_t redo_cap = 0;
    uint32_t *pack_meta = nullptr; // member starts and indices of the run's packs
    size_t pack_meta_cap = 0;
    void *pin = nullptr;           // pinned host staging for the lists coming back
    size_t pin_cap = 0;
    hipStream_t up_stream = nullptr;  // database uploads: side stream and pinned staging
    uint8_t *up_buf[2] = {nullptr, nullptr};
    hipEvent_t up_ev[2] = {nullptr, nullptr};
    int64_t *items = nullptr;      // work items of the run's seed launches (pack, first, end subject)
    size_t items_cap = 0;
    uint32_t *long_ring = nullptr; // seed rings for subjects the LDS ring does not take
    size_t long_ring_cap = 0;
    int64_t *retry_list = nullptr; // seed pairs past a cap, rescanned by the retry launch
    size_t retry_list_cap = 0;
    InitHit *retry_stage = nullptr;
    size_t retry_stage_cap = 0;
    uint32_t *retry_ring = nullptr;
    size_t retry_ring_cap = 0;
    // caps the kernels of the run just enqueued were given
    size_t run_work_cap = 0, run_pool_cap = 0, run_hsp_cap = 0, run_ovf_cap = 0;
    uint32_t run_ring_max = 8;  // longest diagonal ring of the run's queries (retry scans)
    // sizes later runs start from once a run filled a work buffer (then run again)
    size_t grow_work = 0, grow_pool = 0, grow_hsps = 0, grow_ovf = 0;
    // overflow queue: its own counters and buffers, sized to the pairs it runs
    DevCounters *q_ctr = nullptr;
    size_t q_ctr_cap = 0;
    int64_t *q_list = nullptr;
    size_t q_list_cap = 0;
    InitHit *q_stage = nullptr;
    size_t q_stage_cap = 0;
    uint32_t *q_ring = nullptr;
    size_t q_ring_cap = 0;
    InitHit *q_pool = nullptr;
    size_t q_pool_cap = 0;
    WorkItem *q_work = nullptr;
    size_t q_work_cap = 0;
    uint32_t *q_order = nullptr;
    size_t q_order_cap = 0;
    int64_t *q_ovf = nullptr;
    size_t q_ovf_cap = 0;
    int2 *q_dp = nullptr;
    size_t q_dp_cap = 0;
    Hsp *q_saved = nullptr;
    size_t q_saved_cap = 0;
    Hsp *q_hsps = nullptr;
    size_t q_hsps_cap = 0;
    ResultItem *q_results = nullptr;
    size_t q_results_cap = 0;
    // packs of the last batch (keyed by query ids), reused while the same queries come again
    std::vector<uint64_t> pack_key;
    std::vector<Pack> packs;
    uint64_t gen = 0;              // bumped whenever a work buffer is reallocated
    // lists of the last run, query-major and subjects ascending, for blz_result_*
    std::vector<ResultItem> last_res;
    std::vector<Hsp> last_hs;
    std::vector<size_t> last_first;      // [nq + 1] first result item of each query
    std::vector<int64_t> last_nh;        // [nq] HSPs of each query
    std::vector<int64_t> last_ovf;       // overflow pairs: query << 40 | subject
    std::shared_ptr<const std::vector<int64_t>> last_oid_of;  // the last run's database's id map
    TbState *tb = nullptr;               // traceback stage, created by its first call
    // launch sequence of the last batch, captured; valid while the same batch comes again and
    // no other run has rewritten the device tables in between
    hipGraph_t graph[2] = {nullptr, nullptr};          // seed stage, gapped stage
    hipGraphExec_t graph_exec[2] = {nullptr, nullptr};
    std::vector<uint64_t> graph_key;  // query ids, then the database id
    blz_tuning graph_tune{};
    uint64_t graph_gen = 0;
    int graph_launches = 0;
    void drop_graph() {
        for (int k = 0; k < 2; k++) {
            if (graph_exec[k]) (void)hipGraphExecDestroy(graph_exec[k]);
            if (graph[k]) (void)hipGraphDestroy(graph[k]);
            graph_exec[k] = nullptr;
            graph[k] = nullptr;
        }
        graph_key.clear();
    }
};

struct blz_db {
    blz_ctx *ctx = nullptr;
    uint64_t uid = 0;           // never reused, unlike addresses
    uint8_t *res = nullptr;     // device
    uint64_t *offs = nullptr;   // device
    int64_t nsubj = 0;
    uint64_t nres = 0;
    uint64_t max_len = 0;          // longest subject
    std::vector<uint64_t> h_offs;  // host copy of the offsets, for slicing by residues
    int64_t *long_ids = nullptr;   // device: subjects longer than the seed kernel's LDS ring takes
    int64_t n_long = 0;
    // ordinal id of each uploaded position when the upload reordered a database (a BLAST database
    // uploaded length-sorted); null: positions are the ids
    std::shared_ptr<const std::vector<int64_t>> oid_of;
};

struct blz_query {
    blz_ctx *ctx = nullptr;
    uint64_t uid = 0;           // never reused, unlike addresses
    int32_t qlen = 0;
    blz_params params{};
    std::vector<uint8_t> letters;
    DevTables t;
};

namespace blz {
// The kernels' view of a prepared query (lookup tables, ring geometry, search parameters).
QueryDev query_dev(const blz_query *q);
// Records a failed HIP call's message in the context; true on success.
bool check(blz_ctx *ctx, hipError_t e, const char *what);
// Per-device scratch of the allocation-free entry point, looked up once per device (blz_noalloc.hip).
bool noalloc_setup(int device);

}  // namespace blz

