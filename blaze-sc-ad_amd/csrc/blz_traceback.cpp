// blz_traceback.cpp -- host side of the traceback stage (the kernels are in blz_traceback.hip):
// blz_traceback, blz_traceback_sizes, blz_traceback_copy.
//
// A call uploads the pairs and their prelim HSPs, runs the main pass (one wave per pair, the DP
// band in LDS, per-wave traceback storage bounded), then the wide pass over the pairs the main
// pass gave back, with storage sized to the largest of them.  Edit-script runs are reserved from
// one buffer as HSPs are traced; a call that filled it runs again with the buffer grown to what it
// needed.  The result is compacted on the host: each pair's kept HSPs in score order, their runs
// contiguous.
#include <algorithm>
#include <chrono>
#include <cstring>
#include <vector>

#include "blz_host.h"

namespace blz {

struct TbState {
    TbCounters *ctr = nullptr;
    size_t ctr_cap = 0;
    TbQuery *qs = nullptr;
    size_t qs_cap = 0;
    TbPair *pairs = nullptr;
    size_t pairs_cap = 0;
    Hsp *pre = nullptr;
    size_t pre_cap = 0;
    TbHsp *out = nullptr;
    size_t out_cap = 0;
    int32_t *out_n = nullptr;
    size_t out_n_cap = 0;
    uint32_t *runs = nullptr;
    size_t runs_cap = 0;
    uint32_t *back = nullptr;
    size_t back_cap = 0;
    uint8_t *tb = nullptr;
    size_t tb_cap = 0;
    int4 *rows = nullptr;
    size_t rows_cap = 0;
    uint32_t *wruns = nullptr;
    size_t wruns_cap = 0;
    int2 *ring = nullptr;
    size_t ring_cap = 0;
    int blocks_main[2] = {0, 0}, blocks_wide = 0;  // resident workgroups per CU (main: 16, 64 lanes)
    // last result
    std::vector<int32_t> counts;
    std::vector<blz_tb_hsp> hsps;
    std::vector<uint32_t> runs_h;
};

void tb_release(TbState *s) {
    if (!s) return;
    void *bufs[] = {s->ctr, s->qs,   s->pairs, s->pre,   s->out, s->out_n,
                    s->runs, s->back, s->tb,   s->rows, s->wruns, s->ring};
    for (void *b : bufs)
        if (b) (void)hipFree(b);
    delete s;
}

}  // namespace blz

namespace {

static_assert(sizeof(TbHsp) == sizeof(blz_tb_hsp), "TbHsp mirrors blz_tb_hsp");
static_assert(sizeof(Hsp) == sizeof(blz_hsp), "Hsp mirrors blz_hsp");

constexpr size_t kMainTbBytes = 4ull << 30;   // all main-pass pair slots' traceback rows together
constexpr size_t kMainTbPerWave = 8u << 20;   // at most this much per pair slot
constexpr size_t kMainTbMin = 64u << 10;
constexpr int kMainRuns = 4096;               // edit-script runs per pair slot (both directions of one HSP)
constexpr size_t kWideBytes = 4ull << 30;     // wide pass: all waves' storage together
constexpr size_t kTbCellsMax = 0x7FFFFFFF;    // row offsets are 32-bit
constexpr int kWideWavesMax = 1024;

template <class T>
bool grow(blz_ctx *ctx, T *&ptr, size_t &cap, size_t need, const char *what) {
    if (need <= cap && ptr) return true;
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    cap = 0;
    const size_t n = std::max<size_t>(need, 1);
    if (!check(ctx, hipMalloc((void **)&ptr, n * sizeof(T)), what)) return false;
    cap = n;
    return true;
}

size_t pow2_at_least(size_t n) {
    size_t m = 1;
    while (m < n) m <<= 1;
    return m;
}

double elapsed_ms(hipEvent_t a, hipEvent_t b) {
    float ms = 0;
    return hipEventElapsedTime(&ms, a, b) == hipSuccess ? (double)ms : 0.0;
}

}  // namespace

extern "C" {

void blz_tb_default_params(blz_tb_params *p) {
    if (!p) return;
    p->x_dropoff = 64;
    p->gap_open = 11;
    p->gap_extend = 1;
    p->cutoff = 0;
}

blz_status blz_traceback(blz_ctx *ctx, const blz_db *db, const blz_query *const *queries, int32_t nq,
                         const int32_t *query_cutoffs, const blz_tb_pair *pairs, int64_t n_pairs,
                         const blz_hsp *hsps, const blz_tb_params *p, blz_tb_stats *stats) {
    if (!ctx || !db || !p || n_pairs < 0 || nq < 0 || (n_pairs > 0 && (!pairs || !queries)) ||
        n_pairs > (int64_t)0xFFFFFFF0 || db->ctx == nullptr || db->ctx->device != ctx->device)
        return BLZ_EINVAL;
    if (p->gap_extend <= 0 || p->gap_open < 0 || p->x_dropoff < 0) {
        ctx->err = "traceback: gap_extend must be positive, gap_open and x_dropoff non-negative";
        return BLZ_EINVAL;
    }
    const auto t0 = std::chrono::steady_clock::now();
    blz_tb_stats st{};
    st.pairs = (uint64_t)n_pairs;
    // validate and lay out the pairs: only the gapped starts address memory
    std::vector<TbQuery> qh((size_t)nq);
    int max_qlen = 0;
    for (int i = 0; i < nq; i++) {
        const blz_query *q = queries[i];
        if (!q || !q->t.d_query || q->ctx == nullptr || q->ctx->device != ctx->device) return BLZ_EINVAL;
        qh[i] = TbQuery{q->t.d_query, q->qlen, query_cutoffs ? query_cutoffs[i] : p->cutoff};
        max_qlen = std::max(max_qlen, q->qlen);
    }
    std::vector<TbPair> ph((size_t)n_pairs);
    uint64_t total = 0;
    for (int64_t i = 0; i < n_pairs; i++) {
        const blz_tb_pair &x = pairs[i];
        if (x.query < 0 || x.query >= nq || x.oid < 0 || x.oid >= db->nsubj || x.n_hsps < 0) {
            ctx->err = "traceback: pair " + std::to_string(i) + " names no query, subject or HSP count";
            return BLZ_EINVAL;
        }
        if (x.n_hsps > 0 && !hsps) return BLZ_EINVAL;
        const int qlen = qh[x.query].qlen;
        const int64_t slen = (int64_t)(db->h_offs[x.oid + 1] - db->h_offs[x.oid]);
        for (int k = 0; k < x.n_hsps; k++) {
            const blz_hsp &h = hsps[total + k];
            if (h.q_gs < 0 || h.q_gs >= qlen || h.s_gs < 0 || h.s_gs >= slen) {
                ctx->err = "traceback: HSP " + std::to_string(total + k) + " has its gapped start outside the pair";
                return BLZ_EINVAL;
            }
        }
        ph[i] = TbPair{x.oid, total, x.query, x.n_hsps};
        total += (uint64_t)x.n_hsps;
    }
    st.prelim_hsps = total;
    if (!check(ctx, hipSetDevice(ctx->device), "hipSetDevice")) return BLZ_EDEVICE;
    if (!ctx->tb) ctx->tb = new TbState();
    TbState &S = *ctx->tb;
    S.counts.clear();
    S.hsps.clear();
    S.runs_h.clear();
    const int lanes = ctx->tune.tb_lanes == 16 ? 16 : 64;
    const int pairs_per_wave = 64 / lanes;  // pairs a main-pass wave runs side by side
    if (!S.blocks_wide) {
        S.blocks_main[0] = std::max(1, traceback_blocks_per_cu(16, false));
        S.blocks_main[1] = std::max(1, traceback_blocks_per_cu(64, false));
        S.blocks_wide = std::max(1, traceback_blocks_per_cu(64, true));
    }
    hipStream_t s = ctx->stream;
    const size_t np = (size_t)n_pairs;
    // main pass shape: the resident waves (no more than the pairs need), each pair slot with
    // bounded traceback rows
    const size_t blocks = std::max<size_t>(
        1, std::min<size_t>((np + pairs_per_wave - 1) / pairs_per_wave,
                            (size_t)ctx->num_cus * S.blocks_main[lanes == 64]));
    const size_t waves = blocks * pairs_per_wave;  // pair slots, each with its own storage areas
    size_t tb_per_wave = std::min(kMainTbPerWave, std::max(kMainTbMin, (size_t)256 * (size_t)(max_qlen + 2)));
    tb_per_wave = std::min(tb_per_wave, std::max(kMainTbMin, kMainTbBytes / waves));
    if (ctx->tune.cap_tb_bytes > 0) tb_per_wave = (size_t)ctx->tune.cap_tb_bytes;
    const size_t rows_cap = (size_t)max_qlen + 2;
    if (!grow(ctx, S.ctr, S.ctr_cap, 1, "traceback counters") || !grow(ctx, S.qs, S.qs_cap, qh.size(), "traceback queries") ||
        !grow(ctx, S.pairs, S.pairs_cap, np, "traceback pairs") || !grow(ctx, S.pre, S.pre_cap, total, "traceback prelim HSPs") ||
        !grow(ctx, S.out, S.out_cap, total, "traceback HSPs") || !grow(ctx, S.out_n, S.out_n_cap, np, "traceback counts") ||
        !grow(ctx, S.back, S.back_cap, np, "traceback wide list") ||
        !grow(ctx, S.tb, S.tb_cap, waves * tb_per_wave, "traceback rows") ||
        !grow(ctx, S.rows, S.rows_cap, waves * rows_cap, "traceback row table") ||
        !grow(ctx, S.wruns, S.wruns_cap, waves * (size_t)kMainRuns, "traceback run scratch"))
        return BLZ_ENOMEM;
    if ((nq > 0 && !check(ctx, hipMemcpyAsync(S.qs, qh.data(), qh.size() * sizeof(TbQuery), hipMemcpyHostToDevice, s),
                          "traceback queries")) ||
        (np > 0 && !check(ctx, hipMemcpyAsync(S.pairs, ph.data(), np * sizeof(TbPair), hipMemcpyHostToDevice, s),
                          "traceback pairs")) ||
        (total > 0 &&
         !check(ctx, hipMemcpyAsync(S.pre, hsps, total * sizeof(Hsp), hipMemcpyHostToDevice, s), "traceback HSPs")))
        return BLZ_EDEVICE;
    TbArgs a{};
    a.res = db->res;
    a.offs = db->offs;
    a.qs = S.qs;
    a.pairs = S.pairs;
    a.n_pairs = (uint32_t)np;
    a.pre = S.pre;
    a.out = S.out;
    a.out_n = S.out_n;
    a.back = S.back;
    a.back_cap = (uint32_t)np;
    a.ctr = S.ctr;
    a.x_final = p->x_dropoff;
    a.gap_open = p->gap_open;
    a.gap_extend = p->gap_extend;
    size_t runs_need =
        ctx->tune.cap_tb_runs > 0 ? (size_t)ctx->tune.cap_tb_runs : std::max<size_t>(1u << 20, 32 * total);
    TbCounters hc{};
    for (int attempt = 0;; attempt++) {
        if (!grow(ctx, S.runs, S.runs_cap, runs_need, "traceback edit scripts")) return BLZ_ENOMEM;
        a.runs = S.runs;
        a.runs_cap = (attempt == 0 && ctx->tune.cap_tb_runs > 0) ? runs_need : S.runs_cap;
        // main pass
        a.list = nullptr;
        a.n_list = 0;
        a.tb = S.tb;
        a.tb_cap = tb_per_wave;
        a.rows = S.rows;
        a.rows_cap = (int32_t)rows_cap;
        a.wruns = S.wruns;
        a.wruns_cap = kMainRuns;
        a.ring = nullptr;
        a.ring_mask = 0;
        if (!check(ctx, hipMemsetAsync(S.ctr, 0, sizeof(TbCounters), s), "traceback counters") ||
            !check(ctx, hipEventRecord(ctx->ev[0], s), "event") ||
            !check(ctx, launch_traceback(a, (int)blocks, lanes, false, s), "traceback main pass") ||
            !check(ctx, hipEventRecord(ctx->ev[1], s), "event") ||
            !check(ctx, hipMemcpyAsync(&hc, S.ctr, sizeof(hc), hipMemcpyDeviceToHost, s), "traceback counters") ||
            !check(ctx, hipStreamSynchronize(s), "traceback main pass"))
            return BLZ_EDEVICE;
        st.ms_main = elapsed_ms(ctx->ev[0], ctx->ev[1]);
        st.ms_wide = 0;
        if (hc.bad_walk) {
            ctx->err = "traceback: a path left the stored rows";
            return BLZ_EDEVICE;
        }
        const size_t n_back = std::min<size_t>(hc.back_count, np);
        st.wide_pairs = n_back;
        if (n_back > 0) {
            // wide pass: storage for the largest pair given back (a direction's cells fit in
            // (query letters + 1) x (subject letters + 2), its band in subject letters + 2)
            std::vector<uint32_t> back(n_back);
            if (!check(ctx, hipMemcpyAsync(back.data(), S.back, n_back * sizeof(uint32_t), hipMemcpyDeviceToHost, s),
                       "traceback wide list") ||
                !check(ctx, hipStreamSynchronize(s), "traceback wide list"))
                return BLZ_EDEVICE;
            size_t tb_need = 1, ring_need = 1, rows_need = 1, wr_need = 1;
            for (uint32_t pi : back) {
                const size_t ql = (size_t)qh[ph[pi].qi].qlen;
                const size_t sl = (size_t)(db->h_offs[ph[pi].sid + 1] - db->h_offs[ph[pi].sid]);
                tb_need = std::max(tb_need, std::min(kTbCellsMax, (ql + 1) * (sl + 2)));
                ring_need = std::max(ring_need, pow2_at_least(sl + 2));
                rows_need = std::max(rows_need, ql + 2);
                wr_need = std::max(wr_need, 2 * (ql + sl) + 8);
            }
            const size_t per_wave = tb_need + ring_need * sizeof(int2) + rows_need * sizeof(int4) + wr_need * 4;
            size_t wide_waves = std::max<size_t>(1, kWideBytes / per_wave);
            wide_waves = std::min({wide_waves, n_back, (size_t)kWideWavesMax, (size_t)ctx->num_cus * S.blocks_wide});
            if (!grow(ctx, S.tb, S.tb_cap, wide_waves * tb_need, "traceback rows") ||
                !grow(ctx, S.rows, S.rows_cap, wide_waves * rows_need, "traceback row table") ||
                !grow(ctx, S.wruns, S.wruns_cap, wide_waves * wr_need, "traceback run scratch") ||
                !grow(ctx, S.ring, S.ring_cap, wide_waves * ring_need, "traceback band rings"))
                return BLZ_ENOMEM;
            a.list = S.back;
            a.n_list = (uint32_t)n_back;
            a.tb = S.tb;
            a.tb_cap = tb_need;
            a.rows = S.rows;
            a.rows_cap = (int32_t)rows_need;
            a.wruns = S.wruns;
            a.wruns_cap = (int32_t)wr_need;
            a.ring = S.ring;
            a.ring_mask = (int32_t)(ring_need - 1);
            // next_pair and back_count restart; the run reservations and cell counts go on
            if (!check(ctx, hipMemsetAsync(&S.ctr->next_pair, 0, 2 * sizeof(unsigned int), s), "traceback counters") ||
                !check(ctx, hipEventRecord(ctx->ev[2], s), "event") ||
                !check(ctx, launch_traceback(a, (int)wide_waves, 64, true, s), "traceback wide pass") ||
                !check(ctx, hipEventRecord(ctx->ev[3], s), "event") ||
                !check(ctx, hipMemcpyAsync(&hc, S.ctr, sizeof(hc), hipMemcpyDeviceToHost, s), "traceback counters") ||
                !check(ctx, hipStreamSynchronize(s), "traceback wide pass"))
                return BLZ_EDEVICE;
            st.ms_wide = elapsed_ms(ctx->ev[2], ctx->ev[3]);
            if (hc.bad_walk) {
                ctx->err = "traceback: a path left the stored rows";
                return BLZ_EDEVICE;
            }
            if (hc.back_count) {
                ctx->err = "traceback: " + std::to_string(hc.back_count) +
                           " pairs outgrow the wide pass (2^31 traceback cells in one direction)";
                return BLZ_EOVERFLOW;
            }
        }
        if (hc.runs_used <= a.runs_cap) break;
        if (attempt >= 2) {
            ctx->err = "traceback: the edit-script buffer filled again after growing";
            return BLZ_ENOMEM;
        }
        runs_need = (size_t)hc.runs_used;
        st.reruns++;
    }
    // compact: each pair's kept HSPs, their runs contiguous
    std::vector<int32_t> out_n(np);
    std::vector<TbHsp> out(total);
    std::vector<uint32_t> runs((size_t)hc.runs_used);
    if ((np > 0 && !check(ctx, hipMemcpyAsync(out_n.data(), S.out_n, np * 4, hipMemcpyDeviceToHost, s), "traceback")) ||
        (total > 0 &&
         !check(ctx, hipMemcpyAsync(out.data(), S.out, total * sizeof(TbHsp), hipMemcpyDeviceToHost, s), "traceback")) ||
        (!runs.empty() && !check(ctx, hipMemcpyAsync(runs.data(), S.runs, runs.size() * 4, hipMemcpyDeviceToHost, s),
                                 "traceback")) ||
        !check(ctx, hipStreamSynchronize(s), "traceback"))
        return BLZ_EDEVICE;
    S.counts.assign(out_n.begin(), out_n.end());
    for (size_t i = 0; i < np; i++) {
        for (int t = 0; t < out_n[i]; t++) {
            const TbHsp &o = out[ph[i].off + t];
            blz_tb_hsp r;
            std::memcpy(&r, &o, sizeof(r));
            r.op_first = (int32_t)S.runs_h.size();
            S.runs_h.insert(S.runs_h.end(), runs.begin() + o.op_first, runs.begin() + o.op_first + o.op_count);
            S.hsps.push_back(r);
        }
    }
    st.hsps = S.hsps.size();
    st.runs = S.runs_h.size();
    st.alignments = hc.dps;
    st.dp_cells = hc.cells;
    st.ms_total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (stats) *stats = st;
    return BLZ_OK;
}

blz_status blz_traceback_sizes(const blz_ctx *ctx, int64_t *n_pairs, int64_t *n_hsps, int64_t *n_runs) {
    if (!ctx) return BLZ_EINVAL;
    const TbState *S = ctx->tb;
    if (n_pairs) *n_pairs = S ? (int64_t)S->counts.size() : 0;
    if (n_hsps) *n_hsps = S ? (int64_t)S->hsps.size() : 0;
    if (n_runs) *n_runs = S ? (int64_t)S->runs_h.size() : 0;
    return BLZ_OK;
}

blz_status blz_traceback_copy(const blz_ctx *ctx, int32_t *counts, int64_t n_pairs, blz_tb_hsp *hsps,
                              int64_t hsp_cap, uint32_t *runs, int64_t runs_cap) {
    if (!ctx || n_pairs < 0 || hsp_cap < 0 || runs_cap < 0) return BLZ_EINVAL;
    const TbState *S = ctx->tb;
    const size_t np = S ? S->counts.size() : 0, nh = S ? S->hsps.size() : 0, nr = S ? S->runs_h.size() : 0;
    if ((size_t)n_pairs < np || (size_t)hsp_cap < nh || (size_t)runs_cap < nr) return BLZ_ECAPACITY;
    if ((np && !counts) || (nh && !hsps) || (nr && !runs)) return BLZ_EINVAL;
    if (np) std::memcpy(counts, S->counts.data(), np * sizeof(int32_t));
    if (nh) std::memcpy(hsps, S->hsps.data(), nh * sizeof(blz_tb_hsp));
    if (nr) std::memcpy(runs, S->runs_h.data(), nr * sizeof(uint32_t));
    return BLZ_OK;
}

}  // extern "C"
