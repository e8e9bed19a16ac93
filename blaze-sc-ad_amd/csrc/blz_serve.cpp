// blz_serve.cpp -- request loop over a resident database: continuous batching of queries.
//
// Requests arrive from any thread (blz_server_submit).  A preparation thread builds each query's
// lookup tables on the device.  A dispatcher hands prepared requests to free slots as batches:
// whenever a slot is free it takes the oldest waiting request and, from the requests waiting
// with it, those of the same query-size class, up to the batch caps (queries and query letters:
// the LDS and HBM footprint of a batch grows with both).  Each slot owns a context (its own
// stream and work buffers) and runs its batch with blz_prelim_run_batch, so the batches of
// different slots overlap on the device; requests that arrive while batches run are admitted
// into the next free slot.  A request's latency is measured from submit to its lists being on
// the host.
//
// No reference counterpart: the reference is a per-search hook (blast_engine.c:2983-2996) with
// no request loop.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "blz_prelim.h"

namespace {

using Clock = std::chrono::steady_clock;

struct Request {
    uint64_t ticket = 0;
    std::vector<uint8_t> letters;
    blz_params params{};
    blz_query *q = nullptr;
    Clock::time_point t_submit, t_ready, t_start, t_done;
    int32_t batch_queries = 0;
    blz_status status = BLZ_OK;
    bool done = false;
    std::vector<int64_t> oids;
    std::vector<int32_t> counts;
    std::vector<blz_hsp> hsps;
    std::vector<int64_t> overflow;  // subjects its batch handed back
};

// Work items a query is expected to make: (query, subject) pairs with init hits, about one per
// 10,000 letter-subject pairs at the default parameters (99,759 for 500 letters x 2,000,000
// subjects); the server starts from half again as much and then learns from its batches.
constexpr double kPriorWorkRate = 1.5e-4;
constexpr int64_t kSlotWorkItems = 48ll << 20;   // a slot context's work list at start
constexpr int64_t kBytesPerWorkItem = 100;       // work item, two init hits, order entry, result

// Query-size class of a request: the batch former keeps a batch within one class, so the
// batch's seed launches are sized for similar queries.
int size_class(int32_t qlen) { return qlen <= 400 ? 0 : qlen <= 1000 ? 1 : qlen <= 2000 ? 2 : 3; }

// Device time a batch is expected to take: its query letters x the database's residues at a rate
// (letter-residue pairs per second) learned from the batches already run.  The prior is the
// measured rate of one 500-letter query over 825.6 M residues in 11.7 ms.
constexpr double kPriorPairRate = 3.5e13;
constexpr double kDefaultOverlapMs = 5.0;

struct Slot {
    blz_ctx *ctx = nullptr;
    std::thread th;
    std::vector<Request *> batch;  // assigned by the dispatcher, cleared by the slot
    bool busy = false;
    Clock::time_point started;     // when its batch was handed out
    double est_s = 0;              // the batch's expected device time
};

}  // namespace

struct blz_server {
    blz_server_config cfg{};
    const blz_db *db = nullptr;
    blz_ctx *prep_ctx = nullptr;
    std::vector<Slot> slots;
    std::mutex mu;
    std::condition_variable cv;  // any state change
    std::deque<Request *> to_prepare, waiting;
    std::unordered_map<uint64_t, Request *> requests;
    uint64_t next_ticket = 1;
    bool stopping = false;
    std::thread prep_th, dispatch_th;
    blz_server_stats stats{};
    int64_t nsubj = 0;
    double work_rate = kPriorWorkRate;  // work items per query letter and subject
    int64_t work_budget = 0;            // estimated work items a batch may hold
    uint64_t nres = 0;
    double pair_rate = kPriorPairRate;  // query letters x residues per second of device time
};

namespace {

void prepare_loop(blz_server *s) {
    for (;;) {
        Request *r;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            s->cv.wait(lk, [&] { return s->stopping || !s->to_prepare.empty(); });
            if (s->stopping) return;
            r = s->to_prepare.front();
            s->to_prepare.pop_front();
        }
        const blz_status st =
            blz_query_prepare(s->prep_ctx, r->letters.data(), (int32_t)r->letters.size(), &r->params, &r->q);
        std::lock_guard<std::mutex> lk(s->mu);
        r->t_ready = Clock::now();
        if (st != BLZ_OK) {
            r->status = st;
            r->done = true;
            r->t_done = r->t_ready;
        } else {
            s->waiting.push_back(r);
        }
        s->cv.notify_all();
    }
}

// Whether a request may share a batch: a query that packs (at most half a pack's letters, the pack
// being 1024 - 128 - window letters) shares the seed scan with its batch-mates; a longer one scans
// the database alone either way, so where its scan takes long (more than kBatchLongS) batching it
// would only hold its batch-mates' results back until it is done.
constexpr double kBatchLongS = 0.02;
bool batchable(const blz_server *s, const Request *r) {
    const int pack = std::max(64, 1024 - 128 - r->params.window);
    if ((int)r->letters.size() <= pack / 2) return true;
    return (double)r->letters.size() * (double)s->nres / s->pair_rate <= kBatchLongS;
}

// Oldest waiting request plus same-class requests behind it, within the caps: queries, query
// letters, and the estimated device footprint (work items, each request's at most one per subject).
std::vector<Request *> form_batch(blz_server *s) {
    std::vector<Request *> b;
    if (s->waiting.empty()) return b;
    if (!batchable(s, s->waiting.front())) {
        b.push_back(s->waiting.front());
        s->waiting.pop_front();
        return b;
    }
    const int cls = size_class((int32_t)s->waiting.front()->letters.size());
    int64_t letters = 0;
    double work = 0;
    for (auto it = s->waiting.begin(); it != s->waiting.end() && (int)b.size() < s->cfg.max_batch;) {
        Request *r = *it;
        const int64_t L = (int64_t)r->letters.size();
        const double w = std::min((double)s->nsubj, s->work_rate * (double)L * (double)s->nsubj);
        if (size_class((int32_t)L) == cls && batchable(s, r) &&
            (b.empty() || (letters + L <= s->cfg.max_batch_letters && work + w <= (double)s->work_budget))) {
            b.push_back(r);
            letters += L;
            work += w;
            it = s->waiting.erase(it);
        } else {
            ++it;
        }
    }
    return b;
}

// Device time the batches in flight are still expected to need.
double inflight_s(const blz_server *s, Clock::time_point now) {
    double rem = 0;
    for (const auto &sl : s->slots)
        if (sl.busy) rem += std::max(0.0, sl.est_s - std::chrono::duration<double>(now - sl.started).count());
    return rem;
}

void dispatch_loop(blz_server *s) {
    std::unique_lock<std::mutex> lk(s->mu);
    const double overlap_s = s->cfg.overlap_ms == 0 ? kDefaultOverlapMs * 1e-3 : s->cfg.overlap_ms * 1e-3;
    for (;;) {
        if (s->stopping) return;
        Slot *free_slot = nullptr;
        for (auto &sl : s->slots)
            if (!sl.busy) {
                free_slot = &sl;
                break;
            }
        if (s->waiting.empty() || !free_slot) {
            s->cv.wait(lk);
            continue;
        }
        const Clock::time_point now = Clock::now();
        const double rem = inflight_s(s, now);
        if (overlap_s >= 0 && rem > overlap_s) {  // the GPU is still taken: wait, and let requests gather
            s->cv.wait_for(lk, std::chrono::duration<double>(rem - overlap_s));
            continue;
        }
        if (s->cfg.linger_us > 0 && (int)s->waiting.size() < s->cfg.max_batch) {  // let others join
            const Clock::time_point until = s->waiting.front()->t_ready + std::chrono::microseconds(s->cfg.linger_us);
            if (now < until) {
                s->cv.wait_until(lk, until);
                continue;
            }
        }
        free_slot->batch = form_batch(s);
        int64_t letters = 0;
        for (const Request *r : free_slot->batch) letters += (int64_t)r->letters.size();
        free_slot->busy = true;
        free_slot->started = now;
        free_slot->est_s = (double)letters * (double)s->nres / s->pair_rate;
        s->stats.batches++;
        s->cv.notify_all();
    }
}

void slot_loop(blz_server *s, Slot *sl) {
    for (;;) {
        std::vector<Request *> batch;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            s->cv.wait(lk, [&] { return s->stopping || (sl->busy && !sl->batch.empty()); });
            if (s->stopping) return;
            batch.swap(sl->batch);
        }
        const Clock::time_point t0 = Clock::now();
        std::vector<const blz_query *> qs;
        for (Request *r : batch) qs.push_back(r->q);
        blz_run_stats st{};
        const blz_status rc = blz_prelim_run_batch(sl->ctx, qs.data(), (int32_t)qs.size(), s->db, nullptr, &st);
        // pairs handed back, per request
        std::vector<int64_t> handed;
        if (rc == BLZ_EOVERFLOW) {
            int64_t n = 0;
            blz_overflow_drain(sl->ctx, nullptr, 0, &n);
            handed.resize((size_t)n);
            if (n > 0 && blz_overflow_drain(sl->ctx, handed.data(), n, &n) != BLZ_OK) handed.clear();
        }
        int64_t letters = 0;
        for (size_t i = 0; i < batch.size(); i++) {
            Request *r = batch[i];
            letters += (int64_t)r->letters.size();
            for (int64_t e : handed)
                if ((e >> 40) == (int64_t)i) r->overflow.push_back(e & ((int64_t(1) << 40) - 1));
            r->status = rc;
            if (rc == BLZ_OK || rc == BLZ_EOVERFLOW) {
                int64_t ns = 0, nh = 0;
                blz_result_sizes(sl->ctx, (int32_t)i, &ns, &nh);
                r->oids.resize((size_t)std::max<int64_t>(ns, 1));
                r->counts.resize((size_t)std::max<int64_t>(ns, 1));
                r->hsps.resize((size_t)std::max<int64_t>(nh, 1));
                blz_hsp_sink sink{r->oids.data(), r->counts.data(), ns, r->hsps.data(), nh, 0, 0};
                const blz_status c = blz_result_copy(sl->ctx, (int32_t)i, &sink);
                if (c != BLZ_OK) r->status = c;
                r->oids.resize((size_t)ns);
                r->counts.resize((size_t)ns);
                r->hsps.resize((size_t)nh);
            }
            blz_query_free(r->q);
            r->q = nullptr;
        }
        std::lock_guard<std::mutex> lk(s->mu);
        const Clock::time_point t1 = Clock::now();
        for (Request *r : batch) {
            r->t_start = t0;
            r->t_done = t1;
            r->batch_queries = (int32_t)batch.size();
            r->done = true;
        }
        s->stats.completed += batch.size();
        s->stats.residues += st.residues;
        s->stats.device_ms += st.ms_total;
        s->stats.reruns += st.reruns > 0 ? 1 : 0;
        // the batch's letter-residue rate, from its device time (the part of a wall time shared
        // with overlapping batches is not its own); the estimate follows it smoothly
        if (letters > 0 && st.ms_total > 0) {
            const double r = (double)letters * (double)s->nres / (st.ms_total * 1e-3);
            s->pair_rate = std::min(4 * s->pair_rate, std::max(0.25 * s->pair_rate, 0.7 * s->pair_rate + 0.3 * r));
            s->stats.pair_rate = s->pair_rate;
        }
        // the admission's rate follows the batches: a quarter above the highest rate seen lately
        if (letters > 0 && s->nsubj > 0) {
            const double seen = (double)st.gapped_subjects / ((double)letters * (double)s->nsubj);
            s->work_rate = std::max(1.25 * seen, std::max(0.95 * s->work_rate, kPriorWorkRate / 16));
            s->stats.work_rate = s->work_rate;
        }
        sl->busy = false;
        s->cv.notify_all();
    }
}

}  // namespace

extern "C" {

void blz_server_default_config(blz_server_config *c) {
    c->slots = 2;
    c->max_batch = 64;
    c->max_batch_letters = 32768;
    c->linger_us = 0;
    c->max_batch_work = 0;
    c->overlap_ms = 0;
}

blz_status blz_server_create(int32_t device, const blz_db *db, const blz_server_config *cfg,
                             const blz_tuning *tuning, blz_server **out) {
    if (!db || !out) return BLZ_EINVAL;
    *out = nullptr;
    blz_server *s = new blz_server();
    blz_server_default_config(&s->cfg);
    if (cfg) {
        if (cfg->slots > 0) s->cfg.slots = cfg->slots;
        if (cfg->max_batch > 0) s->cfg.max_batch = cfg->max_batch;
        if (cfg->max_batch_letters > 0) s->cfg.max_batch_letters = cfg->max_batch_letters;
        if (cfg->max_batch_work > 0) s->cfg.max_batch_work = cfg->max_batch_work;
        s->cfg.overlap_ms = cfg->overlap_ms;
        s->cfg.linger_us = cfg->linger_us > 0 ? cfg->linger_us : 0;
    }
    if (s->cfg.slots > 16) s->cfg.slots = 16;
    s->db = db;
    s->nsubj = blz_db_num_subjects(db);
    s->nres = blz_db_num_residues(db);
    s->stats.work_rate = s->work_rate;
    s->stats.pair_rate = s->pair_rate;
    blz_status st = blz_ctx_create(device, &s->prep_ctx);
    // a batch's footprint budget: what a slot's buffers hold at start, within its share of half
    // the free device memory
    s->work_budget = s->cfg.max_batch_work;
    if (s->work_budget <= 0) {
        size_t free_b = 0, total_b = 0;
        if (st == BLZ_OK && hipSetDevice(device) == hipSuccess && hipMemGetInfo(&free_b, &total_b) == hipSuccess)
            s->work_budget = std::min<int64_t>(kSlotWorkItems, (int64_t)(free_b / 2 / (size_t)s->cfg.slots) /
                                                                   kBytesPerWorkItem);
        else
            s->work_budget = kSlotWorkItems;
    }
    s->slots.resize((size_t)s->cfg.slots);
    for (auto &sl : s->slots) {
        if (st != BLZ_OK) break;
        st = blz_ctx_create(device, &sl.ctx);
        if (st == BLZ_OK && tuning) st = blz_ctx_set_tuning(sl.ctx, tuning);
        if (st == BLZ_OK) {  // batches differ from one to the next: no graph capture
            blz_tuning t;
            blz_ctx_get_tuning(sl.ctx, &t);
            t.graph = 1;
            st = blz_ctx_set_tuning(sl.ctx, &t);
        }
    }
    if (st != BLZ_OK) {
        for (auto &sl : s->slots) blz_ctx_destroy(sl.ctx);
        blz_ctx_destroy(s->prep_ctx);
        delete s;
        return st;
    }
    s->prep_th = std::thread(prepare_loop, s);
    s->dispatch_th = std::thread(dispatch_loop, s);
    for (auto &sl : s->slots) sl.th = std::thread(slot_loop, s, &sl);
    *out = s;
    return BLZ_OK;
}

void blz_server_destroy(blz_server *s) {
    if (!s) return;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stopping = true;
        s->cv.notify_all();
    }
    s->prep_th.join();
    s->dispatch_th.join();
    for (auto &sl : s->slots) sl.th.join();
    for (auto &kv : s->requests) {
        if (kv.second->q) blz_query_free(kv.second->q);
        delete kv.second;
    }
    for (auto &sl : s->slots) blz_ctx_destroy(sl.ctx);
    blz_ctx_destroy(s->prep_ctx);
    delete s;
}

blz_status blz_server_submit(blz_server *s, const uint8_t *query, int32_t qlen, const blz_params *params,
                             uint64_t *ticket) {
    if (!s || !query || qlen <= 0 || !ticket) return BLZ_EINVAL;
    Request *r = new Request();
    r->t_submit = Clock::now();
    r->letters.assign(query, query + qlen);
    if (params)
        r->params = *params;
    else
        blz_default_params(&r->params);
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->stopping) {
        delete r;
        return BLZ_EINVAL;
    }
    r->ticket = s->next_ticket++;
    s->requests[r->ticket] = r;
    s->to_prepare.push_back(r);
    s->stats.submitted++;
    *ticket = r->ticket;
    s->cv.notify_all();
    return BLZ_OK;
}

blz_status blz_server_wait(blz_server *s, uint64_t ticket, int32_t timeout_ms, blz_request_info *info) {
    if (!s) return BLZ_EINVAL;
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->requests.find(ticket);
    if (it == s->requests.end()) return BLZ_EINVAL;
    Request *r = it->second;
    auto ready = [&] { return r->done; };
    if (timeout_ms < 0)
        s->cv.wait(lk, ready);
    else if (!s->cv.wait_for(lk, std::chrono::milliseconds(timeout_ms), ready))
        return BLZ_ETIMEOUT;
    if (info) {
        auto us = [](Clock::duration d) { return std::chrono::duration<double, std::micro>(d).count(); };
        info->status = r->status;
        info->batch_queries = r->batch_queries;
        info->latency_us = us(r->t_done - r->t_submit);
        info->queue_us = us(r->t_start - r->t_submit);
        info->n_subjects = (int64_t)r->oids.size();
        info->n_hsps = (int64_t)r->hsps.size();
        info->n_overflow = (int64_t)r->overflow.size();
    }
    return BLZ_OK;
}

blz_status blz_server_overflow(blz_server *s, uint64_t ticket, int64_t *subjects, int64_t cap, int64_t *n) {
    if (!s || !n || cap < 0 || (cap > 0 && !subjects)) return BLZ_EINVAL;
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->requests.find(ticket);
    if (it == s->requests.end() || !it->second->done) return BLZ_EINVAL;
    const std::vector<int64_t> &v = it->second->overflow;
    *n = (int64_t)v.size();
    if (*n > cap) return BLZ_ECAPACITY;
    std::copy(v.begin(), v.end(), subjects);
    return BLZ_OK;
}

blz_status blz_server_copy(blz_server *s, uint64_t ticket, blz_hsp_sink *sink) {
    if (!s || !sink) return BLZ_EINVAL;
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->requests.find(ticket);
    if (it == s->requests.end() || !it->second->done) return BLZ_EINVAL;
    Request *r = it->second;
    sink->n_subjects = (int64_t)r->oids.size();
    sink->n_hsps = (int64_t)r->hsps.size();
    if (sink->n_subjects > sink->subject_cap || sink->n_hsps > sink->hsp_cap) return BLZ_ECAPACITY;
    std::copy(r->oids.begin(), r->oids.end(), sink->oids);
    std::copy(r->counts.begin(), r->counts.end(), sink->counts);
    std::copy(r->hsps.begin(), r->hsps.end(), sink->hsps);
    return BLZ_OK;
}

blz_status blz_server_release(blz_server *s, uint64_t ticket) {
    if (!s) return BLZ_EINVAL;
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->requests.find(ticket);
    if (it == s->requests.end() || !it->second->done) return BLZ_EINVAL;
    delete it->second;
    s->requests.erase(it);
    return BLZ_OK;
}

blz_status blz_server_get_stats(blz_server *s, blz_server_stats *out) {
    if (!s || !out) return BLZ_EINVAL;
    std::lock_guard<std::mutex> lk(s->mu);
    *out = s->stats;
    return BLZ_OK;
}

}  // extern "C"
