// blz_seed.hip -- seed stage with LDS-resident query tables and batched ungapped extensions.
//
// A workgroup stages the query's lookup table (compact: u16 offsets inside 64-word blocks plus
// one u32 base per block), the query letters and BLOSUM62 in LDS once; each of its waves then
// takes subjects from a device counter.  Per subject a wave:
//   1. reads 64 consecutive subject words per step (one coalesced load per lane) and expands
//      their query hits in rounds of 64 (lane = hit);
//   2. applies each hit to its diagonal's two-hit state in a per-wave LDS ring; hits of one
//      round on the same diagonal are applied earliest-first (LDS atomicMin on the hit rank,
//      hashed on the diagonal);
//   3. does NOT extend a pair on the spot: the pair becomes a task and its diagonal is marked
//      pending; later hits on a pending diagonal are queued behind it.  When the task queue
//      fills, and at the end of the subject, all tasks are extended at once, one lane each,
//      their outcome written back to the diagonals, and the queued hits replayed in order.
// The result is the same state machine as NCBI's two-hit word finder (per diagonal, hits in
// subject order); only independent diagonals are interleaved.
// Reference counterpart: gpuBLAZE_full, BLAZE_gpu/gpuBLAZE.cu:885-1252.
#include <hip/hip_runtime.h>

#include "blz_common.h"
#include "blz_dev_common.h"
#include "blz_device.h"

namespace blz {

namespace {

constexpr uint32_t kFlag = 0x80000000u;     // right extension done: low bits = barrier + 1
constexpr uint32_t kPending = 0x40000000u;  // extension queued: low bits = task index
constexpr uint32_t kPosMask = 0x3FFFFFFFu;
constexpr int kTaskCap = 128;
constexpr int kDeferCap = 256;

struct Task {
    int32_t q, s, s_left, slot;
};

struct WaveState {
    uint32_t *ring;
    uint32_t *claim;
    uint32_t *excl;
    uint32_t *p0;
    Task *task;
    int2 *defer;
    uint32_t *cnt;  // [0] tasks, [1] deferred, [2] init hits
};

__device__ __forceinline__ uint32_t lds_bytes_align(uint32_t b) { return (b + 15u) & ~15u; }

struct Ctx {
    const int16_t (*mat)[28];
    const uint8_t *Q;
    int qlen;
    const uint8_t *S;
    int slen;
    uint32_t abs0;
    uint32_t ring_mask;
    uint32_t claim_mask;
    int window;
    int xdrop;
    int cutoff;
    InitHit *stage;
    int stage_cap;
};

// Two-hit state machine for one hit (q, s) on a diagonal this lane owns for now.
__device__ __forceinline__ void apply_hit(const Ctx &c, WaveState &w, int q, int s) {
    const uint32_t slot = (uint32_t)(q - s) & c.ring_mask;
    const uint32_t st = w.ring[slot];
    const uint32_t sabs = c.abs0 + (uint32_t)s;
    if (st & kPending) {
        const uint32_t k = atomicAdd(&w.cnt[1], 1u);
        w.defer[k] = make_int2(q, s);
        return;
    }
    if (st & kFlag) {
        if (sabs >= (st & kPosMask) - 1u) w.ring[slot] = sabs + 1u;
        return;
    }
    const uint32_t diff = st ? sabs - (st - 1u) : 0x7FFFFFFFu;
    if (diff >= (uint32_t)c.window) {
        w.ring[slot] = sabs + 1u;
        return;
    }
    if (diff < (uint32_t)kWord) return;
    const uint32_t t = atomicAdd(&w.cnt[0], 1u);
    w.task[t] = Task{q, s, (int)(st - 1u - c.abs0) + kWord, (int)slot};
    w.ring[slot] = kPending | t;
}

// Applies up to one hit per lane; same-diagonal hits go earliest-rank first.
__device__ __forceinline__ void resolve_round(const Ctx &c, WaveState &w, bool pend, int q, int s,
                                              uint32_t rank) {
    const uint32_t h = (uint32_t)(q - s) & c.claim_mask;
    while (__ballot(pend)) {
        if (pend) atomicMin(&w.claim[h], rank);
        wave_sync();
        const bool win = pend && w.claim[h] == rank;
        wave_sync();
        if (win) {
            w.claim[h] = 0xFFFFFFFFu;
            pend = false;
            apply_hit(c, w, q, s);
        }
        wave_sync();
    }
}

// Extends every queued pair (one lane each), writes the outcome to its diagonal, then replays
// the hits queued behind pending diagonals; repeats until nothing is pending.
__device__ void flush(const Ctx &c, WaveState &w, unsigned long long &n_ext) {
    const int lane = lane_id();
    for (;;) {
        wave_sync();
        const uint32_t nt = w.cnt[0];
        const uint32_t nd = w.cnt[1];
        if (nt == 0 && nd == 0) break;
        for (uint32_t t0 = 0; t0 < nt; t0 += 64) {
            const uint32_t t = t0 + lane;
            if (t < nt) {
                const Task tk = w.task[t];
                int hq, hs, hl, rext, slast;
                const int score = two_hit(c.mat, c.Q, c.qlen, c.S, c.slen, tk.s_left, tk.s, tk.q, c.xdrop, &hq,
                                          &hs, &hl, &rext, &slast);
                n_ext++;
                if (score >= c.cutoff) {
                    const uint32_t k = atomicAdd(&w.cnt[2], 1u);
                    if (k < (uint32_t)c.stage_cap) c.stage[k] = InitHit{hq, hs, hl, score};
                }
                w.ring[tk.slot] = rext ? ((c.abs0 + (uint32_t)slast - 2u + 1u) | kFlag) : (c.abs0 + (uint32_t)tk.s + 1u);
            }
        }
        wave_sync();
        if (lane == 0) w.cnt[0] = 0;
        wave_sync();
        // replay the queued hits [0, nd) in order; new deferrals land at nd and beyond
        for (uint32_t r = 0; r < nd; r += 64) {
            const uint32_t i = r + lane;
            int q = 0, s = 0;
            const bool pend = i < nd;
            if (pend) {
                const int2 d = w.defer[i];
                q = d.x;
                s = d.y;
            }
            resolve_round(c, w, pend, q, s, i);
        }
        wave_sync();
        const uint32_t nd2 = w.cnt[1];
        for (uint32_t i = nd + lane; i < nd2; i += 64) w.defer[i - nd] = w.defer[i];
        wave_sync();
        if (lane == 0) w.cnt[1] = nd2 - nd;
    }
}

}  // namespace

template <bool kLdsTable>
__global__ void __launch_bounds__(1024) k_seed2(SeedArgs a) {
    extern __shared__ __align__(16) uint8_t smem[];
    const int lane = lane_id();
    const int wave = threadIdx.x >> 6;
    const int nwaves = blockDim.x >> 6;
    // LDS is sized for the largest query of the launch (a.qlen, a.ring_mask, a.claim_mask,
    // a.n_entries); each item uses its own query's sizes
    const uint32_t M = a.ring_mask + 1;
    const uint32_t C = a.claim_mask + 1;

    // block-shared section
    uint32_t off = 0;
    int32_t *blk = reinterpret_cast<int32_t *>(smem + off);  // [0] item, [1] next subject of the item
    off += 16;
    int16_t(*mat)[28] = reinterpret_cast<int16_t(*)[28]>(smem + off);
    off += lds_bytes_align(28 * 28 * 2);
    uint8_t *Qs = smem + off;
    off += lds_bytes_align((uint32_t)a.qlen + 16);
    uint16_t *rel = nullptr;
    uint32_t *bb = nullptr;
    uint16_t *ent = nullptr;
    if (kLdsTable) {
        rel = reinterpret_cast<uint16_t *>(smem + off);
        off += lds_bytes_align((kWords + 1) * 2);
        bb = reinterpret_cast<uint32_t *>(smem + off);
        off += lds_bytes_align(kRelBlocks * 4);
        ent = reinterpret_cast<uint16_t *>(smem + off);
        off += lds_bytes_align(a.n_entries * 2 + 4);
    }
    // per-wave section
    const uint32_t wave_bytes = 4 * (M + C + 68 + 64 + 4) + sizeof(Task) * kTaskCap + sizeof(int2) * kDeferCap;
    uint8_t *wb = smem + off + (size_t)wave * lds_bytes_align(wave_bytes);
    WaveState w;
    w.ring = reinterpret_cast<uint32_t *>(wb);
    w.claim = w.ring + M;
    w.excl = w.claim + C;
    w.p0 = w.excl + 68;
    w.cnt = w.p0 + 64;
    w.task = reinterpret_cast<Task *>(w.cnt + 4);
    w.defer = reinterpret_cast<int2 *>(w.task + kTaskCap);

    for (int i = threadIdx.x; i < 28 * 28; i += blockDim.x) mat[i / 28][i % 28] = kBlosum62Dev[i / 28][i % 28];
    for (uint32_t i = lane; i < M; i += 64) w.ring[i] = 0;
    for (uint32_t i = lane; i < C; i += 64) w.claim[i] = 0xFFFFFFFFu;
    if (lane < 4) w.cnt[lane] = 0;

    const uint64_t gwave = (uint64_t)blockIdx.x * nwaves + wave;
    Ctx c;
    c.mat = mat;
    c.Q = Qs;
    c.stage = a.stage + gwave * (uint64_t)a.stage_cap;
    c.stage_cap = a.stage_cap;
    uint32_t abs_base = 0;  // wave-relative position: increases over every subject the wave sees
    unsigned long long n_hits = 0, n_ext = 0;
    int cur_q = -1;
    const uint32_t n_items = (uint32_t)a.nq * a.items_per_query;

    for (;;) {
        __syncthreads();  // the previous item (and its query tables) is finished by every wave
        if (threadIdx.x == 0) {
            blk[0] = (int32_t)atomicAdd(a.item_ctr, 1u);
            blk[1] = 0;
        }
        __syncthreads();
        const uint32_t item = (uint32_t)blk[0];
        if (item >= n_items) break;
        const int qi = (int)(item / a.items_per_query);
        const int64_t sub0 = a.s_begin + (int64_t)(item % a.items_per_query) * a.subj_per_item;
        const int64_t sub1 = sub0 + a.subj_per_item < a.s_end ? sub0 + a.subj_per_item : a.s_end;
        const QueryDev &qd = a.qd[qi];
        if (qi != cur_q) {  // stage this query's tables
            for (int i = threadIdx.x; i < qd.qlen; i += blockDim.x) Qs[i] = qd.query[i];
            if (kLdsTable) {
                for (int i = threadIdx.x; i <= kWords; i += blockDim.x) rel[i] = qd.rel16[i];
                for (int i = threadIdx.x; i < kRelBlocks; i += blockDim.x) bb[i] = qd.blockbase[i];
                const uint32_t *src = reinterpret_cast<const uint32_t *>(qd.entries);
                uint32_t *dst = reinterpret_cast<uint32_t *>(ent);
                for (uint32_t i = threadIdx.x; i < (qd.n_entries + 1) / 2; i += blockDim.x) dst[i] = src[i];
            }
            __syncthreads();
            cur_q = qi;
        }
        c.qlen = qd.qlen;
        c.ring_mask = qd.ring_mask;
        c.claim_mask = qd.claim_mask;
        c.window = qd.window;
        c.xdrop = qd.x_ungapped;
        c.cutoff = qd.ungapped_cutoff;
        const uint32_t *gprefix = qd.prefix;
        const uint16_t *gent = qd.entries;

        for (;;) {
            int64_t sid = 0;
            if (lane == 0) sid = sub0 + (int64_t)atomicAdd(reinterpret_cast<uint32_t *>(&blk[1]), 1u);
            sid = __shfl(sid, 0, 64);
            if (sid >= sub1) break;
            const uint64_t o0 = a.offs[sid];
            const int slen = (int)(a.offs[sid + 1] - o0);
            c.S = a.res + o0;
            c.slen = slen;
            c.abs0 = abs_base;
            abs_base += (uint32_t)slen + 128u;

            for (int c0 = 0; c0 + kWord <= slen; c0 += 64) {
                const int s = c0 + lane;
                uint32_t pa = 0, nh = 0;
                if (s + kWord <= slen) {
                    const uint32_t wd = c.S[s] * 784u + c.S[s + 1] * 28u + c.S[s + 2];
                    if (kLdsTable) {
                        pa = bb[wd >> 6] + rel[wd];
                        nh = bb[(wd + 1) >> 6] + rel[wd + 1] - pa;
                    } else {
                        pa = gprefix[wd];
                        nh = gprefix[wd + 1] - pa;
                    }
                }
                const uint32_t inc = wave_inclusive_sum(nh);
                const uint32_t H = __shfl(inc, 63, 64);
                w.excl[lane] = inc - nh;
                w.p0[lane] = pa;
                wave_sync();
                n_hits += nh;
                for (uint32_t r = 0; r < H; r += 64) {
                    if (w.cnt[0] > (uint32_t)(kTaskCap - 64) || w.cnt[1] > 64u) flush(c, w, n_ext);
                    const uint32_t g = r + lane;
                    const bool pend = g < H;
                    int src = 0, qo = 0;
                    if (pend) {
                        int lo = 0, hi = 63;
                        while (lo < hi) {
                            const int mid = (lo + hi + 1) >> 1;
                            if (w.excl[mid] <= g) lo = mid;
                            else hi = mid - 1;
                        }
                        src = lo;
                        const uint32_t e = w.p0[src] + (g - w.excl[src]);
                        qo = kLdsTable ? ent[e] : gent[e];
                    }
                    resolve_round(c, w, pend, qo, c0 + src, g);
                }
                wave_sync();
            }
            flush(c, w, n_ext);

            // hand the subject's init hits to the gapped stage
            wave_sync();
            const uint32_t nih = w.cnt[2];
            wave_sync();
            if (nih > 0) {
                const int64_t tag = ((int64_t)(a.qi_base + qi) << kOvfShift) | sid;
                if (nih > (uint32_t)a.stage_cap) {
                    if (lane == 0) {
                        const uint32_t k = atomicAdd(&a.ctr->overflow_count, 1u);
                        if (k < a.ovf_cap) a.ovf[k] = tag;
                    }
                } else {
                    unsigned long long base = 0;
                    if (lane == 0) base = atomicAdd(&a.ctr->pool_used, (unsigned long long)nih);
                    base = __shfl(base, 0, 64);
                    if (base + nih > a.pool_cap) {
                        if (lane == 0) {
                            const uint32_t k = atomicAdd(&a.ctr->overflow_count, 1u);
                            if (k < a.ovf_cap) a.ovf[k] = tag;
                        }
                    } else {
                        for (uint32_t k = lane; k < nih; k += 64) a.pool[base + k] = c.stage[k];
                        if (lane == 0) {
                            const uint32_t wi = atomicAdd(&a.ctr->work_count, 1u);
                            a.work[wi] = WorkItem{sid, (uint64_t)base, (int32_t)nih, a.qi_base + qi};
                            atomicAdd(&a.ctr->init_hits, (unsigned long long)nih);
                        }
                    }
                }
                wave_sync();
                if (lane == 0) w.cnt[2] = 0;
                wave_sync();
            }
        }
    }
    unsigned long long ext_tot = n_ext, hit_tot = n_hits;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        ext_tot += __shfl_xor(ext_tot, o, 64);
        hit_tot += __shfl_xor(hit_tot, o, 64);
    }
    if (lane == 0) {
        atomicAdd(&a.ctr->word_hits, hit_tot);
        atomicAdd(&a.ctr->ungapped_ext, ext_tot);
    }
}

size_t seed2_lds_bytes(const SeedArgs &a, bool lds_table, int waves) {
    auto al = [](size_t b) { return (b + 15) & ~size_t(15); };
    size_t b = 16 + al(28 * 28 * 2) + al((size_t)a.qlen + 16);
    if (lds_table) b += al((kWords + 1) * 2) + al(kRelBlocks * 4) + al((size_t)a.n_entries * 2 + 4);
    const size_t wave_bytes = 4 * ((size_t)a.ring_mask + 1 + a.claim_mask + 1 + 68 + 64 + 4) +
                              sizeof(Task) * kTaskCap + sizeof(int2) * kDeferCap;
    return b + (size_t)waves * al(wave_bytes);
}

hipError_t launch_seed2(const SeedArgs &a, bool lds_table, int blocks, int waves, hipStream_t st) {
    const size_t lds = seed2_lds_bytes(a, lds_table, waves);
    if (lds_table)
        hipLaunchKernelGGL(k_seed2<true>, dim3(blocks), dim3(64 * waves), lds, st, a);
    else
        hipLaunchKernelGGL(k_seed2<false>, dim3(blocks), dim3(64 * waves), lds, st, a);
    return hipGetLastError();
}

}  // namespace blz
