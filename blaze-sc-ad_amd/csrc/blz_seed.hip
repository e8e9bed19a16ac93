// blz_seed.hip -- seed stage: word lookup, two-hit diagonal state, batched ungapped X-drop.
//
// Work items are (query pack, range of subjects), taken pack-major from a device counter by
// persistent workgroups; ranges shrink towards the end of the item list (guided sizes).  A workgroup stages its item's query in LDS: letters, BLOSUM62 and,
// depending on the table placement kTable, the lookup entries (u16 query offsets per word) and
// the compact lookup prefix (u32 base per 64-word block + u16 offset per word):
//   kTable 2: prefix and entries in LDS, 1: entries in LDS, prefix in global memory,
//   0: both in global memory (long queries).
// Each wave then takes subjects of the item from an LDS counter.  Per subject a wave:
//   1. copies the subject into its LDS buffer (when it fits) with 16-byte loads;
//   2. reads 64 consecutive words per step (lane = position), finds each word's hit range, lays
//      the step's hits out in an LDS list with a wave prefix sum and expands them 64 at a time
//      (lane = hit);
//   3. applies each hit to its diagonal's two-hit state in a per-wave LDS ring; hits of one
//      round on the same diagonal are applied earliest-first (LDS atomicMin on the hit rank,
//      hashed on the diagonal);
//   4. does not extend a pair on the spot: the pair becomes a task and its diagonal is marked
//      pending; later hits on a pending diagonal are queued behind it.  When the task queue
//      fills, and at the end of the subject, all tasks are extended at once, one lane each,
//      their outcome written back to the diagonals, and the queued hits replayed in order.
// The result is NCBI's two-hit word finder (per diagonal, hits in subject order); only
// independent diagonals are interleaved.
// Reference counterpart: gpuBLAZE_full, BLAZE_gpu/gpuBLAZE.cu:885-1252.
#include <hip/hip_runtime.h>

#include <map>
#include <mutex>
#include <utility>

#include "blz_common.h"
#include "blz_dev_common.h"
#include "blz_device.h"

namespace blz {

namespace {

// Diagonal ring entry: 0 empty; last hit s + 1; kFlag | (barrier + 1) after a right extension
// (NCBI's flag and last_hit = s_last_off - 2); kPending | task while an extension is queued.
// Positions are offsets in the current subject: the ring is cleared for every subject.  Rings
// of 16-bit entries (LDS) take subjects of up to kNarrowMax letters, longer ones use a ring of
// 32-bit entries in global memory.
template <class R>
struct RingFmt;
template <>
struct RingFmt<uint16_t> {
    static constexpr uint32_t kFlag = 0x8000u, kPending = 0x4000u, kPosMask = 0x3FFFu;
};
template <>
struct RingFmt<uint32_t> {
    static constexpr uint32_t kFlag = 0x80000000u, kPending = 0x40000000u, kPosMask = 0x3FFFFFFFu;
};
constexpr int kNarrowMax = kSeedNarrowMax;  // entries hold at most slen - 1 below the flag bits

// Orders the wave's ring accesses: LDS for the 16-bit ring; the 32-bit ring lives in global
// memory, whose accesses need a workgroup-scope fence.
template <class R>
__device__ __forceinline__ void ring_sync() {
    if constexpr (sizeof(R) == 4) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    } else {
        wave_sync();
    }
}
constexpr int kTaskCap = 128;
constexpr int kDeferCap = 256;
constexpr int kHitBuf = 128;                // most hits of one step that go through the queue
constexpr int kHitQueue = 192;              // queued hits: fewer than 64 waiting + one step
constexpr int kHitSlack = 8;                // entries past the queue that a step's inline stores may hit
constexpr int kSubjBuf = 1024;              // subject letters staged in LDS per wave
constexpr int kSeqPad = 16;                 // zero letters before and after a sequence staged in LDS
constexpr int kMatStride = 32;              // LDS matrix row: 28 scores, 4 unused (64 bytes)
constexpr uint32_t kMatBytes = 28 * kMatStride * 2;
static_assert(kSubjBuf <= kNarrowMax, "staged subjects use the LDS ring");
static_assert(kWordRecInline == 6, "scan_subject writes six inline entries");

// Positions in the hit queue, the task list and the deferred list are 16-bit offsets from the
// first position of the current window (see scan_subject); the diagonal ring holds absolute
// positions.  A task's first-hit end is stored plus kEndBias: it may lie up to window - 1 letters
// before the window's first position.
constexpr uint32_t kEndBias = 64;
struct Task {
    uint32_t q_slot;  // query offset | diagonal slot << 16
    uint32_t s_pack;  // second hit's offset | (end of the first hit's word - window start + kEndBias) << 16
};

// Short subjects scanned together by one wave (a group).  Their letters sit in the wave's LDS
// buffer at 64-aligned starts with at least 4 zero letters between them, so a step of 64 words
// never spans two subjects and a walk stops at a subject's end.  The two-hit state sees virtual
// positions: subject i's letters shifted by i * gap, gap >= window, so no pair, flag or diagonal
// state carries from one subject to the next, and one diagonal ring, task queue and flush serve
// the whole group (fuller extension batches, one ring clear per group).
constexpr int kGroupMax = 8;
struct Group {
    int64_t sid[kGroupMax];
    uint32_t start[kGroupMax];   // first letter in the buffer
    uint32_t len[kGroupMax];
    uint32_t vstart[kGroupMax];  // start + i * gap
};
// Subjects too long for the buffer (up to kNarrowMax letters) are grouped the same way while
// their letters stay in global memory: subject i's positions are shifted by vstart[i] (the
// previous subjects' lengths plus gap each), up to kNarrowMax virtual positions in all (the
// 16-bit ring's position field).  The table lies in the (then unused) letter buffer, laid out as
// a Group (start unused) followed by the subjects' DB offsets.
struct GlobalGroup {
    Group g;
    uint64_t o0[kGroupMax];
};

template <class R>
struct WaveState {
    R *ring;
    uint32_t *claim;
    uint32_t *excl;
    uint32_t *p0;
    uint32_t *cnt;  // [2] init hits staged ([0], [1], [3] unused)
    uint32_t *hits;
    Task *task;
    uint32_t *defer;  // query offset | subject offset << 16
    uint8_t *subj;
    Group *grp;
};

__host__ __device__ constexpr uint32_t al16(uint32_t b) { return (b + 15u) & ~15u; }

__host__ __device__ inline uint32_t wave_lds_bytes(uint32_t M, uint32_t C) {
    return al16(2 * M + 4 * (C + 68 + 64 + 4 + kHitQueue + kHitSlack) + sizeof(Task) * kTaskCap + sizeof(uint32_t) * kDeferCap +
                kSubjBuf + 2 * kSeqPad + sizeof(Group));
}

__host__ __device__ inline uint32_t block_lds_bytes(int table, uint32_t qlen, uint32_t n_entries) {
    uint32_t b = 16 + al16(kMatBytes) + kSeqPad + al16(qlen + kSeqPad) + al16(2 * (qlen + 2 * kSeqPad));
    if (table >= 1) b += al16(n_entries * 2 + 4);
    if (table >= 2) b += al16((kWords + 1) * 2) + al16(kRelBlocks * 4);
    return b;
}

struct Ctx {
    const int16_t (*mat)[kMatStride];
    const uint8_t *Q;
    const uint16_t *Qx;  // LDS: the query letters times a matrix row's bytes (kPad walks)
    int qlen;
    uint32_t ring_len, ring_base, ring_magic;
    uint32_t claim_mask;
    int window;
    int xdrop;
    int cutoff;
    int gap;  // subject groups: virtual distance between subjects (0: no groups)
    InitHit *stage;
    int stage_cap;
};

// Ring slot of diagonal q - s: (q + ring_base - s) mod ring_len by multiply-high.
__device__ __forceinline__ uint32_t ring_slot(const Ctx &c, int q, int s) {
    const uint32_t x = (uint32_t)q + c.ring_base - (uint32_t)s;
    return x - __umulhi(x, c.ring_magic) * c.ring_len;
}

// Ungapped walks (NCBI s_BlastAaExtendLeft / s_BlastAaExtendRight, the stop rules of
// extend_left / extend_right), four letters per step with no branch inside a step: the loads of
// a step are independent, and a lane that stops keeps its score, best and stop position while
// the step's remaining letters are masked out.
//
// kPad: the query and the subject sit in LDS with at least 4 letters of value 0 before and after
// them.  Letter 0 scores -32768 against every letter, so at a sequence end the walk stops by its
// own X-drop test (right: score <= 0) with the maximum, length and stop position it had there,
// exactly where the bounded loop ends; no bounds test per letter.  Otherwise a position past an
// end reads an in-bounds letter and scores kEndScore instead.
constexpr int kEndScore = -32768;

// Score of the query letter at qi against a subject letter.  With kPad both come from LDS in
// byte-offset form: c.Qx holds each query letter times a matrix row's bytes and the wave's
// subject buffer each subject letter doubled, so the lookup address is one add.  Otherwise sv is
// the subject letter itself and the query letter comes from c.Q.
template <bool kPad>
__device__ __forceinline__ int pair_score(const Ctx &c, int qi, uint32_t sv) {
    if constexpr (kPad)
        return *reinterpret_cast<const int16_t *>(reinterpret_cast<const uint8_t *>(c.mat) + (uint32_t)c.Qx[qi] + sv);
    return c.mat[c.Q[qi]][sv];
}

template <bool kPad, class SP>
__device__ __forceinline__ int walk_left(const Ctx &c, SP S, int q_off, int s_off, int dropoff, int *length) {
    const int n = s_off < q_off ? s_off : q_off;  // last k inside both sequences
    int maxscore = 0, score = 0, best = -1;
    bool live = true;
    for (int k0 = 0;; k0 += 4) {
        int sc[4];
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int k = k0 + j;
            if constexpr (kPad) {
                sc[j] = pair_score<true>(c, q_off - k, S[s_off - k]);
            } else {
                const int kk = k <= n ? k : n;
                const int v = pair_score<false>(c, q_off - kk, S[s_off - kk]);
                sc[j] = k <= n ? v : kEndScore;
            }
        }
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int s1 = score + sc[j];
            const bool up = live && s1 > maxscore;
            best = up ? k0 + j : best;
            maxscore = up ? s1 : maxscore;
            live = live && maxscore - s1 < dropoff;
            score = s1;
        }
        if (!live) break;
    }
    *length = best + 1;
    return maxscore;
}

template <bool kPad, class SP>
__device__ __forceinline__ int walk_right(const Ctx &c, SP S, int slen, int q_off, int s_off, int dropoff,
                                          int maxscore, int *length, int *s_last_off) {
    const int n = slen - s_off < c.qlen - q_off ? slen - s_off : c.qlen - q_off;  // letters inside both
    const int nl = n > 0 ? n - 1 : 0;
    int score = maxscore, best = -1, stop_k = 0;
    bool live = true;
    for (int k0 = 0;; k0 += 4) {
        int sc[4];
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int k = k0 + j;
            if constexpr (kPad) {
                sc[j] = pair_score<true>(c, q_off + k, S[s_off + k]);
            } else {
                const int kk = k < n ? k : nl;
                const int v = pair_score<false>(c, q_off + kk, S[s_off + kk]);
                sc[j] = k < n ? v : kEndScore;
            }
        }
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int s1 = score + sc[j];
            const bool up = live && s1 > maxscore;
            best = up ? k0 + j : best;
            maxscore = up ? s1 : maxscore;
            const bool stop = live && (s1 <= 0 || maxscore - s1 >= dropoff);
            stop_k = stop ? k0 + j : stop_k;
            live = live && !stop;
            score = s1;
        }
        if (!live) break;
    }
    *length = best + 1;
    *s_last_off = s_off + stop_k;
    return maxscore;
}

// NCBI s_BlastAaExtendTwoHit (reference restatement gpuBLAZE.cu:1116-1154).
template <bool kPad, class SP>
__device__ __forceinline__ int two_hit_walk(const Ctx &c, SP S, int slen, int s_left_off, int s_right_off,
                                            int q_right_off, int *hsp_q, int *hsp_s, int *hsp_len,
                                            int *right_extend, int *s_last_off) {
    int left_d = 0, right_d = 0, left_score = 0, right_score = 0, score = 0;
#pragma unroll
    for (int i = 0; i < kWord; i++) {
        score += pair_score<kPad>(c, q_right_off + i, S[s_right_off + i]);
        if (score > left_score) {
            left_score = score;
            right_d = i + 1;
        }
    }
    q_right_off += right_d;
    s_right_off += right_d;
    right_d = 0;
    *right_extend = 0;
    *s_last_off = s_right_off;
    left_score = walk_left<kPad>(c, S, q_right_off - 1, s_right_off - 1, c.xdrop, &left_d);
    if (left_d >= s_right_off - s_left_off) {
        *right_extend = 1;
        right_score = walk_right<kPad>(c, S, slen, q_right_off, s_right_off, c.xdrop, left_score, &right_d, s_last_off);
    }
    *hsp_q = q_right_off - left_d;
    *hsp_s = s_right_off - left_d;
    *hsp_len = left_d + right_d;
    return left_score > right_score ? left_score : right_score;
}

// Lengths of the wave's task and deferred-hit lists.  Wave-uniform, so they live in scalar
// registers; entries are placed by ballot prefix counts.
struct Lists {
    uint32_t nt = 0;  // tasks
    uint32_t nd = 0;  // deferred hits
};

__device__ __forceinline__ uint32_t lanes_below(uint64_t m) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// Two-hit state machine for the hits of the lanes with win set, each (q, s) on its own diagonal
// whose slot holds st (NCBI's word finder, reference restatement gpuBLAZE.cu:1025-1049):
//   pending diagonal: the hit waits behind the queued extension (deferred list);
//   flag (a right extension reached barrier - 1): the hit is skipped unless it lies at or past
//     the barrier, which clears the flag and makes it the last hit;
//   otherwise, with diff the distance to the diagonal's last hit: diff >= window makes it the
//     last hit, diff < word overlaps the last hit and is skipped, anything between becomes an
//     extension task and marks the diagonal pending.
// Computed without branches; every lane of the wave calls it.
template <class R>
__device__ __forceinline__ void apply_hits(const Ctx &c, WaveState<R> &w, bool win, int q, int s, uint32_t slot,
                                           uint32_t st, Lists &l, uint32_t w0) {
    constexpr uint32_t kFlag = RingFmt<R>::kFlag, kPending = RingFmt<R>::kPending, kPosMask = RingFmt<R>::kPosMask;
    const uint32_t sloc = (uint32_t)s;  // offset in the window starting at w0
    const uint32_t sabs = sloc + w0;
    const bool pnd = win && (st & kPending) != 0u;
    const bool plain = win && (st & (kPending | kFlag)) == 0u;
    const uint32_t diff = st ? sabs - (st - 1u) : 0x7FFFFFFFu;
    const bool far = diff >= (uint32_t)c.window;
    const bool tsk = plain && !far && diff >= (uint32_t)kWord;
    const bool rec = (plain && far) || (win && (st & kFlag) != 0u && sabs >= (st & kPosMask) - 1u);
    const uint64_t dm = __ballot(pnd), tm = __ballot(tsk);
    const uint32_t t = l.nt + lanes_below(tm);
    if (pnd) w.defer[l.nd + lanes_below(dm)] = (uint32_t)q | (sloc << 16);
    if (tsk) w.task[t] = Task{(uint32_t)q | (slot << 16), sloc | ((st - 1u + kWord - w0 + kEndBias) << 16)};
    if (tsk || rec) w.ring[slot] = (R)(tsk ? (kPending | t) : sabs + 1u);
    l.nd += (uint32_t)__popcll(dm);
    l.nt += (uint32_t)__popcll(tm);
}

// One claim pass: of the pending lanes whose diagonals hash to the same claim entry, the
// earliest rank wins and applies its hit.  The claim and the diagonal's state are read in the
// same pass: a winner's state is current, since no lane writes a ring slot before the claims
// are settled.
//
// The lanes that lost then read their diagonal's state as the winners left it, and settle their
// hit at once where it is skipped or deferred.  This is exact for every losing hit, whether an
// earlier hit of its diagonal lost too or not: against one state, the skipped hits of a
// diagonal are its earliest (a plain state skips diff < min(word, window), and the distance
// grows with the position; a flag skips positions before the barrier), a skipped hit leaves
// the state as it was, and on a pending diagonal every hit is deferred, in rank order.  The
// first hit of a diagonal that would change its state is left, with every later one, for the
// next pass.
// Returns true for the lanes whose hit is still to be applied.
template <class R>
__device__ __forceinline__ bool claim_pass(const Ctx &c, WaveState<R> &w, bool pend, int q, int s, uint32_t slot,
                                           uint32_t rank, Lists &l, uint32_t w0) {
    constexpr uint32_t kFlag = RingFmt<R>::kFlag, kPending = RingFmt<R>::kPending, kPosMask = RingFmt<R>::kPosMask;
    // every lane takes part, without branches: an idle lane claims an entry of its own with the
    // largest rank (no effect on the minimum), reads a ring slot that lies in the ring, and all
    // lanes free their entries once the claims are read, so the table is free again after the pass
    const uint32_t h = (pend ? slot : (uint32_t)lane_id()) & c.claim_mask;
    atomicMin(&w.claim[h], pend ? rank : 0xFFFFFFFFu);
    const uint32_t st = w.ring[slot];
    ring_sync<R>();
    const bool win = pend & (w.claim[h] == rank);
    ring_sync<R>();
    w.claim[h] = 0xFFFFFFFFu;
    apply_hits(c, w, win, q, s, slot, st, l, w0);
    ring_sync<R>();
    // (every lane's slot lies in the ring, so all of them read it; bitwise tests, no branches)
    const bool lost = pend & !win;
    const uint32_t st2 = w.ring[slot];
    const uint32_t sabs = (uint32_t)s + w0;
    const uint32_t kind = st2 & (kPending | kFlag);
    const bool pnd = lost & (kind == kPending);
    const bool skip = lost & (st2 != 0u) &
                      (((kind == 0u) & (sabs - (st2 - 1u) < (uint32_t)min(kWord, c.window))) |
                       ((kind == kFlag) & (sabs < (st2 & kPosMask) - 1u)));
    const uint64_t dm = __ballot(pnd);
    if (pnd) w.defer[l.nd + lanes_below(dm)] = (uint32_t)q | ((uint32_t)s << 16);
    l.nd += (uint32_t)__popcll(dm);
    return lost && !pnd && !skip;
}

// Applies up to one hit per lane; same-diagonal hits go earliest-rank first.
template <class R>
__device__ __forceinline__ void resolve_round(const Ctx &c, WaveState<R> &w, bool pend, int q, int s, uint32_t rank,
                                              Lists &l, uint32_t w0) {
    const uint32_t slot = ring_slot(c, q, s);
    while (__ballot(pend)) pend = claim_pass(c, w, pend, q, s, slot, rank, l, w0);
}

// Extends every queued pair (one lane each), writes the outcome to its diagonal, then replays
// the hits queued behind pending diagonals.  With drain (the end of a scan or window) it repeats
// until nothing is pending.  Otherwise it runs full walk passes only (the oldest 64 tasks) and
// stops once both lists are back to 64 entries or fewer: the other tasks, and those a replay
// creates, wait for the next flush and fill its walk pass (a pending task with hits deferred
// behind it is the scan's normal state).
// kGrouped: the tasks belong to the subjects of w.grp (ng of them); a task's subject follows from
// its virtual position, its letters from the subject's place in the buffer.  Staged init hits
// carry their subject's index in bits 24.. of len and subject-relative offsets.
template <bool kPad, bool kGrouped, bool kGlobal, class R, class SP>
__device__ void flush(const Ctx &c, WaveState<R> &w, SP S, int slen, int ng, uint32_t &n_ext, Lists &l, uint32_t w0,
                      bool drain) {
    constexpr uint32_t kFlag = RingFmt<R>::kFlag;
    const int lane = lane_id();
    for (;;) {
        ring_sync<R>();
        const uint32_t nt = l.nt;
        const uint32_t nd = l.nd;
        if (drain ? nt == 0 && nd == 0 : nt <= 64u && nd <= 64u) break;
        // in the middle of a scan one full walk pass of the oldest 64 tasks, while the replay
        // below cannot overfill the task list (it adds at most one task per replayed hit)
        const uint32_t ne = !drain && nt > 64u && nd <= 64u ? 64u : nt;
        for (uint32_t t0 = 0; t0 < ne; t0 += 64) {
            const uint32_t t = t0 + lane;
            if (t < ne) {
                const Task tk = w.task[t];
                const int q = (int)(tk.q_slot & 0xFFFFu);
                const uint32_t slot = tk.q_slot >> 16;
                const int ts = (int)(tk.s_pack & 0xFFFFu) + (int)w0;  // absolute
                int sub = 0, shift = 0, rel = 0;  // subject in the group, virtual shift, first letter
                SP St = S;                        // its letters
                int slen_t = slen;
                if constexpr (kGrouped) {
                    for (int j = 1; j < ng; j++) sub += ts >= (int)w.grp->vstart[j];
                    if constexpr (kGlobal) {  // letters in global memory, positions from the subject's start
                        shift = (int)w.grp->vstart[sub];
                        St = S + reinterpret_cast<const GlobalGroup *>(w.grp)->o0[sub];
                        slen_t = (int)w.grp->len[sub];
                    } else {
                        shift = sub * c.gap;
                        rel = (int)w.grp->start[sub];
                    }
                }
                int hq, hs, hl, rext, slast;
                const int s_left = (int)(tk.s_pack >> 16) - (int)kEndBias + (int)w0;
                const int score = two_hit_walk<kPad>(c, St, slen_t, s_left - shift, ts - shift, q, &hq, &hs, &hl, &rext,
                                                     &slast);
                n_ext++;
                if (score >= c.cutoff) {
                    const uint32_t k = atomicAdd(&w.cnt[2], 1u);
                    if (k < (uint32_t)c.stage_cap) c.stage[k] = InitHit{hq, hs - rel, hl | (sub << 24), score};
                }
                w.ring[slot] = (R)(rext ? (((uint32_t)(slast + shift) - 2u + 1u) | kFlag) : ((uint32_t)ts + 1u));
            }
        }
        ring_sync<R>();
        if (ne < nt) {  // the other tasks stay queued, moved to the front (a ring entry only says "pending")
            const bool mv = (uint32_t)lane < nt - ne;
            const Task tk = mv ? w.task[ne + lane] : Task{0u, 0u};
            wave_sync();
            if (mv) w.task[lane] = tk;
            wave_sync();
        }
        l.nt = nt - ne;
        // replay the queued hits [0, nd) in order; new deferrals land at nd and beyond
        for (uint32_t r = 0; r < nd; r += 64) {
            const uint32_t i = r + lane;
            int q = 0, s = 0;
            const bool pend = i < nd;
            if (pend) {
                const uint32_t d = w.defer[i];
                q = (int)(d & 0xFFFFu);
                s = (int)(d >> 16);
            }
            resolve_round(c, w, pend, q, s, i, l, w0);
        }
        ring_sync<R>();
        const uint32_t nd2 = l.nd;
        for (uint32_t i = nd + lane; i < nd2; i += 64) w.defer[i - nd] = w.defer[i];
        ring_sync<R>();
        l.nd = nd2 - nd;
    }
}

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

// Global-memory view of a table pointer: loads become global loads (no flat address-space test,
// and LDS waits do not wait on them).
template <class T>
__device__ __forceinline__ const __attribute__((address_space(1))) T *gmem(const T *p) {
    return (const __attribute__((address_space(1))) T *)p;
}

// Lookup tables of the current query, wherever kTable puts them.
struct Tables {
    const uint16_t *rel;  // LDS (kTable 2)
    const uint32_t *bb;   // LDS (kTable 2)
    const uint32_t *gprefix;
    const uint16_t *ent;  // LDS (kTable >= 1) or global
    const uint4 *rec;     // word records in global memory (kTable 0)
};

// Lookup entry i: a global load for kTable 0, an LDS load otherwise.
template <int kTable>
__device__ __forceinline__ uint32_t entry(const Tables &t, uint32_t i) {
    if constexpr (kTable == 0) return gmem(t.ent)[i];
    return t.ent[i];
}

// Lookup entries [pa, pb) of word wd.
template <int kTable>
__device__ __forceinline__ void word_range(const Tables &t, uint32_t wd, uint32_t &pa, uint32_t &pb) {
    if (kTable == 2) {
        pa = t.bb[wd >> 6] + t.rel[wd];
        pb = t.bb[(wd + 1) >> 6] + t.rel[wd + 1];
    } else {
        pa = t.gprefix[wd];
        pb = t.gprefix[wd + 1];
    }
}

// Seed scan of one subject whose letters are read through S (LDS buffer or global memory).
// Seed scan of one subject (slen letters read through S), or with kGrouped of the ng subjects of
// w.grp in the wave's LDS buffer S.
//
// Hits wait in a queue (w.hits[head, tail), entries q << 16 | position, subject order) and go
// through the two-hit state 64 at a time, so a round is full even when a step yields a few more
// than 64 hits.  A round makes one claim pass; the hits that lost their claim go back to the
// front of the queue, in order, and lead the next round: per diagonal the order stays the
// subject order, since a lost hit is behind its diagonal's winner and ahead of every later hit.
// A step whose hits exceed kHitBuf (low-complexity letters) is expanded into the queue 64 hits
// at a time, each hit finding its position by a prefix max over the marked starts of the lanes'
// ranges.
// The loop is one state machine with a single site for each of its phases (claim round with its
// extension flush, dense expansion, word step), which keeps the inlined code and its registers
// small.
// A subject longer than kWindowFrom letters is scanned in windows of word positions, each a
// multiple of the diagonal ring's length long (at most kWindowMax), so that positions in the
// queue and the lists stay 16-bit offsets from the window's first position while a diagonal keeps
// its ring slot; the ring holds absolute positions and carries over from window to window.  Each
// window ends with its queue drained and a flush, which only extends pairs earlier than a later
// flush would: the result does not depend on when flushes happen.
constexpr int kWindowFrom = 60000;
constexpr int kWindowMax = 32768;
template <int kTable, bool kPad, bool kGrouped = false, bool kGlobal = false, class R, class SP>
__device__ void scan_subject(const Ctx &c, WaveState<R> &w, const Tables &t, SP S, int slen, uint32_t &n_hits,
                             uint32_t &n_ext, int ng = 1) {
    const int lane = lane_id();
    Lists l;                              // the task and deferred lists (empty between subjects)
    uint32_t head = 0, tail = 0;
    int gi = 0;                           // subject of the group
    int p_begin = 0, p_end = slen, voff = 0;  // its letters [p_begin, p_end), virtual shift
    SP Sc = S;                            // kGlobal: the current subject's letters
    auto subject = [&](int i) {           // subject i of the group
        if constexpr (kGlobal) {
            p_begin = 0;
            p_end = __builtin_amdgcn_readfirstlane((int)w.grp->len[i]);
            voff = __builtin_amdgcn_readfirstlane((int)w.grp->vstart[i]);
            Sc = S + first_lane64(reinterpret_cast<const GlobalGroup *>(w.grp)->o0[i]);
        } else {
            p_begin = __builtin_amdgcn_readfirstlane((int)w.grp->start[i]);
            p_end = p_begin + __builtin_amdgcn_readfirstlane((int)w.grp->len[i]);
            voff = i * c.gap;
        }
    };
    if constexpr (kGrouped) subject(0);
    // only the launches with 32-bit rings in global memory (long subjects, the retry scan) take
    // subjects past kWindowFrom; the others compile without windows
    constexpr bool kWindows = sizeof(R) == 4 && !kGrouped;
    uint32_t w0 = 0;                      // first word position of the window
    int win = 0, w_end = 0x7FFFFFFF;      // window length (0: one window), end of the window
    if (kWindows && slen > kWindowFrom) {
        win = (int)(c.ring_len * ((uint32_t)kWindowMax / c.ring_len));
        w_end = win;
    }
    int c0 = p_begin;                     // next step
    uint32_t dense_h = 0, dense_r = 0;    // dense step in progress: its hits, those expanded
    int dense_c0 = 0, dense_voff = 0;
    bool dense_pf = false;                // the next pass's owner lanes and entries are loaded
    uint32_t pf_q = 0, pf_lo = 0;         // (this lane's hit of that pass: entry, owner lane)
  for (;;) {                              // windows
    bool input = true;                    // steps (or dense hits) remain
    for (;;) {
        // a claim round over the queue's first 64 hits, after the extensions when the task or
        // deferred lists fill
        if (tail - head >= 64u || (!input && tail > head)) {
            if (l.nt > (uint32_t)(kTaskCap - 64) || l.nd > 64u) flush<kPad, kGrouped, kGlobal>(c, w, S, slen, ng, n_ext, l, w0, false);
            const uint32_t n = tail - head < 64u ? tail - head : 64u;
            const bool pend = (uint32_t)lane < n;
            const uint32_t e0 = w.hits[head + lane];  // (past the queue: the task list, in the wave's LDS)
            const uint32_t e = pend ? e0 : 0u;
            const int q = (int)(e >> 16), sp = (int)(e & 0xFFFFu);
            const bool lost = claim_pass(c, w, pend, q, sp, ring_slot(c, q, sp), (uint32_t)lane, l, w0);
            const uint64_t lm = __ballot(lost);
            head += n - (uint32_t)__popcll(lm);
            // (the others store to the deferred list's last 64 words, unused here: at most 128
            // deferred hits after a round)
            *(volatile __attribute__((address_space(3))) uint32_t *)(lost ? w.hits + head + lanes_below(lm)
                                                                           : w.defer + kDeferCap - 64 + lane) = e;
            wave_sync();
            continue;
        }
        if (!input) break;
        // fewer than 64 hits wait; room for `need` more: move the waiting hits to the front
        auto make_room = [&](uint32_t need) {
            if (tail + need > (uint32_t)kHitQueue) {
                const uint32_t rest = tail - head;
                const uint32_t v = (uint32_t)lane < rest ? w.hits[head + lane] : 0u;
                wave_sync();
                if ((uint32_t)lane < rest) w.hits[lane] = v;
                wave_sync();
                head = 0;
                tail = rest;
            }
        };
        // 64 more hits of a dense step: hit g belongs to the last position lane whose range of
        // the step's hits, [excl[i], excl[i + 1]), starts at or before g.  Each lane with hits
        // marks the slot where its range starts in this pass's 64 (slot 0 for a range begun
        // earlier) with an LDS max, in the queue slots the pass is about to fill, and a prefix
        // max over the slots gives every hit its lane.
        if (dense_r < dense_h) {
            make_room(64u);
            // owner lanes of the pass starting at hit g0, marked in mark[0, 64)
            auto owners = [&](uint32_t g0, uint32_t *mark) {
                mark[lane] = 0u;
                wave_sync();
                const uint32_t ex_i = w.excl[lane];
                if (w.excl[lane + 1] > ex_i && ex_i < g0 + 64u) atomicMax(&mark[ex_i > g0 ? ex_i - g0 : 0u], (uint32_t)lane + 1u);
                wave_sync();
                const int lo = wave_inclusive_max((int)mark[lane] - 1);  // slot 0 always holds a mark
                wave_sync();
                return lo;
            };
            const uint32_t g = dense_r + lane;
            const uint32_t n = dense_h - dense_r < 64u ? dense_h - dense_r : 64u;
            uint32_t qo = pf_q, lo = pf_lo;
            if (!dense_pf) {
                lo = (uint32_t)owners(dense_r, w.hits + tail);
                qo = g < dense_h ? entry<kTable>(t, w.p0[lo] + (g - w.excl[lo])) : 0u;
            }
            if (g < dense_h) w.hits[tail + lane] = (qo << 16) | (uint32_t)(dense_c0 + (int)lo + dense_voff);
            // the next pass's owners (marked in the claim table's first 64 entries, free between
            // claim passes, then freed again) and its entry loads, in flight while rounds run
            const uint32_t g1 = dense_r + n;
            dense_pf = g1 < dense_h;
            if (dense_pf) {
                pf_lo = (uint32_t)owners(g1, w.claim);
                w.claim[lane] = 0xFFFFFFFFu;
                pf_q = g1 + lane < dense_h ? entry<kTable>(t, w.p0[pf_lo] + (g1 + lane - w.excl[pf_lo])) : 0u;
            }
            wave_sync();
            tail += n;
            dense_r += n;
            continue;
        }
        // the next step of 64 words
        if (c0 + kWord > p_end || c0 >= w_end) {  // this subject (window) is done: the group's next, or the end
            if constexpr (kGrouped) {
                if (gi + 1 < ng) {
                    gi++;
                    subject(gi);
                    c0 = p_begin;
                    continue;
                }
            }
            input = false;
            continue;
        }
        const int p = c0 + lane;
        const int s = p + voff - (int)w0;  // position seen by the two-hit state, in the window
        uint32_t pa = 0, nh = 0;
        u32x4 r = {0u, 0u, 0u, 0u};  // kTable 0: the word's record
        if constexpr (kTable == 0) {
            // every lane loads a record, lanes past the last word that of the last word (clamped
            // position); their count is then taken as 0
            const int p_last = min(p_end - kWord, w_end - 1);
            const int pc = min(p, p_last);
            const uint32_t w2 = Sc[pc] * 784u + Sc[pc + 1] * 28u + Sc[pc + 2];
            const uint32_t wd = w2 >> (kPad ? 1 : 0);
            r = *reinterpret_cast<const __attribute__((address_space(1))) u32x4 *>(
                gmem(reinterpret_cast<const uint8_t *>(t.rec)) + (w2 << (kPad ? 3 : 4)));
            nh = p <= p_last ? r.x & 0xFFu : 0u;
            pa = r.x >> 8;
            if (nh == 0xFFu) {  // 255 entries or more, or a first entry past 2^24
                pa = gmem(t.gprefix)[wd];
                nh = gmem(t.gprefix)[wd + 1] - pa;
            }
        } else if (p + kWord <= p_end && p < w_end) {
            // (letters in the LDS buffer are doubled, and so is their word index: a 16-byte
            // record's byte offset is then the index times 8)
            const uint32_t w2 = Sc[p] * 784u + Sc[p + 1] * 28u + Sc[p + 2];
            const uint32_t wd = w2 >> (kPad ? 1 : 0);
            if constexpr (kTable == 0) {
                r = *reinterpret_cast<const __attribute__((address_space(1))) u32x4 *>(
                    gmem(reinterpret_cast<const uint8_t *>(t.rec)) + (w2 << (kPad ? 3 : 4)));
                nh = r.x & 0xFFu;
                pa = r.x >> 8;
                if (nh == 0xFFu) {  // 255 entries or more, or a first entry past 2^24
                    pa = gmem(t.gprefix)[wd];
                    nh = gmem(t.gprefix)[wd + 1] - pa;
                }
            } else {
                uint32_t pb;
                word_range<kTable>(t, wd, pa, pb);
                nh = pb - pa;
            }
        }
        const uint32_t inc = wave_inclusive_sum(nh);
        const uint32_t H = last_lane(inc);
        const uint32_t ex = inc - nh;
        n_hits += nh;
        if (H <= (uint32_t)kHitBuf) {
            make_room(H);
            // append: the record's inline entries, then four independent lookup loads per pass
            // (indices clamped to the word's last entry).
            // The six inline entries are stored unconditionally, highest index first, through one
            // base address (a lane without hits uses its own word of w.excl, idle outside dense
            // steps).  A store past a lane's nh entries lands either past the step's H hits (in
            // the queue's slack) or on entry j' < j of a later lane, whose own store of entry j'
            // comes later in program order and overwrites it; two lanes with hits never share an
            // address in one store, since their lists start at distinct offsets.
            uint32_t *const dst = w.hits + tail + ex;
            uint32_t k0 = 0;
            if constexpr (kTable == 0) {
                const uint32_t su = (uint32_t)s;
                const uint32_t v[kWordRecInline] = {(r.y << 16) | su, (r.y & 0xFFFF0000u) | su,
                                                    (r.z << 16) | su, (r.z & 0xFFFF0000u) | su,
                                                    (r.w << 16) | su, (r.w & 0xFFFF0000u) | su};
                // volatile (LDS): the six stores stay separate and in this order
                typedef volatile __attribute__((address_space(3))) uint32_t lds_u32;
                lds_u32 *const d = (lds_u32 *)(nh ? dst : w.excl + lane);
#pragma unroll
                for (int j = kWordRecInline - 1; j >= 0; j--) d[j] = v[j];
                k0 = kWordRecInline;
            }
            for (uint32_t k = k0; k < nh; k += 4) {
                uint32_t e[4];
#pragma unroll
                for (int j = 0; j < 4; j++) e[j] = entry<kTable>(t, pa + min(k + j, nh - 1));
                // stores past the lane's entries go to its own idle word of w.excl (no branches,
                // so the four loads stay ahead of the stores)
                typedef volatile __attribute__((address_space(3))) uint32_t lds_u32x;
#pragma unroll
                for (uint32_t j = 0; j < 4; j++)
                    *(lds_u32x *)(k + j < nh ? dst + k + j : w.excl + lane) = (e[j] << 16) | (uint32_t)s;
            }
            wave_sync();
            tail += H;
        } else {
            // dense step: its hits enter the queue 64 at a time (above)
            w.excl[lane] = ex;
            if (lane == 0) w.excl[64] = H;
            w.p0[lane] = pa;
            wave_sync();
            dense_h = H;
            dense_r = 0;
            dense_pf = false;
            dense_c0 = c0;
            dense_voff = voff - (int)w0;
        }
        c0 += 64;
    }
    flush<kPad, kGrouped, kGlobal>(c, w, S, slen, ng, n_ext, l, w0, true);
    if (!kWindows || win == 0 || w_end + kWord > p_end) break;
    w0 = (uint32_t)w_end;  // the next window; the ring carries over
    w_end = w_end + win < p_end ? w_end + win : p_end;
    c0 = (int)w0;
  }
}

// The subject's init hits (staged by this wave) go to the pool, one work item per query of
// the item's pack that has any.  Query offsets become offsets in that query.
template <class R>
__device__ void emit_init_hits(const SeedArgs &a, const QueryDev &qd, WaveState<R> &w, const Ctx &c, int64_t sid,
                               int slen, uint32_t nih, bool count_init = true) {
    const int lane = lane_id();
    const int nm = qd.n_members;
    auto overflow = [&](int qi) {
        if (lane == 0) {
            const uint32_t k = atomicAdd(&a.ctr->overflow_count, 1u);
            if (k < a.ovf_cap) a.ovf[k] = ((int64_t)qi << kOvfShift) | sid;
        }
    };
    if (nih > (uint32_t)a.stage_cap) {
        for (int m = 0; m < nm; m++) overflow((int)qd.member_idx[m]);
        return;
    }
    unsigned long long base = 0;
    if (lane == 0) base = atomicAdd(&a.ctr->pool_used, (unsigned long long)nih);
    base = first_lane64(base);
    if (base + nih > a.pool_cap) {
        for (int m = 0; m < nm; m++) overflow((int)qd.member_idx[m]);
        return;
    }
    if (nm == 1) {
        int best = 0;
        for (uint32_t k = lane; k < nih; k += 64) {
            const InitHit h = c.stage[k];
            uint32_t r = k;
            if (nih > 32) {  // long lists go out in the gapped stage's order (its sort is then linear)
                r = 0;
                for (uint32_t j = 0; j < nih; j++) {
                    const InitHit o = c.stage[j];
                    r += (uint32_t)(init_before(o, h) || (j < k && !init_before(h, o)));
                }
            }
            a.pool[base + r] = h;
            best = h.score > best ? h.score : best;
        }
        best = (int)last_lane((uint32_t)wave_inclusive_max(best));
        if (lane == 0) {
            const int qi = (int)qd.member_idx[0];
            const uint32_t wi = atomicAdd(&a.ctr->work_count, 1u);
            if (wi < a.work_cap) {
                const int bucket = cost_bucket(best, (int)nih, slen);
                a.work[wi] = WorkItem{sid, (uint64_t)base, (int32_t)nih, qi, bucket, 0};
                atomicAdd(&a.ctr->bucket_count[bucket], 1u);
                if (count_init) atomicAdd(&a.ctr->init_hits, (unsigned long long)nih);
            } else {
                const uint32_t k = atomicAdd(&a.ctr->overflow_count, 1u);
                if (k < a.ovf_cap) a.ovf[k] = ((int64_t)qi << kOvfShift) | sid;
            }
        }
        return;
    }
    // a pack of several queries: count, lay out and relabel the hits per query
    uint32_t *cnt = w.p0;          // [kMaxMembers] counts, then cursors
    uint32_t *bestm = w.excl;      // [kMaxMembers] best score per query
    uint32_t *offm = w.excl + 32;  // [kMaxMembers] first slot per query
    if (lane < nm) {
        cnt[lane] = 0;
        bestm[lane] = 0;
    }
    wave_sync();
    auto member_of = [&](int q) {
        int m = 0;
        while (m + 1 < nm && (int)qd.member_start[m + 1] <= q) m++;
        return m;
    };
    for (uint32_t k = lane; k < nih; k += 64) {
        const InitHit h = c.stage[k];
        const int m = member_of(h.q_start);
        atomicAdd(&cnt[m], 1u);
        atomicMax(&bestm[m], (uint32_t)h.score);
    }
    wave_sync();
    const uint32_t mine = lane < nm ? cnt[lane] : 0u;
    const uint32_t off = wave_inclusive_sum(mine) - mine;
    wave_sync();
    if (lane < nm) {
        offm[lane] = off;
        cnt[lane] = 0;
    }
    wave_sync();
    for (uint32_t k = lane; k < nih; k += 64) {
        InitHit h = c.stage[k];
        const int m = member_of(h.q_start);
        const uint32_t pos = offm[m] + atomicAdd(&cnt[m], 1u);
        h.q_start -= (int)qd.member_start[m];
        a.pool[base + pos] = h;
    }
    wave_sync();
    if (lane < nm && mine > 0) {
        const int qi = (int)qd.member_idx[lane];
        const uint32_t wi = atomicAdd(&a.ctr->work_count, 1u);
        if (wi < a.work_cap) {
            const int bucket = cost_bucket((int)bestm[lane], (int)mine, slen);
            a.work[wi] = WorkItem{sid, (uint64_t)(base + off), (int32_t)mine, qi, bucket, 0};
            atomicAdd(&a.ctr->bucket_count[bucket], 1u);
            atomicAdd(&a.ctr->init_hits, (unsigned long long)mine);
        } else {
            const uint32_t k = atomicAdd(&a.ctr->overflow_count, 1u);
            if (k < a.ovf_cap) a.ovf[k] = ((int64_t)qi << kOvfShift) | sid;
        }
    }
}

// A group's init hits (staged by this wave, tagged with their subject) go to the pool, one work
// item per (subject, query of the item's pack) that has any.  Query offsets become offsets in
// that query.  Bins are subject-major; their counts, best scores and first slots sit in the
// wave's task, deferred and hit-queue lists, all empty once the group is scanned, so a group
// holds at most kHitQueue / n_members subjects.
template <class R>
__device__ void emit_group(const SeedArgs &a, const QueryDev &qd, WaveState<R> &w, const Ctx &c, int ng, uint32_t nih) {
    const int lane = lane_id();
    const int nm = qd.n_members;
    const int nb = ng * nm;
    auto record = [&](int bin) {
        const uint32_t k = atomicAdd(&a.ctr->overflow_count, 1u);
        if (k < a.ovf_cap) a.ovf[k] = ((int64_t)qd.member_idx[bin % nm] << kOvfShift) | w.grp->sid[bin / nm];
    };
    if (nih > (uint32_t)a.stage_cap) {
        for (int bin = lane; bin < nb; bin += 64) record(bin);
        return;
    }
    unsigned long long base = 0;
    if (lane == 0) base = atomicAdd(&a.ctr->pool_used, (unsigned long long)nih);
    base = first_lane64(base);
    if (base + nih > a.pool_cap) {
        for (int bin = lane; bin < nb; bin += 64) record(bin);
        return;
    }
    uint32_t *cnt = reinterpret_cast<uint32_t *>(w.task);  // [nb] counts, then cursors
    uint32_t *bestm = w.defer;                             // [nb] best score
    uint32_t *offm = w.hits;                               // [nb] first slot
    for (int bin = lane; bin < nb; bin += 64) {
        cnt[bin] = 0;
        bestm[bin] = 0;
    }
    wave_sync();
    auto bin_of = [&](const InitHit &h, int &m) {
        m = 0;
        while (m + 1 < nm && (int)qd.member_start[m + 1] <= h.q_start) m++;
        return (h.len >> 24) * nm + m;
    };
    for (uint32_t k = lane; k < nih; k += 64) {
        const InitHit h = c.stage[k];
        int m;
        const int bin = bin_of(h, m);
        atomicAdd(&cnt[bin], 1u);
        atomicMax(&bestm[bin], (uint32_t)h.score);
    }
    wave_sync();
    uint32_t run = 0;  // exclusive scan of the counts, 64 bins per pass
    for (int b0 = 0; b0 < nb; b0 += 64) {
        const uint32_t mine = b0 + lane < nb ? cnt[b0 + lane] : 0u;
        const uint32_t inc = wave_inclusive_sum(mine);
        if (b0 + lane < nb) offm[b0 + lane] = run + inc - mine;
        run += last_lane(inc);
    }
    wave_sync();
    for (int bin = lane; bin < nb; bin += 64) cnt[bin] = 0;
    wave_sync();
    for (uint32_t k = lane; k < nih; k += 64) {
        InitHit h = c.stage[k];
        int m;
        const int bin = bin_of(h, m);
        h.len &= 0xFFFFFF;
        h.q_start -= (int)qd.member_start[m];
        a.pool[base + offm[bin] + atomicAdd(&cnt[bin], 1u)] = h;
    }
    wave_sync();
    for (int bin = lane; bin < nb; bin += 64) {
        const uint32_t n = cnt[bin];
        if (n == 0) continue;
        const int qi = (int)qd.member_idx[bin % nm];
        const int64_t sid = w.grp->sid[bin / nm];
        const uint32_t wi = atomicAdd(&a.ctr->work_count, 1u);
        if (wi < a.work_cap) {
            const int bucket = cost_bucket((int)bestm[bin], (int)n, (int)w.grp->len[bin / nm]);
            a.work[wi] = WorkItem{sid, (uint64_t)(base + offm[bin]), (int32_t)n, qi, bucket, 0};
            atomicAdd(&a.ctr->bucket_count[bucket], 1u);
            atomicAdd(&a.ctr->init_hits, (unsigned long long)n);
        } else {
            const uint32_t k = atomicAdd(&a.ctr->overflow_count, 1u);
            if (k < a.ovf_cap) a.ovf[k] = ((int64_t)qi << kOvfShift) | sid;
        }
    }
    wave_sync();
}

}  // namespace

// kLong: the launch over the subjects longer than kNarrowMax (a.long_ids), with diagonal rings
// of 32-bit entries in global memory; the other launches skip those subjects.
// At least six waves per SIMD (80 VGPRs): the kernel is bound by LDS latency, so resident
// waves are what hides it.
template <int kTable, bool kLong>
__global__ void __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(6, 8))) k_seed2(SeedArgs a) {
    extern __shared__ __align__(16) uint8_t smem[];
    const int lane = lane_id();
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // uniform: per-wave LDS bases live in SGPRs
    const int nwaves = blockDim.x >> 6;
    // LDS is sized for the largest query of the launch (a.qlen, a.ring_len, a.claim_mask,
    // a.n_entries); each item uses its own query's sizes
    const uint32_t M = a.ring_len;
    const uint32_t C = a.claim_mask + 1;

    uint32_t off = 0;
    int32_t *blk = reinterpret_cast<int32_t *>(smem);  // [0] item, [1] next subject of the item
    off += 16;
    int16_t(*mat)[kMatStride] = reinterpret_cast<int16_t(*)[kMatStride]>(smem + off);
    off += al16(kMatBytes);
    uint8_t *Qs = smem + off + kSeqPad;  // kSeqPad zero letters on both sides
    off += kSeqPad + al16((uint32_t)a.qlen + kSeqPad);
    uint16_t *Qx = reinterpret_cast<uint16_t *>(smem + off) + kSeqPad;  // the same letters times 2 * kMatStride
    off += al16(2 * ((uint32_t)a.qlen + 2 * kSeqPad));
    uint16_t *ent_l = nullptr;
    uint16_t *rel = nullptr;
    uint32_t *bb = nullptr;
    if (kTable >= 1) {
        ent_l = reinterpret_cast<uint16_t *>(smem + off);
        off += al16(a.n_entries * 2 + 4);
    }
    if (kTable >= 2) {
        rel = reinterpret_cast<uint16_t *>(smem + off);
        off += al16((kWords + 1) * 2);
        bb = reinterpret_cast<uint32_t *>(smem + off);
        off += al16(kRelBlocks * 4);
    }
    uint8_t *wb = smem + off + (size_t)wave * wave_lds_bytes(M, C);
    WaveState<uint16_t> w;  // LDS ring of 16-bit entries
    w.ring = reinterpret_cast<uint16_t *>(wb);
    w.claim = reinterpret_cast<uint32_t *>(w.ring + M);
    w.excl = w.claim + C;
    w.p0 = w.excl + 68;
    w.cnt = w.p0 + 64;
    w.hits = w.cnt + 4;
    w.task = reinterpret_cast<Task *>(w.hits + kHitQueue + kHitSlack);
    w.defer = reinterpret_cast<uint32_t *>(w.task + kTaskCap);
    w.subj = reinterpret_cast<uint8_t *>(w.defer + kDeferCap) + kSeqPad;  // kSeqPad zero letters on both sides
    w.grp = reinterpret_cast<Group *>(w.subj + kSubjBuf + kSeqPad);

    for (int i = threadIdx.x; i < 28 * kMatStride; i += blockDim.x)
        mat[i / kMatStride][i % kMatStride] = i % kMatStride < 28 ? kBlosum62Dev[i / kMatStride][i % kMatStride] : 0;
    if (lane < kSeqPad) w.subj[lane - kSeqPad] = 0;
    for (uint32_t i = lane; i < C; i += 64) w.claim[i] = 0xFFFFFFFFu;
    if (lane < 4) w.cnt[lane] = 0;

    const uint64_t gwave = (uint64_t)blockIdx.x * nwaves + wave;
    WaveState<uint32_t> wl;  // the same lists with a ring of 32-bit entries in global memory
    wl.ring = kLong ? a.long_ring + gwave * M : nullptr;
    wl.claim = w.claim;
    wl.excl = w.excl;
    wl.p0 = w.p0;
    wl.cnt = w.cnt;
    wl.hits = w.hits;
    wl.task = w.task;
    wl.defer = w.defer;
    wl.subj = w.subj;
    wl.grp = nullptr;
    Ctx c;
    c.mat = mat;
    c.Q = Qs;
    c.Qx = Qx;
    c.stage = a.stage + gwave * (uint64_t)a.stage_cap;
    c.stage_cap = a.stage_cap;
    uint32_t n_hits = 0, n_ext = 0;  // this lane's word hits and extensions
    int cur_q = -1;
    const uint32_t n_items = a.n_items;
    Tables t{};

    for (;;) {
        __syncthreads();  // every wave is done with the previous item (and its query tables)
        if (threadIdx.x == 0) {
            blk[0] = (int32_t)atomicAdd(a.item_ctr, 1u);
            blk[1] = 0;
        }
        __syncthreads();
        const uint32_t item = first_lane((uint32_t)blk[0]);  // uniform: the item's values live in scalars
        if (item >= n_items) break;
        const int qi = (int)a.items[3 * (uint64_t)item];
        const int64_t sub0 = a.items[3 * (uint64_t)item + 1];
        const int64_t sub1 = a.items[3 * (uint64_t)item + 2];
        const QueryDev &qd = a.qd[qi];
        if (qi != cur_q) {  // stage this query
            for (int i = (int)threadIdx.x - kSeqPad; i < qd.qlen + kSeqPad; i += blockDim.x) {
                const uint8_t l = i >= 0 && i < qd.qlen ? qd.query[i] : 0;
                Qs[i] = l;
                Qx[i] = (uint16_t)(l * (2 * kMatStride));
            }
            if (kTable >= 1) {
                const uint32_t *src = reinterpret_cast<const uint32_t *>(qd.entries);
                uint32_t *dst = reinterpret_cast<uint32_t *>(ent_l);
                for (uint32_t i = threadIdx.x; i < (qd.n_entries + 1) / 2; i += blockDim.x) dst[i] = src[i];
            }
            if (kTable >= 2) {
                for (int i = threadIdx.x; i <= kWords; i += blockDim.x) rel[i] = qd.rel16[i];
                for (int i = threadIdx.x; i < kRelBlocks; i += blockDim.x) bb[i] = qd.blockbase[i];
            }
            __syncthreads();
            cur_q = qi;
        }
        c.qlen = qd.qlen;
        c.ring_len = qd.ring_len;
        c.ring_base = qd.ring_base;
        c.ring_magic = qd.ring_magic;
        c.claim_mask = qd.claim_mask;
        c.window = qd.window;
        c.xdrop = qd.x_ungapped;
        c.cutoff = qd.ungapped_cutoff;
        c.gap = kLong ? 0 : qd.group_gap;
        t.rel = rel;
        t.bb = bb;
        // table pointers in scalar registers: uniform, and not worth a vector register each
        t.gprefix = reinterpret_cast<const uint32_t *>(first_lane64(reinterpret_cast<uint64_t>(qd.prefix)));
        t.ent = kTable >= 1 ? ent_l
                            : reinterpret_cast<const uint16_t *>(first_lane64(reinterpret_cast<uint64_t>(qd.entries)));
        t.rec = reinterpret_cast<const uint4 *>(first_lane64(reinterpret_cast<uint64_t>(qd.rec)));

        // subjects of up to kSubjBuf letters join the wave's group while they fit its buffer; a
        // full group is scanned as one
        int ng = 0;
        uint32_t used = 0;
        const int group_cap = c.gap > 0 ? min(kGroupMax, kHitQueue / qd.n_members) : 1;
        auto run_group = [&]() {
            uint4 *r4 = reinterpret_cast<uint4 *>(w.ring);
            for (uint32_t i = lane; i < c.ring_len / 8; i += 64) r4[i] = uint4{0, 0, 0, 0};
            wave_sync();
            scan_subject<kTable, true, true>(c, w, t, (const uint8_t *)w.subj, 0, n_hits, n_ext, ng);
            wave_sync();
            const uint32_t nih = w.cnt[2];
            wave_sync();
            if (nih > 0) {
                if (ng == 1 && qd.n_members == 1)
                    emit_init_hits(a, qd, w, c, w.grp->sid[0], (int)w.grp->len[0], nih);  // long lists leave sorted
                else
                    emit_group(a, qd, w, c, ng, nih);
                wave_sync();
                if (lane == 0) w.cnt[2] = 0;
                wave_sync();
            }
            ng = 0;
            used = 0;
        };
        // subjects longer than the buffer (to kNarrowMax letters) form groups of their own, their
        // table in the buffer (run before short subjects need it, and the other way round)
        int ngg = 0;
        uint32_t gv = 0;  // virtual positions the global group spans
        GlobalGroup *const gg = reinterpret_cast<GlobalGroup *>(w.subj);
        auto run_global_group = [&]() {
            uint4 *r4 = reinterpret_cast<uint4 *>(w.ring);
            for (uint32_t i = lane; i < c.ring_len / 8; i += 64) r4[i] = uint4{0, 0, 0, 0};
            wave_sync();
            WaveState<uint16_t> wg = w;
            wg.grp = &gg->g;
            scan_subject<kTable, false, true, true>(c, wg, t, a.res, 0, n_hits, n_ext, ngg);
            wave_sync();
            const uint32_t nih = w.cnt[2];
            wave_sync();
            if (nih > 0) {
                if (ngg == 1 && qd.n_members == 1)
                    emit_init_hits(a, qd, wg, c, gg->g.sid[0], (int)gg->g.len[0], nih);  // long lists leave sorted
                else
                    emit_group(a, qd, wg, c, ngg, nih);
                wave_sync();
                if (lane == 0) w.cnt[2] = 0;
                wave_sync();
            }
            ngg = 0;
            gv = 0;
        };
        for (;;) {
            int64_t k = 0;
            if (lane == 0) k = sub0 + (int64_t)atomicAdd(reinterpret_cast<uint32_t *>(&blk[1]), 1u);
            k = (int64_t)first_lane64((uint64_t)k);
            if (k >= sub1) break;
            const int64_t sid = kLong ? a.long_ids[k] : k;
            const uint64_t o0 = a.offs[sid];
            const int slen = (int)(a.offs[sid + 1] - o0);
            const uint8_t *Sg = a.res + o0;
            const uint32_t Mq = c.ring_len;
            if (!kLong && slen <= kSubjBuf) {  // staged in LDS, alone when groups are off (gap 0)
                if (ngg > 0) run_global_group();  // the buffer holds its table
                if (ng == group_cap || used + (uint32_t)slen > (uint32_t)kSubjBuf) run_group();
                // letters at [used, used + slen), doubled, then zeros up to the next 64-aligned
                // start (at least 4 zeros) or the buffer's end pad: one dword per lane, shifted
                // into place from two aligned loads (the DB keeps 64 zero bytes past its end, and
                // the source offset is clamped to the subject's last letter, so every load stays
                // inside the DB)
                const uint32_t next = (used + (uint32_t)slen + 4u + 63u) & ~63u;
                const uint32_t zend = next < (uint32_t)(kSubjBuf + kSeqPad) ? next : (uint32_t)(kSubjBuf + kSeqPad);
                const auto *R32 = gmem(reinterpret_cast<const uint32_t *>(a.res));
                uint32_t *dst = reinterpret_cast<uint32_t *>(w.subj + used);
                for (uint32_t j = lane; j < (zend - used) >> 2; j += 64) {
                    const int rem = slen - 4 * (int)j;  // letters from this dword on
                    const uint64_t src = o0 + (uint64_t)max(min(4 * (int)j, slen - 1), 0);
                    const uint32_t lo = R32[src >> 2], hi = R32[(src >> 2) + 1];
                    const uint32_t x = (uint32_t)((((uint64_t)hi << 32) | lo) >> (8u * ((uint32_t)src & 3u)));
                    const uint32_t keep = rem >= 4 ? 0xFFFFFFFFu : rem <= 0 ? 0u : (1u << (8 * rem)) - 1u;
                    dst[j] = (x & keep) << 1;  // doubled (letters are below 128)
                }
                if (lane == 0) {
                    w.grp->sid[ng] = sid;
                    w.grp->start[ng] = used;
                    w.grp->len[ng] = (uint32_t)slen;
                    w.grp->vstart[ng] = used + (uint32_t)(ng * c.gap);
                }
                wave_sync();
                ng++;
                used = next;
                continue;
            }
            if (ng > 0) run_group();  // the paths below reuse the buffer
            if constexpr (!kLong) {
                if (slen > kNarrowMax) continue;  // the long-subject launch takes it
                const uint32_t v0 = ngg > 0 ? gv + (uint32_t)c.gap : 0u;
                if (ngg > 0 && (ngg == group_cap || v0 + (uint32_t)slen > (uint32_t)kNarrowMax)) run_global_group();
                const uint32_t vs = ngg > 0 ? gv + (uint32_t)c.gap : 0u;
                if (lane == 0) {
                    gg->g.sid[ngg] = sid;
                    gg->g.start[ngg] = 0;
                    gg->g.len[ngg] = (uint32_t)slen;
                    gg->g.vstart[ngg] = vs;
                    gg->o0[ngg] = o0;
                }
                wave_sync();
                ngg++;
                gv = vs + (uint32_t)slen;
                continue;
            }
            if constexpr (kLong) {
                for (uint32_t i = lane; i < Mq; i += 64) wl.ring[i] = 0;
                ring_sync<uint32_t>();
                scan_subject<kTable, false>(c, wl, t, Sg, slen, n_hits, n_ext);
            }

            // hand the subject's init hits to the gapped stage
            wave_sync();
            const uint32_t nih = w.cnt[2];
            wave_sync();
            if (nih > 0) {
                emit_init_hits(a, qd, w, c, sid, slen, nih);
                wave_sync();
                if (lane == 0) w.cnt[2] = 0;
                wave_sync();
            }
        }
        if (ng > 0) run_group();
        if (ngg > 0) run_global_group();
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

size_t seed2_lds_bytes(const SeedArgs &a, int table, int waves) {
    return block_lds_bytes(table, (uint32_t)a.qlen, a.n_entries) +
           (size_t)waves * wave_lds_bytes(a.ring_len, a.claim_mask + 1);
}

// Second chance for subjects the seed launches recorded as past a cap (init hits beyond the
// per-wave stage, pool or work list full).  One workgroup moves the overflow list to the retry
// list (as long as the overflow list) and restarts the overflow list at 0; records that did not
// fit the overflow list are added to ovf_lost, and the host then runs the batch again with a
// longer list.
__global__ void __launch_bounds__(256) k_seed_retry_prep(DevCounters *ctr, const int64_t *ovf, uint32_t ovf_cap,
                                                         int64_t *retry) {
    const uint32_t total = ctr->overflow_count;
    const uint32_t n = total < ovf_cap ? total : ovf_cap;
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) retry[i] = ovf[i];
    __syncthreads();
    if (threadIdx.x == 0) {
        ctr->retry_count = n;
        ctr->next_retry = 0;
        ctr->ovf_lost += total - n;
        ctr->overflow_count = 0;
    }
}

// Rescans each retried (query, subject) pair with the query's own tables in global memory, a
// 32-bit diagonal ring in global memory and a stage of a.stage_cap records per wave.  The pair's
// word hits and extensions were counted by the first pass and are not counted again.
__global__ void __launch_bounds__(256) k_seed_retry(SeedArgs a) {
    extern __shared__ __align__(16) uint8_t smem[];
    const int lane = lane_id();
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // uniform: per-wave LDS bases live in SGPRs
    const uint32_t C = a.claim_mask + 1;
    int16_t(*mat)[kMatStride] = reinterpret_cast<int16_t(*)[kMatStride]>(smem);
    uint8_t *wb = smem + al16(kMatBytes) + (size_t)wave * wave_lds_bytes(8, C);
    WaveState<uint32_t> w;
    const uint64_t gwave = (uint64_t)blockIdx.x * (blockDim.x >> 6) + wave;
    w.ring = a.long_ring + gwave * a.ring_len;
    w.claim = reinterpret_cast<uint32_t *>(wb + 16);
    w.excl = w.claim + C;
    w.p0 = w.excl + 68;
    w.cnt = w.p0 + 64;
    w.hits = w.cnt + 4;
    w.task = reinterpret_cast<Task *>(w.hits + kHitQueue + kHitSlack);
    w.defer = reinterpret_cast<uint32_t *>(w.task + kTaskCap);
    w.subj = reinterpret_cast<uint8_t *>(w.defer + kDeferCap);
    w.grp = nullptr;
    for (int i = threadIdx.x; i < 28 * kMatStride; i += blockDim.x)
        mat[i / kMatStride][i % kMatStride] = i % kMatStride < 28 ? kBlosum62Dev[i / kMatStride][i % kMatStride] : 0;
    for (uint32_t i = lane; i < C; i += 64) w.claim[i] = 0xFFFFFFFFu;
    if (lane < 4) w.cnt[lane] = 0;
    __syncthreads();
    Ctx c;
    c.mat = mat;
    c.Qx = nullptr;  // subjects from global memory only: the walks read c.Q
    c.stage = a.stage + gwave * (uint64_t)a.stage_cap;
    c.stage_cap = a.stage_cap;
    c.claim_mask = a.claim_mask;
    c.gap = 0;
    uint32_t n_hits = 0, n_ext = 0;  // not counted again
    for (;;) {
        uint32_t k = 0;
        if (lane == 0) k = atomicAdd(&a.ctr->next_retry, 1u);
        k = first_lane(k);
        if (k >= a.ctr->retry_count) break;
        const int64_t e = a.retry[k];
        const int qi = (int)((e >> kOvfShift) & kOvfQiMask);
        const int64_t sid = e & kOvfSidMask;
        // a pair the gapped stage gave up on had its init hits counted when they were first found
        const bool count_init = (e & (int64_t)kOvfGapped) == 0;
        const QueryDev &qd = a.qd[qi];
        c.Q = qd.query;
        c.qlen = qd.qlen;
        c.ring_len = qd.ring_len;
        c.ring_base = qd.ring_base;
        c.ring_magic = qd.ring_magic;
        c.window = qd.window;
        c.xdrop = qd.x_ungapped;
        c.cutoff = qd.ungapped_cutoff;
        Tables t{};
        t.gprefix = qd.prefix;
        t.ent = qd.entries;
        t.rec = qd.rec;
        const uint64_t o0 = a.offs[sid];
        const int slen = (int)(a.offs[sid + 1] - o0);
        for (uint32_t i = lane; i < qd.ring_len; i += 64) w.ring[i] = 0;
        ring_sync<uint32_t>();
        scan_subject<0, false>(c, w, t, a.res + o0, slen, n_hits, n_ext);
        wave_sync();
        const uint32_t nih = w.cnt[2];
        wave_sync();
        if (nih > 0) {
            emit_init_hits(a, qd, w, c, sid, slen, nih, count_init);
            wave_sync();
            if (lane == 0) w.cnt[2] = 0;
            wave_sync();
        }
    }
}

size_t seed_retry_lds_bytes(const SeedArgs &a, int waves) {
    return al16(kMatBytes) + (size_t)waves * wave_lds_bytes(8, a.claim_mask + 1);
}

hipError_t launch_seed_retry(const SeedArgs &a, int64_t *retry, int blocks, int waves, hipStream_t st) {
    hipLaunchKernelGGL(k_seed_retry_prep, dim3(1), dim3(256), 0, st, a.ctr, a.ovf, a.ovf_cap, retry);
    SeedArgs r = a;
    r.retry = retry;
    hipLaunchKernelGGL(k_seed_retry, dim3(blocks), dim3(64 * waves), seed_retry_lds_bytes(a, waves), st, r);
    return hipGetLastError();
}

hipError_t launch_seed_pairs(const SeedArgs &a, int blocks, int waves, hipStream_t st) {
    hipLaunchKernelGGL(k_seed_retry, dim3(blocks), dim3(64 * waves), seed_retry_lds_bytes(a, waves), st, a);
    return hipGetLastError();
}

namespace {
template <int kTable, bool kLong>
int occupancy_of(int waves, size_t lds) {
    // the runtime query costs microseconds and a run asks it dozens of times: keep the answers
    static std::mutex mu;
    static std::map<std::pair<int, size_t>, int> known;
    std::lock_guard<std::mutex> lk(mu);
    const auto key = std::make_pair(waves, lds);
    auto it = known.find(key);
    if (it != known.end()) return it->second;
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_seed2<kTable, kLong>, 64 * waves, lds) != hipSuccess) n = 0;
    known[key] = n;
    return n;
}
}  // namespace

int seed2_blocks_per_cu(const SeedArgs &a, int table, int waves, bool long_subjects) {
    const size_t lds = seed2_lds_bytes(a, table, waves);
    if (long_subjects) return occupancy_of<0, true>(waves, lds);
    if (table == 2) return occupancy_of<2, false>(waves, lds);
    if (table == 1) return occupancy_of<1, false>(waves, lds);
    return occupancy_of<0, false>(waves, lds);
}

hipError_t launch_seed2(const SeedArgs &a, int table, int blocks, int waves, bool long_subjects, hipStream_t st) {
    const size_t lds = seed2_lds_bytes(a, table, waves);
    if (long_subjects)
        hipLaunchKernelGGL((k_seed2<0, true>), dim3(blocks), dim3(64 * waves), lds, st, a);
    else if (table == 2)
        hipLaunchKernelGGL((k_seed2<2, false>), dim3(blocks), dim3(64 * waves), lds, st, a);
    else if (table == 1)
        hipLaunchKernelGGL((k_seed2<1, false>), dim3(blocks), dim3(64 * waves), lds, st, a);
    else
        hipLaunchKernelGGL((k_seed2<0, false>), dim3(blocks), dim3(64 * waves), lds, st, a);
    return hipGetLastError();
}

}  // namespace blz
