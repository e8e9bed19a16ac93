// blz_gapped.hip -- gapped stage: one lane per subject of the work list.
//
// Per subject (NCBI BLAST_GetGappedScore, called at BLAZE_ncbi/blast_engine.c:767):
//   init hits in score order; an init hit inside an HSP already saved with at least its score
//   is skipped; otherwise the gapped start is found on its diagonal, the score-only semi-gapped
//   X-drop DP runs left (reversed prefixes, start included) and right (suffixes), and the HSP is
//   saved when left + right reaches the gapped cutoff.  Then the common-endpoint purge
//   (blast_engine.c:792) and the score sort (blast_engine.c:801).
// Reference counterpart: gpuBLAZE_everything + semiGappedKernel*, BLAZE_gpu/gpuBLAZE.cu:94-677,
// 1254-1397 (which runs the full rectangle, not the X-drop band, and only flags survivors).
//
// Two DP storages, same arithmetic:
//   LDS ring (k_gapped_lds): kLdsRing cells per lane, each one 32-bit word holding the cell's
//     best and best-gap scores as 16-bit offsets from the best score at the start of the row
//     that wrote it.  Every live cell holds a value within [best - X - gap_open - gap_extend,
//     best + 11], so 16 bits are exact; anything further below the running best can never
//     reach the X-drop window again and is stored as minus infinity.  Lanes of a wave take
//     consecutive entries of a cost-ordered work list, so they run DPs of similar size.
//   Global ring (k_gapped_wide): 32-bit pairs in a per-lane ring of dp_wmask+1 cells, for the
//     few subjects whose band does not fit the LDS ring.
#include <hip/hip_runtime.h>

#include <type_traits>

#include "blz_common.h"
#include "blz_dev_common.h"
#include "blz_device.h"

namespace blz {

namespace {

// --- DP storage policies ------------------------------------------------------------------
// get(b) -> (best, best_gap) of column b as written by the previous row;
// put(b, best, best_gap) for the current row.  start_row() is called before each row with
// the best score at the start of that row.

struct GlobalRing {
    int2 *ring;
    int wmask;
    __device__ __forceinline__ void start_row(int) {}
    __device__ __forceinline__ int2 get(int b) const { return ring[b & wmask]; }
    __device__ __forceinline__ void put(int b, int h, int e) { ring[b & wmask] = make_int2(h, e); }
    __device__ __forceinline__ void put_h(int b, int h) { ring[b & wmask].x = h; }
    __device__ __forceinline__ int capacity() const { return wmask + 1; }
};

struct LdsRing16 {
    uint32_t *ring;  // [kLdsRing][64], this lane's column
    int base_prev = 0, base_cur = 0;
    static constexpr int kNeg16 = -32768;
    __device__ __forceinline__ void start_row(int best) {
        base_prev = base_cur;
        base_cur = best;
    }
    __device__ __forceinline__ static int dec(int v16, int base) { return v16 == kNeg16 ? kMinIntDp : base + v16; }
    __device__ __forceinline__ static uint32_t enc1(int v, int base) {
        const int d = v - base;
        return (uint32_t)(uint16_t)(int16_t)(d < -32767 ? kNeg16 : d);
    }
    __device__ __forceinline__ int2 get(int b) const {
        const uint32_t w = ring[(b & (kLdsRing - 1)) * 64];
        return make_int2(dec((int16_t)(w & 0xFFFFu), base_prev), dec((int16_t)(w >> 16), base_prev));
    }
    __device__ __forceinline__ void put(int b, int h, int e) {
        ring[(b & (kLdsRing - 1)) * 64] = enc1(h, base_cur) | (enc1(e, base_cur) << 16);
    }
    // a cell that failed the X-drop: both scores become minus infinity (its gap score can no
    // longer reach the X-drop window either)
    __device__ __forceinline__ void put_h(int b, int) { ring[(b & (kLdsRing - 1)) * 64] = 0x80008000u; }
    __device__ __forceinline__ int capacity() const { return kLdsRing; }
};

// Score-only semi-gapped X-drop DP (NCBI Blast_SemiGappedAlign, score-only path).  Rows walk
// A, the live band [first_b_index, b_size) walks B.  Sets *ovf when the band outgrows R.
template <class R>
__device__ int semi_gapped(const int16_t (*mat)[28], const uint8_t *A, const uint8_t *B, int M, int N,
                           int x_dropoff, int gap_open, int gap_extend, bool reverse, int *a_off, int *b_off,
                           R &ring, bool *ovf, unsigned long long *cells) {
    const int goe = gap_open + gap_extend;
    *a_off = 0;
    *b_off = 0;
    if (x_dropoff < goe) x_dropoff = goe;
    if (N <= 0 || M <= 0) return 0;
    const int W = ring.capacity();
    ring.start_row(0);
    ring.start_row(0);
    int score = -goe;
    ring.put(0, 0, -goe);
    int i;
    for (i = 1; i <= N; i++) {
        if (score < -x_dropoff) break;
        if (i >= W) {
            *ovf = true;
            return 0;
        }
        ring.put(i, score, score - goe);
        score -= gap_extend;
    }
    int b_size = i;
    int best_score = 0;
    int first_b_index = 0;
    unsigned long long nc = 0;
    for (int a_index = 1; a_index <= M; a_index++) {
        ring.start_row(best_score);
        const int16_t *mrow = mat[reverse ? A[M - a_index] : A[a_index]];
        int sc = kMinIntDp;
        int score_gap_row = kMinIntDp;
        int last_b_index = first_b_index;
        for (int b_index = first_b_index; b_index < b_size; b_index++) {
            const int2 cell = ring.get(b_index);
            int score_gap_col = cell.y;
            // diagonal score of the next column; the last column (b_index == N) has none
            const int next_score =
                b_index < N ? cell.x + mrow[reverse ? B[N - 1 - b_index] : B[b_index + 1]] : kMinIntDp;
            nc++;
            if (sc < score_gap_col) sc = score_gap_col;
            if (sc < score_gap_row) sc = score_gap_row;
            if (best_score - sc > x_dropoff) {
                if (b_index == first_b_index)
                    first_b_index++;
                else
                    ring.put_h(b_index, kMinIntDp);
            } else {
                last_b_index = b_index;
                if (sc > best_score) {
                    best_score = sc;
                    *a_off = a_index;
                    *b_off = b_index;
                }
                score_gap_row -= gap_extend;
                score_gap_col -= gap_extend;
                const int open = sc - goe;
                const int e = open > score_gap_col ? open : score_gap_col;
                score_gap_row = open > score_gap_row ? open : score_gap_row;
                ring.put(b_index, sc, e);
            }
            sc = next_score;
        }
        if (first_b_index == b_size) break;
        if (last_b_index < b_size - 1) {
            b_size = last_b_index + 1;
        } else {
            while (score_gap_row >= best_score - x_dropoff && b_size <= N) {
                if (b_size - first_b_index >= W) {
                    *ovf = true;
                    return 0;
                }
                ring.put(b_size, score_gap_row, score_gap_row - goe);
                score_gap_row -= gap_extend;
                b_size++;
            }
        }
        if (b_size <= N) {
            if (b_size - first_b_index >= W) {
                *ovf = true;
                return 0;
            }
            ring.put(b_size, kMinIntDp, kMinIntDp);
            b_size++;
        }
    }
    *cells += nc;
    return best_score;
}

// Gapped start: middle of a short segment, else the centre of its best 11-letter window when
// that window scores above 0, else the segment start (reference restatement at
// gpuBLAZE.cu:1166-1205).
template <int kStride>
__device__ __forceinline__ int gapped_start(const int16_t (*mat)[kStride], const uint8_t *q, const uint8_t *s,
                                            int q_start, int s_start, int len) {
    if (len <= kHspMaxWindow) return q_start + len / 2;
    // both sequences are in global memory: global loads, not flat ones
    const __attribute__((address_space(1))) uint8_t *qa = (const __attribute__((address_space(1))) uint8_t *)(q + q_start);
    const __attribute__((address_space(1))) uint8_t *sa = (const __attribute__((address_space(1))) uint8_t *)(s + s_start);
    int score = 0;
    for (int i = 0; i < kHspMaxWindow; i++) score += mat[qa[i]][sa[i]];
    int max_score = score;
    int max_k = kHspMaxWindow - 1;  // end of the best window, relative to the segment start
    int k = kHspMaxWindow;
    for (; k + 4 <= len; k += 4) {  // four window moves per pass: their sixteen letters load together
        uint32_t qo[4], so[4], qn[4], sn[4];
#pragma unroll
        for (int j = 0; j < 4; j++) {
            qo[j] = qa[k + j - kHspMaxWindow];
            so[j] = sa[k + j - kHspMaxWindow];
            qn[j] = qa[k + j];
            sn[j] = sa[k + j];
        }
#pragma unroll
        for (int j = 0; j < 4; j++) {
            score += mat[qn[j]][sn[j]] - mat[qo[j]][so[j]];
            if (score > max_score) {
                max_score = score;
                max_k = k + j;
            }
        }
    }
    for (; k < len; k++) {
        score += mat[qa[k]][sa[k]] - mat[qa[k - kHspMaxWindow]][sa[k - kHspMaxWindow]];
        if (score > max_score) {
            max_score = score;
            max_k = k;
        }
    }
    return max_score > 0 ? q_start + max_k - kHspMaxWindow / 2 : q_start;
}

enum ItemOutcome { kDone = 0, kRedo = 1, kOverflow = 2 };  // kRedo: band or HSP count past this pass's caps

// Wave-parallel semi-gapped X-drop DP for one extension (all 64 lanes call it with the same
// arguments).  Rows are swept in order; the cells of a row's band are spread over the lanes, 64
// per step.  Within a row: D(b) = H_prev(b-1) + s(A_a, B_b) (none for the band's first cell),
// E(b) = the column gap of the previous row, the row gap F(b) = max over b' < b of
// max(D, E)(b') - open - extend * (b - b') is a prefix max, H = max(D, E, F), and a cell survives
// when (best before it in row order) - H <= X, the running best being a prefix max too.  The band
// bounds and the gap run appended past a surviving last cell follow the same rules as the
// lane-serial loop; the row-vector form is pinned against NCBI's loop in tests/test_oracle.py.
// Cells that fail the X-drop store minus infinity for both scores (see LdsRing16).
template <int kSeg>
__device__ int wave_semi_gapped(const int16_t (*mat)[28], const uint8_t *A, const uint8_t *B, int M, int N,
                                int x_dropoff, int gap_open, int gap_extend, bool reverse, int *a_off, int *b_off,
                                int2 *ring, int wmask, bool *ovf, unsigned long long *cells) {
    using S = Seg<kSeg>;
    const int lane = S::lane();
    const int goe = gap_open + gap_extend;
    const int ge = gap_extend;
    *a_off = 0;
    *b_off = 0;
    if (x_dropoff < goe) x_dropoff = goe;
    const int X = x_dropoff;
    if (N <= 0 || M <= 0) return 0;
    const int W = wmask + 1;
    // row 0: a run of gaps in A while it stays within the X-drop
    const int k0 = (X - goe) / ge + 1;
    int size = (N < k0 ? N : k0) + 1;
    if (size > W) {
        *ovf = true;
        return 0;
    }
    for (int b = lane; b < size; b += kSeg) {
        const int h = b == 0 ? 0 : -goe - (b - 1) * ge;
        ring[b & wmask] = make_int2(h, h - goe);
    }
    wave_sync();
    int best = 0, first = 0, ba = 0, bb = 0;
    unsigned long long nc = 0;
    for (int a = 1; a <= M; a++) {
        const int16_t *mrow = mat[reverse ? A[M - a] : A[a]];
        int carry_g = kMinIntDp;   // max of max(D,E)(b') + b' * extend over earlier steps
        int carry_h = best;        // running best over earlier steps
        int carry_hp = kMinIntDp;  // previous row's H left of this step (none for the first cell)
        int fa = -1, la = -1, rmax = kMinIntDp, rmax_b = -1;
        for (int c0 = first; c0 < size; c0 += kSeg) {
            const int b = c0 + lane;
            const bool valid = b < size;
            const int2 prev = valid ? ring[b & wmask] : make_int2(kMinIntDp, kMinIntDp);
            const int hp_left = S::shift_up1(prev.x, carry_hp);
            const int sb = (valid && b >= 1 && b <= N) ? mrow[reverse ? B[N - b] : B[b]] : 0;
            const int de = valid ? max(hp_left + sb, prev.y) : kMinIntDp;
            const int g = valid ? de + b * ge : kMinIntDp;
            const int gin = S::incl_max(g);
            const int gex = max(S::shift_up1(gin, kMinIntDp), carry_g);
            const int f = gex - goe - (b - 1) * ge;
            const int h = valid ? max(de, f) : kMinIntDp;
            const int hin = S::incl_max(h);
            const int hex = max(S::shift_up1(hin, kMinIntDp), carry_h);
            const bool alive = valid && hex - h <= X;
            const uint64_t am = S::ballot(alive);
            if (am) {
                if (fa < 0) fa = c0 + (int)__builtin_ctzll(am);
                la = c0 + 63 - (int)__builtin_clzll(am);
            }
            // The row's best live score: a dead cell lies below the running best, so only a chunk
            // whose maximum tops the running best (carry_h) can raise it, and the first cell
            // reaching that maximum is live (everything before it in row order is lower).
            const int cm = S::last(hin);
            if (cm > carry_h) {
                rmax = cm;
                rmax_b = c0 + (int)__builtin_ctzll(S::ballot(valid && h == cm));
            }
            if (valid)
                ring[b & wmask] = alive ? make_int2(h, max(h - goe, prev.y - ge)) : make_int2(kMinIntDp, kMinIntDp);
            carry_g = max(carry_g, S::last(gin));
            carry_h = max(carry_h, cm);
            carry_hp = S::last(prev.x);
        }
        nc += (unsigned long long)(size - first);
        if (fa < 0) break;
        if (rmax > best) {
            best = rmax;
            ba = a;
            bb = rmax_b;
        }
        first = fa;
        if (la < size - 1) {
            size = la + 1;
        } else {
            const int gr = carry_g - goe - (size - 1) * ge;  // row gap one past the last cell
            if (gr >= best - X) {
                int k = (ge == 1 ? gr - (best - X) : (gr - (best - X)) / ge) + 1;
                if (k > N + 1 - size) k = N + 1 - size;
                if (size + k - first > W) {
                    *ovf = true;
                    return 0;
                }
                for (int i = lane; i < k; i += kSeg) {
                    const int hv = gr - i * ge;
                    ring[(size + i) & wmask] = make_int2(hv, hv - goe);
                }
                size += k;
            }
        }
        if (size <= N) {
            if (size + 1 - first > W) {
                *ovf = true;
                return 0;
            }
            if (lane == 0) ring[size & wmask] = make_int2(kMinIntDp, kMinIntDp);
            size++;
        }
        wave_sync();
    }
    *a_off = ba;
    *b_off = bb;
    *cells += nc;
    return best;
}

__device__ int finish_hsps(const GappedArgs &a, const WorkItem &wi, Hsp *saved, int ns);

// Whole gapped stage of one subject, by one lane (kWave false) or by a whole wave (kWave true:
// every lane runs the same control flow, the DP is shared across the lanes and lane 0 does the
// sorting and the writes).  Returns kRedo if the DP band or the HSP count outgrew this pass's
// storage (nothing was written), kOverflow if an output pool is full.
template <bool kWave, class R, int kSeg = 64>
__device__ int gapped_item(const GappedArgs &a, const int16_t (*mat)[28], uint32_t widx, R &ring, Hsp *saved,
                           unsigned long long &cells, unsigned long long &gext) {
    const bool writer = !kWave || Seg<kSeg>::lane() == 0;
    const WorkItem wi = a.work[widx];
    InitHit *ih = a.pool + wi.off;
    const int n = wi.n;
    if (writer) {
        for (int i = 1; i < n; i++) {  // score order of the init hits (sorted in place once)
            InitHit x = ih[i];
            int j = i - 1;
            while (j >= 0 && init_before(x, ih[j])) {
                ih[j + 1] = ih[j];
                j--;
            }
            ih[j + 1] = x;
        }
    }
    if (kWave) wave_sync();
    const uint64_t o0 = a.offs[wi.sid];
    const int slen = (int)(a.offs[wi.sid + 1] - o0);
    const uint8_t *S = a.res + o0;
    const QueryDev &qd = a.qd[wi.qi];
    const uint8_t *Q = qd.query;
    const int qlen = qd.qlen;
    int ns = 0;
    bool ovf = false;
    unsigned long long my_cells = 0, my_ext = 0;
    for (int k = 0; k < n && !ovf; k++) {
        const InitHit h = ih[k];
        const int q0 = h.q_start, q1 = h.q_start + h.len - 1;
        const int s0 = h.s_start, s1 = h.s_start + h.len - 1;
        bool skip = false;
        for (int t = 0; t < ns && !skip; t++) {
            const Hsp &u = saved[t];
            skip = h.score <= u.score && u.q_off <= q0 && q0 <= u.q_end && u.s_off <= s0 && s0 <= u.s_end &&
                   u.q_off <= q1 && q1 <= u.q_end && u.s_off <= s1 && s1 <= u.s_end;
        }
        if (skip) continue;
        const int q_gs = gapped_start(mat, Q, S, h.q_start, h.s_start, h.len);
        const int s_gs = h.s_start + (q_gs - h.q_start);
        int la, lb, ra, rb, left, right;
        if constexpr (kWave) {
            left = wave_semi_gapped<kSeg>(mat, Q, S, q_gs + 1, s_gs + 1, qd.x_gapped, qd.gap_open, qd.gap_extend, true, &la,
                                    &lb, ring.ring, ring.wmask, &ovf, &my_cells);
            if (ovf) break;
            right = wave_semi_gapped<kSeg>(mat, Q + q_gs, S + s_gs, qlen - q_gs - 1, slen - s_gs - 1, qd.x_gapped,
                                     qd.gap_open, qd.gap_extend, false, &ra, &rb, ring.ring, ring.wmask, &ovf,
                                     &my_cells);
        } else {
            left = semi_gapped(mat, Q, S, q_gs + 1, s_gs + 1, qd.x_gapped, qd.gap_open, qd.gap_extend, true, &la,
                               &lb, ring, &ovf, &my_cells);
            if (ovf) break;
            right = semi_gapped(mat, Q + q_gs, S + s_gs, qlen - q_gs - 1, slen - s_gs - 1, qd.x_gapped, qd.gap_open,
                                qd.gap_extend, false, &ra, &rb, ring, &ovf, &my_cells);
        }
        if (ovf) break;
        my_ext++;
        const int score = left + right;
        if (score >= qd.gapped_cutoff) {
            if (ns >= a.saved_cap) return kRedo;  // more HSPs than this pass can hold
            if (writer) saved[ns] = Hsp{score, q_gs + 1 - la, q_gs + 1 + ra, s_gs + 1 - lb, s_gs + 1 + rb, q_gs, s_gs};
            ns++;
            if (kWave) wave_sync();
        }
    }
    if (ovf) return kRedo;  // counters are only taken from the pass that completes the subject
    if (writer) {
        cells += my_cells;
        gext += my_ext;
    }
    int outcome = writer ? finish_hsps(a, wi, saved, ns) : kDone;
    if (kWave) outcome = Seg<kSeg>::first(outcome);
    return outcome;
}

// The subject's saved HSPs: common-endpoint purge (NCBI blast_engine.c:792), score sort
// (blast_engine.c:801), then the list and its result item out.  One lane.
__device__ int finish_hsps(const GappedArgs &a, const WorkItem &wi, Hsp *saved, int ns) {
    if (ns == 0) return kDone;
    int outcome = kDone;
    {
        insertion_sort<cmp_start>(saved, ns);
        int m = 0;
        for (int i = 0; i < ns; i++) {
            if (m > 0 && saved[m - 1].q_off == saved[i].q_off && saved[m - 1].s_off == saved[i].s_off) continue;
            saved[m++] = saved[i];
        }
        ns = m;
        insertion_sort<cmp_end>(saved, ns);
        m = 0;
        for (int i = 0; i < ns; i++) {
            if (m > 0 && saved[m - 1].q_end == saved[i].q_end && saved[m - 1].s_end == saved[i].s_end) continue;
            saved[m++] = saved[i];
        }
        ns = m;
        insertion_sort<cmp_score>(saved, ns);
        // prelim e-value reap (blast_engine.c:1457): for one query the e-value falls as the
        // score rises, so it keeps the HSPs scoring at least reap_score, a prefix of the order
        const int reap = a.qd[wi.qi].reap_score;
        while (ns > 0 && saved[ns - 1].score < reap) ns--;
        if (ns == 0) return kDone;
        const unsigned long long base = atomicAdd(&a.ctr->hsp_used, (unsigned long long)ns);
        if (base + ns > a.hsp_cap) {
            outcome = kOverflow;
        } else {
            for (int i = 0; i < ns; i++) a.hsps[base + i] = saved[i];
            const uint32_t ri = atomicAdd(&a.ctr->result_count, 1u);
            a.results[ri] = ResultItem{wi.sid, (uint64_t)base, ns, wi.qi};
        }
    }
    return outcome;
}

__device__ __forceinline__ void record_overflow(const GappedArgs &a, uint32_t widx) {
    const WorkItem wi = a.work[widx];
    const uint32_t k = atomicAdd(&a.ctr->overflow_count, 1u);
    if (k < a.ovf_cap) a.ovf[k] = (int64_t)kOvfGapped | ((int64_t)wi.qi << kOvfShift) | wi.sid;
}

}  // namespace

// Work list in cost order: histogram counts per class come from the seed stage.  A workgroup
// orders a contiguous range of items: it counts them per class in LDS, reserves each class's
// share with one global atomic, then scatters.  (With one global atomic per class and wave the
// eight counters' contention cost 4.6 ms at 64 queries x 500,000 subjects.)
__global__ void __launch_bounds__(256) k_order(const WorkItem *work, DevCounters *ctr, uint32_t *order,
                                               uint32_t work_cap) {
    __shared__ uint32_t cnt[kBuckets], start[kBuckets];
    const uint32_t n = ctr->work_count < work_cap ? ctr->work_count : work_cap;
    uint32_t base[kBuckets];
    uint32_t acc = 0;
    for (int b = 0; b < kBuckets; b++) {
        base[b] = acc;
        acc += ctr->bucket_count[b];
    }
    const int lane = lane_id();
    const uint64_t below = (1ull << lane) - 1;
    const uint32_t per = (n + gridDim.x - 1) / gridDim.x;
    const uint32_t r0 = min(n, blockIdx.x * per), r1 = min(n, r0 + per);
    if (threadIdx.x < kBuckets) cnt[threadIdx.x] = 0;
    __syncthreads();
    // one LDS counter update per class present in a wave, the first waiting lane's class first;
    // the second pass places each item at its class's start plus its rank
    auto pass = [&](bool place) {
        for (uint32_t i0 = r0; i0 < r1; i0 += blockDim.x) {
            const uint32_t i = i0 + threadIdx.x;
            const bool has = i < r1;
            const int b = has ? work[i].bucket : -1;
            uint64_t left = __ballot(has);
            while (left) {
                const int k = __shfl(b, __ffsll((long long)left) - 1, 64);
                const uint64_t m = __ballot(b == k);
                const int leader = __ffsll((long long)m) - 1;
                uint32_t at = 0;
                if (lane == leader) at = atomicAdd(&cnt[k], (uint32_t)__popcll(m));
                at = (uint32_t)__shfl((int)at, leader, 64);
                if (place && b == k) order[base[k] + start[k] + at + (uint32_t)__popcll(m & below)] = i;
                left &= ~m;
            }
        }
    };
    pass(false);
    __syncthreads();
    if (threadIdx.x < kBuckets) {
        const uint32_t c = cnt[threadIdx.x];
        start[threadIdx.x] = c ? atomicAdd(&ctr->bucket_cursor[threadIdx.x], c) : 0u;
        cnt[threadIdx.x] = 0;
    }
    __syncthreads();
    pass(true);
}

__global__ void __launch_bounds__(64) k_gapped_lds(GappedArgs a) {
    __shared__ int16_t mat[28][28];
    __shared__ uint32_t ring_mem[kLdsRing * 64];
    for (int i = threadIdx.x; i < 28 * 28; i += blockDim.x) mat[i / 28][i % 28] = kBlosum62Dev[i / 28][i % 28];
    __syncthreads();
    const int lane = lane_id();
    const uint64_t gtid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    Hsp *saved = a.hsp_scratch + gtid * (uint64_t)a.saved_cap;
    LdsRing16 ring;
    ring.ring = ring_mem + lane;
    const uint32_t n = *a.count < a.count_cap ? *a.count : a.count_cap;
    // the heavy cost classes (first in the order) are done by k_gapped_wave
    uint32_t light0 = 0;
    for (int b = 0; b < a.heavy_classes; b++) light0 += a.ctr->bucket_count[b];
    unsigned long long cells = 0, gext = 0;
    for (;;) {
        uint32_t base = 0;
        if (lane == 0) base = atomicAdd(a.next, 64u);
        base = first_lane(base) + light0;
        if (base >= n) break;
        const uint32_t i = base + lane;
        if (i < n) {
            const uint32_t widx = a.order[i];
            const int r = gapped_item<false>(a, mat, widx, ring, saved, cells, gext);
            if (r == kRedo) {
                const uint32_t k = atomicAdd(&a.ctr->redo_count, 1u);
                if (k < a.redo_cap) a.redo[k] = widx;
                else record_overflow(a, widx);
            } else if (r == kOverflow) {
                record_overflow(a, widx);
            }
        }
    }
    unsigned long long c = cells, g = gext;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        c += __shfl_xor(c, o, 64);
        g += __shfl_xor(g, o, 64);
    }
    if (lane == 0) {
        atomicAdd(&a.ctr->dp_cells, c);
        atomicAdd(&a.ctr->gapped_ext, g);
    }
}

__global__ void __launch_bounds__(1024) k_gapped_wide(GappedArgs a) {
    __shared__ int16_t mat[28][28];
    for (int i = threadIdx.x; i < 28 * 28; i += blockDim.x) mat[i / 28][i % 28] = kBlosum62Dev[i / 28][i % 28];
    __syncthreads();
    const uint64_t gtid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t nthreads = (uint64_t)gridDim.x * blockDim.x;
    GlobalRing ring{a.dp_scratch + gtid * (uint64_t)(a.dp_wmask + 1), a.dp_wmask};
    Hsp *saved = a.hsp_scratch + gtid * (uint64_t)a.saved_cap;
    const uint32_t n = *a.count < a.count_cap ? *a.count : a.count_cap;
    unsigned long long cells = 0, gext = 0;
    for (uint64_t i = gtid; i < n; i += nthreads) {
        const uint32_t widx = a.order[i];
        if (gapped_item<false>(a, mat, widx, ring, saved, cells, gext) != kDone) record_overflow(a, widx);
    }
    atomicAdd(&a.ctr->dp_cells, cells);
    atomicAdd(&a.ctr->gapped_ext, gext);
}

// Subjects of the heavy cost classes: one wave per subject, the DP band across the lanes.
// Takes order[0, heavy) where heavy = the item count of the classes below a.heavy_classes.
__global__ void __launch_bounds__(64) k_gapped_wave(GappedArgs a) {
    __shared__ int16_t mat[28][28];
    __shared__ int2 ring_mem[kWaveRing];
    for (int i = threadIdx.x; i < 28 * 28; i += blockDim.x) mat[i / 28][i % 28] = kBlosum62Dev[i / 28][i % 28];
    __syncthreads();
    const int lane = lane_id();
    uint32_t heavy = 0;
    for (int b = 0; b < a.heavy_classes; b++) heavy += a.ctr->bucket_count[b];
    const uint32_t n = *a.count < a.count_cap ? *a.count : a.count_cap;
    if (heavy > n) heavy = n;
    struct {
        int2 *ring;
        int wmask;
    } ring{ring_mem, kWaveRing - 1};
    Hsp *saved = a.hsp_scratch + (uint64_t)blockIdx.x * a.saved_cap;
    unsigned long long cells = 0, gext = 0;
    for (;;) {
        uint32_t i = 0;
        if (lane == 0) i = atomicAdd(a.next_heavy, 1u);
        i = first_lane(i);
        if (i >= heavy) break;
        const uint32_t widx = a.order[i];
        const int r = gapped_item<true>(a, mat, widx, ring, saved, cells, gext);
        if (lane == 0) {
            if (r == kRedo) {
                const uint32_t k = atomicAdd(&a.ctr->redo_count, 1u);
                if (k < a.redo_cap) a.redo[k] = widx;
                else record_overflow(a, widx);
            } else if (r == kOverflow) {
                record_overflow(a, widx);
            }
        }
    }
    if (lane == 0) {
        atomicAdd(&a.ctr->dp_cells, cells);
        atomicAdd(&a.ctr->gapped_ext, gext);
    }
}

// Gapped stage with four subjects per wave, one per 16-lane DPP row.  Each segment runs its own
// sequence of subjects, init hits and DPs as a state machine; every pass of the wave's loop
// advances each segment by one step of what it is doing (take a subject, pick the next init hit
// and its gapped start, one DP row), so a segment whose DP ends starts the next at once instead
// of waiting at a loop exit for the other segments' rows.  Same per-subject steps, arithmetic,
// caps and outcomes as gapped_item / wave_semi_gapped.
//
// kPk: packed 16-bit DP.  Each lane holds two adjacent columns (a 16-lane chunk spans 32), and
// the row's arithmetic runs on pairs of int16 with v_pk_* instructions, adds and subtracts
// saturating.  Scores are kept relative to the DP's start (0): a cell's score is at most 11 per
// row, so for min(rows, columns) <= kPkMaxRows every live value lies between -(X + open + extend)
// and 31,900, and minus infinity (-32,768) stays minus infinity under saturation; the X-drop test
// never sees a clipped live value.  The row-gap prefix max runs on (column - row start) x extend,
// which the band keeps below 128 x extend.  A DP past those bounds sends its subject to the wide
// pass, as a band past the ring does.  The ring holds one 32-bit word per column (H low, E high).
constexpr int kPkMaxRows = 2900;
template <bool kPk>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(5))) k_gapped_seg(GappedArgs a) {
    using S = Seg<16>;
    using RingCell = typename std::conditional<kPk, uint32_t, int2>::type;
    constexpr int kLetRing = 256;   // letters of a DP's columns per segment (the band and 64 ahead)
    constexpr int kZeroLetter = 28;  // scores 0 against every letter: the gapped start's own column
    static_assert(kSegRing + 64 <= kLetRing, "a band and the 64 letters filled past it fit the letter ring");
    __shared__ int16_t mat[28][32];
    __shared__ RingCell ring_mem[4][kSegRing];
    __shared__ uint8_t lett_mem[4][kLetRing];
    for (int i = threadIdx.x; i < 28 * 32; i += blockDim.x) mat[i / 32][i % 32] = i % 32 < 28 ? kBlosum62Dev[i / 32][i % 32] : 0;
    for (int i = threadIdx.x; i < 4 * kLetRing; i += blockDim.x) lett_mem[i / kLetRing][i % kLetRing] = kZeroLetter;
    __syncthreads();
    const int lane = lane_id();
    const int seg = lane >> 4;
    const int sl = S::lane();
    const bool writer = sl == 0;
    uint32_t heavy = 0;
    for (int b = 0; b < a.heavy_classes; b++) heavy += a.ctr->bucket_count[b];
    const uint32_t navail = *a.count < a.count_cap ? *a.count : a.count_cap;
    if (heavy > navail) heavy = navail;
    RingCell *ring = ring_mem[seg];
    uint8_t *lett = lett_mem[seg];
    constexpr int wmask = kSegRing - 1, W = kSegRing;
    Hsp *saved = a.hsp_scratch + ((uint64_t)blockIdx.x * 4 + seg) * a.saved_cap;
    unsigned long long cells = 0, gext = 0;

    enum { kTake, kHit, kDp, kDone };
    int phase = kTake;
    // the segment's subject
    uint32_t widx = 0;
    WorkItem wi{};
    const InitHit *ih = nullptr;
    const uint8_t *Sj = nullptr, *Q = nullptr;
    int slen = 0, qlen = 0, n = 0, k = 0, ns = 0;
    bool redo = false;
    int xg = 0, gopen = 0, gext1 = 1, gcut = 0;
    unsigned long long item_cells = 0, item_ext = 0;
    int q_gs = 0, s_gs = 0, left = 0, la_l = 0, lb_l = 0;
    // its current DP (wave_semi_gapped's state)
    int dir = 0, M = 0, N = 0, X = 0, goe = 0, ge = 1;
    const uint8_t *A = nullptr, *B = nullptr;
    bool rev = false;
    enum { kRun, kEmpty, kOvf };
    int dps = kRun;  // the DP's status: running, empty (no rows or columns), band past the ring
    int arow = 1, first = 0, size = 0, best = 0, ba = 0, bb = 0;
    int filled = 0;  // columns [0, filled) have their letters in the letter ring
    s16x2 lane_bge{0, 0};  // kPk: this lane's two columns' offsets in a chunk, times the extension
    // a ring cell holding (h, e)
    auto cell = [](int h, int e) {
        if constexpr (kPk) {
            const int hh = h < -32768 ? -32768 : h, ee = e < -32768 ? -32768 : e;
            return (RingCell)((uint32_t)(uint16_t)(int16_t)hh | ((uint32_t)(uint16_t)(int16_t)ee << 16));
        } else {
            return (RingCell)make_int2(h, e);
        }
    };
    auto neg_cell = []() {
        if constexpr (kPk) return (RingCell)kNeg16x2;
        else return (RingCell)make_int2(kMinIntDp, kMinIntDp);
    };
    unsigned long long nc = 0;

    // Letters of the DP's columns up to the band's end, 64 columns at a time; column 0 (the
    // gapped start, no letter of B) takes the zero-score letter.  The band spans at most W = 128
    // columns, so the 256 slots never hold two live columns.
    auto fill_letters = [&]() {
        while (filled < size) {
            for (int j = sl; j < 64; j += 16) {
                const int b = filled + j;
                const int lt = b >= 1 && b <= N ? (int)(rev ? B[N - b] : B[b]) : kZeroLetter;
                lett[b & (kLetRing - 1)] = (uint8_t)lt;
            }
            filled += 64;
        }
    };

    // sets up the DP of direction d for init hit k (rows walk the query, the band the subject)
    auto dp_begin = [&](int d) {
        dir = d;
        if (d == 0) {  // left: reversed prefixes, the gapped start included
            A = Q; B = Sj; M = q_gs + 1; N = s_gs + 1; rev = true;
        } else {       // right: the suffixes after the gapped start
            A = Q + q_gs; B = Sj + s_gs; M = qlen - q_gs - 1; N = slen - s_gs - 1; rev = false;
        }
        goe = gopen + gext1;
        ge = gext1;
        X = xg < goe ? goe : xg;
        best = 0; ba = 0; bb = 0; nc = 0; first = 0; arow = 1;
        dps = N <= 0 || M <= 0 ? kEmpty : kRun;
        if (dps == kEmpty) return;
        if constexpr (kPk) {  // outside the 16-bit bounds: the wide pass takes the subject
            if ((M < N ? M : N) > kPkMaxRows || X + goe > 8192 || 128 * ge > 8192) {
                dps = kOvf;
                return;
            }
            lane_bge = s16x2{(short)(2 * sl * ge), (short)((2 * sl + 1) * ge)};
        }
        const int k0 = (X - goe) / ge + 1;  // row 0: a run of gaps while it stays within the X-drop
        size = (N < k0 ? N : k0) + 1;
        if (size > W) {
            dps = kOvf;
            return;
        }
        for (int b = sl; b < size; b += 16) {
            const int h = b == 0 ? 0 : -goe - (b - 1) * ge;
            ring[b & wmask] = cell(h, h - goe);
        }
        filled = 0;
        fill_letters();
    };

    for (;;) {
        if (phase == kTake) {  // the next subject of the cost order
            uint32_t i = 0;
            if (writer) i = atomicAdd(a.next_heavy, 1u);
            i = (uint32_t)__shfl((int)i, lane & 48, 64);
            if (i >= heavy) {
                phase = kDone;
            } else {
                widx = a.order[i];
                wi = a.work[widx];
                InitHit *ihw = a.pool + wi.off;
                n = wi.n;
                if (writer) {
                    for (int x = 1; x < n; x++) {  // score order of the init hits (sorted in place once)
                        InitHit h = ihw[x];
                        int j = x - 1;
                        while (j >= 0 && init_before(h, ihw[j])) {
                            ihw[j + 1] = ihw[j];
                            j--;
                        }
                        ihw[j + 1] = h;
                    }
                }
                ih = ihw;
                const uint64_t o0 = a.offs[wi.sid];
                slen = (int)(a.offs[wi.sid + 1] - o0);
                Sj = a.res + o0;
                const QueryDev &qd = a.qd[wi.qi];
                Q = qd.query;
                qlen = qd.qlen;
                xg = qd.x_gapped;
                gopen = qd.gap_open;
                gext1 = qd.gap_extend;
                gcut = qd.gapped_cutoff;
                k = 0;
                ns = 0;
                redo = false;
                item_cells = 0;
                item_ext = 0;
                phase = kHit;
            }
        }
        wave_sync();  // the sorted init hits before anyone reads them
        if (phase == kHit) {  // the next init hit outside the HSPs saved so far, or the end of the subject
            bool found = false;
            while (k < n) {
                const InitHit h = ih[k];
                const int q0 = h.q_start, q1 = h.q_start + h.len - 1;
                const int s0 = h.s_start, s1 = h.s_start + h.len - 1;
                bool skip = false;
                for (int t = 0; t < ns && !skip; t++) {
                    const Hsp &u = saved[t];
                    skip = h.score <= u.score && u.q_off <= q0 && q0 <= u.q_end && u.s_off <= s0 && s0 <= u.s_end &&
                           u.q_off <= q1 && q1 <= u.q_end && u.s_off <= s1 && s1 <= u.s_end;
                }
                if (!skip) {
                    q_gs = gapped_start(mat, Q, Sj, h.q_start, h.s_start, h.len);
                    s_gs = h.s_start + (q_gs - h.q_start);
                    found = true;
                    break;
                }
                k++;
            }
            if (found) {
                dp_begin(0);
                phase = kDp;
            } else {  // all init hits done: the subject's list out
                if (writer) {
                    int r = kRedo;
                    if (!redo) {
                        cells += item_cells;
                        gext += item_ext;
                        r = finish_hsps(a, wi, saved, ns);
                    }
                    if (r == kRedo) {
                        const uint32_t kk = atomicAdd(&a.ctr->redo_count, 1u);
                        if (kk < a.redo_cap) a.redo[kk] = widx;
                        else record_overflow(a, widx);
                    } else if (r == kOverflow) {
                        record_overflow(a, widx);
                    }
                }
                phase = kTake;
            }
        } else if (phase == kDp) {  // one row of the DP
            bool done = dps != kRun || arow > M;
            if (!done) {
                const int16_t *mrow = mat[rev ? A[M - arow] : A[arow]];
                int carry_g = kMinIntDp;   // max of max(D,E)(b') + b' * extend over earlier chunks
                int carry_h = best;        // running best over earlier chunks
                int carry_hp = kMinIntDp;  // previous row's H left of this chunk
                int fa = 0x7fffffff, la = -1;  // this lane's first and last surviving column
                int rmax = kMinIntDp, rmax_b = -1;
                const int rs = first;      // kPk: the row's first column (gap offsets count from it)
                if constexpr (kPk) {
                    s16x2 cg = pk_splat(-32768), ch = pk_splat(best);  // carry_g, carry_h
                    uint32_t hp_old = kNeg16x2;  // high half: previous row's H left of the chunk
                    s16x2 bge = lane_bge;        // (column - rs) x extend, per column
                    const s16x2 go2 = pk_splat(gopen), goe2 = pk_splat(goe), ge2 = pk_splat(ge), X2 = pk_splat(X);
                    const s16x2 step2 = pk_splat(32 * ge);
                    for (int c0 = first; c0 < size; c0 += 32) {
                        const int b0 = c0 + 2 * sl;
                        // lanes past the band read ring and letter slots they do not use
                        const uint32_t w0 = ring[b0 & wmask], w1 = ring[(b0 + 1) & wmask];
                        const s16x2 hprev = as_s16x2(__builtin_amdgcn_perm(w1, w0, 0x05040100u));
                        const s16x2 eprev = as_s16x2(__builtin_amdgcn_perm(w1, w0, 0x07060302u));
                        const int sc0 = mrow[lett[b0 & (kLetRing - 1)]], sc1 = mrow[lett[(b0 + 1) & (kLetRing - 1)]];
                        const s16x2 sb = as_s16x2(__builtin_amdgcn_perm((uint32_t)sc1, (uint32_t)sc0, 0x05040100u));
                        // H of the column left of each: the previous lane's high column, then our low one
                        const uint32_t hsh = __builtin_amdgcn_update_dpp(hp_old, as_u32(hprev), 0x111, 0xf, 0xf, false);
                        const s16x2 hpl = as_s16x2(__builtin_amdgcn_perm(as_u32(hprev), hsh, 0x05040302u));
                        const int cnt = size - b0;  // columns of this lane inside the band: 2, 1 or none
                        const uint32_t vm = cnt >= 2 ? 0xFFFFFFFFu : (cnt == 1 ? 0x0000FFFFu : 0u);
                        const s16x2 de = as_s16x2(pk_select(vm, as_u32(pk_max(pk_add(hpl, sb), eprev)), kNeg16x2));
                        const s16x2 gin = pk_incl_max16(pk_add(de, bge));
                        const s16x2 din = pk_incl_max16(de);
                        const uint32_t gsh = __builtin_amdgcn_update_dpp(kNeg16x2, as_u32(gin), 0x111, 0xf, 0xf, false);
                        const s16x2 gex = pk_max(as_s16x2(__builtin_amdgcn_perm(as_u32(gin), gsh, 0x05040302u)), cg);
                        const s16x2 h = pk_max(de, pk_sub(gex, pk_add(bge, go2)));
                        // alive: X - ((best before the cell) - H) >= 0, both halves at once
                        const s16x2 dead = pk_sub(X2, pk_sub(pk_max(din, ch), h)) >> 15;
                        const uint32_t alive = ~as_u32(dead) & vm;
                        const bool al0 = (alive & 0xFFFFu) != 0, al1 = (alive >> 16) != 0;
                        fa = min(fa, al0 ? b0 : (al1 ? b0 + 1 : 0x7fffffff));
                        la = al1 ? b0 + 1 : (al0 ? b0 : la);
                        const s16x2 cmv = as_s16x2(S::bcast((int)as_u32(din)));  // high half: the chunk's max
                        const int cm = cmv.y;
                        if (cm > ch.y) {  // a new best in this chunk: its first cell (H equals D/E there)
                            rmax = cm;
                            const uint64_t m0 = S::ballot(de.x == cm), m1 = S::ballot(de.y == cm);
                            const int c_lo = m0 ? 2 * (int)__builtin_ctzll(m0) : 0x7fffffff;
                            const int c_hi = m1 ? 2 * (int)__builtin_ctzll(m1) + 1 : 0x7fffffff;
                            rmax_b = c0 + min(c_lo, c_hi);
                        }
                        // new (H, E) for the band's columns (minus infinity where the X-drop fails);
                        // a column past the band gets its own old word back
                        const uint32_t hs = pk_select(alive, as_u32(h), kNeg16x2);
                        const uint32_t es = pk_select(alive, as_u32(pk_max(pk_sub(h, goe2), pk_sub(eprev, ge2))), kNeg16x2);
                        const uint32_t ho = pk_select(vm, hs, as_u32(hprev)), eo = pk_select(vm, es, as_u32(eprev));
                        ring[b0 & wmask] = __builtin_amdgcn_perm(eo, ho, 0x05040100u);
                        ring[(b0 + 1) & wmask] = __builtin_amdgcn_perm(eo, ho, 0x07060302u);
                        cg = pk_max(cg, as_s16x2(S::bcast((int)as_u32(gin))).yy);
                        ch = pk_max(ch, cmv.yy);
                        hp_old = (uint32_t)S::bcast((int)as_u32(hprev));
                        bge = pk_add(bge, step2);
                    }
                    carry_g = cg.x;
                } else
                for (int c0 = first; c0 < size; c0 += 16) {
                    const int b = c0 + sl;
                    const bool valid = b < size;
                    // lanes past the band read ring and letter slots they do not use
                    const int2 prev = ring[b & wmask];
                    const int sb = mrow[lett[b & (kLetRing - 1)]];
                    const int hp_left = S::shift_up1(prev.x, carry_hp);
                    const int de = valid ? max(hp_left + sb, prev.y) : kMinIntDp;
                    const int g = valid ? de + b * ge : kMinIntDp;
                    // The running best before a cell is a prefix max of max(D,E) alone: a row gap
                    // F(b) is at most max(D,E)(b') - open for some earlier b', so it never sets a
                    // new best.  The two prefix scans are then independent.
                    const int gin = S::incl_max(g);
                    const int din = S::incl_max(de);
                    const int gex = max(S::shift_up1(gin, kMinIntDp), carry_g);
                    const int h = max(de, gex - goe - (b - 1) * ge);
                    // (best before the cell) - H <= X, with the cell's own score counted: equal
                    // when H is below that best, and 0 <= X otherwise
                    const bool alive = valid && max(din, carry_h) - h <= X;
                    fa = min(fa, alive ? b : 0x7fffffff);
                    la = alive ? b : la;
                    const int cm = S::bcast(din);
                    if (cm > carry_h) {  // a new best in this chunk: its first cell (H equals D/E there)
                        rmax = cm;
                        rmax_b = c0 + (int)__builtin_ctzll(S::ballot(de == cm));
                    }
                    if (valid)
                        ring[b & wmask] = alive ? make_int2(h, max(h - goe, prev.y - ge)) : make_int2(kMinIntDp, kMinIntDp);
                    carry_g = max(carry_g, S::last(gin));
                    carry_h = max(carry_h, cm);
                    carry_hp = S::bcast(prev.x);
                }
                nc += (unsigned long long)(size - first);
                la = S::reduce_max(la);
                fa = -S::reduce_max(-fa);
                if (la < 0) {
                    done = true;
                } else {
                    if (rmax > best) {
                        best = rmax;
                        ba = arow;
                        bb = rmax_b;
                    }
                    first = fa;
                    if (la < size - 1) {
                        size = la + 1;
                    } else {
                        // row gap one past the last cell (kPk: carry_g counts columns from rs)
                        const int gr = carry_g - goe - (size - 1 - (kPk ? rs : 0)) * ge;
                        if (gr >= best - X) {
                            // cells of the run (a division only for a gap extension other than 1)
                            int kk = (ge == 1 ? gr - (best - X) : (gr - (best - X)) / ge) + 1;
                            if (kk > N + 1 - size) kk = N + 1 - size;
                            if (size + kk - first > W) {
                                dps = kOvf;
                            } else {
                                for (int i = sl; i < kk; i += 16) {
                                    const int hv = gr - i * ge;
                                    ring[(size + i) & wmask] = cell(hv, hv - goe);
                                }
                                size += kk;
                            }
                        }
                    }
                    if (dps == kRun && size <= N) {
                        if (size + 1 - first > W) {
                            dps = kOvf;
                        } else {
                            if (sl == 0) ring[size & wmask] = neg_cell();
                            size++;
                        }
                    }
                    if (dps == kRun) fill_letters();
                    arow++;
                    done = dps != kRun || arow > M;
                }
            }
            if (done) {
                const bool empty = dps == kEmpty;
                const int score = empty ? 0 : best;
                const int oa = empty ? 0 : ba, ob = empty ? 0 : bb;
                if (dps == kOvf) {  // band past the ring: the subject goes to the wide pass
                    redo = true;
                    k = n;
                    phase = kHit;
                } else if (dir == 0) {
                    if (!empty) item_cells += nc;
                    left = score;
                    la_l = oa;
                    lb_l = ob;
                    dp_begin(1);
                } else {
                    if (!empty) item_cells += nc;
                    item_ext++;
                    const int total = left + score;
                    if (total >= gcut) {
                        if (ns >= a.saved_cap) {  // more HSPs than this pass can hold
                            redo = true;
                            k = n;
                        } else {
                            if (writer)
                                saved[ns] = Hsp{total, q_gs + 1 - la_l, q_gs + 1 + oa, s_gs + 1 - lb_l, s_gs + 1 + ob, q_gs, s_gs};
                            ns++;
                        }
                    }
                    if (k < n) k++;
                    phase = kHit;
                }
            }
        }
        wave_sync();
        if (__ballot(phase != kDone) == 0) break;
    }
    if (writer) {
        atomicAdd(&a.ctr->dp_cells, cells);
        atomicAdd(&a.ctr->gapped_ext, gext);
    }
}

int gapped_seg_blocks_per_cu(bool packed) {
    static const int n32 = [] {
        int k = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&k, k_gapped_seg<false>, 64, 0) != hipSuccess || k <= 0) k = 8;
        return k;
    }();
    static const int n16 = [] {
        int k = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&k, k_gapped_seg<true>, 64, 0) != hipSuccess || k <= 0) k = 8;
        return k;
    }();
    return packed ? n16 : n32;
}

hipError_t launch_gapped_seg(const GappedArgs &a, int blocks, bool packed, hipStream_t st) {
    if (packed)
        hipLaunchKernelGGL(k_gapped_seg<true>, dim3(blocks), dim3(64), 0, st, a);
    else
        hipLaunchKernelGGL(k_gapped_seg<false>, dim3(blocks), dim3(64), 0, st, a);
    return hipGetLastError();
}

int gapped_wave_blocks_per_cu() {
    static const int n = [] {
        int k = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&k, k_gapped_wave, 64, 0) != hipSuccess || k <= 0) k = 8;
        return k;
    }();
    return n;
}

hipError_t launch_gapped_wave(const GappedArgs &a, int blocks, hipStream_t st) {
    hipLaunchKernelGGL(k_gapped_wave, dim3(blocks), dim3(64), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_order(const WorkItem *work, DevCounters *ctr, uint32_t *order, uint32_t work_cap, int blocks,
                        hipStream_t st) {
    hipLaunchKernelGGL(k_order, dim3(blocks), dim3(256), 0, st, work, ctr, order, work_cap);
    return hipGetLastError();
}

hipError_t launch_gapped_lds(const GappedArgs &a, int blocks, hipStream_t st) {
    hipLaunchKernelGGL(k_gapped_lds, dim3(blocks), dim3(64), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_gapped_wide(const GappedArgs &a, int blocks, int threads, hipStream_t st) {
    hipLaunchKernelGGL(k_gapped_wide, dim3(blocks), dim3(threads), 0, st, a);
    return hipGetLastError();
}

}  // namespace blz
