// blz_traceback.hip -- traceback stage: the gapped alignment with traceback of the prelim HSPs of
// the (query, subject) pairs a caller lists (normally each query's hit list), one wave per pair.
//
// Per pair (NCBI BLAST+ 2.13.0 Blast_TracebackFromHSPList for blastp without composition-based
// rescoring: the stage that consumes the lists BLAZE's prelim stage writes to the HSP stream at
// BLAZE_ncbi/blast_engine.c:2499, which the reference leaves to NCBI's CPU code):
//   the prelim HSPs in their score order; one contained in an HSP already traced (score no higher,
//   both end points inside its box; NCBI s_HSPIsContained) is skipped; otherwise the alignment
//   with traceback (BLAST_GappedAlignmentWithTraceback) runs ALIGN_EX left over the reversed
//   prefixes (the gapped start included) and right over the suffixes with the final X-drop and
//   reads each path back from its best cell.  An HSP scoring at least the cutoff is kept with its
//   edit script (runs of one operation: an aligned pair, a subject letter against a gap, a query
//   letter against a gap), identities, columns, gap letters and gap openings.  Then the
//   common-endpoint purge and the score sort of the prelim stage.
//
// ALIGN_EX here is row-parallel: a row's band is spread over the 64 lanes, and the row gap and the
// running best are prefix maxima across the lanes (as in the prelim stage's wave DP).  Each cell
// stores the serial loop's traceback byte: the operation that gave the cell its score (the
// diagonal unless a gap score is strictly greater, the row gap last) and, per gap direction,
// whether the gap leaving the cell extends the one entering it.  The prefix-max form has the
// serial loop's gap values at every cell the X-drop keeps, except where the serial loop carries a
// gap past a dropped cell; such a gap lies below the X-drop bound, so it can neither give a kept
// cell its score nor lie on the path read back, and every byte the path reads is the serial one.
//
// The same code runs a pair per segment of kSeg lanes: by default a whole wave per pair, its
// pair-wide values in scalar registers; or (tuning) 16-lane DPP rows, four pairs per wave, each
// segment with its own LDS ring, letter windows and walk window and its values held per lane, a
// segment that finishes a pair taking the next.  Bands are narrow (about 25 cells per row at the
// final X-drop), but the segments' divergent phases and the vector bookkeeping cost more than the
// idle lanes of a wave-wide row.  The main pass bounds each pair's traceback rows and
// edit-script scratch; a pair outgrowing them (or the LDS ring) goes to the wide pass, one pair
// per wave with a global ring and storage sized to the largest such pair.
#include <hip/hip_runtime.h>

#include "blz_common.h"
#include "blz_dev_common.h"
#include "blz_device.h"

namespace blz {
namespace {

// traceback byte of a cell
enum : int { kTbSub = 0, kTbGapA = 1, kTbGapB = 2, kTbOpMask = 3, kTbExtA = 4, kTbExtB = 8 };
constexpr int kNoCell = 0xFF;  // window byte outside the stored rows
constexpr uint32_t kRunLen = 0x0FFFFFFFu;
constexpr int kMatRow = 32;  // LDS matrix row: 28 scores and 4 unused, so a row's address is a shift

enum PairOutcome { kPairDone = 0, kPairBack = 1, kPairBad = 2 };
enum WalkOutcome { kWalkOk = 0, kWalkFull = 1, kWalkBad = 2 };

// Segment geometry and the per-segment LDS: the subject letters of the band (a window over the
// columns), the query letters of the coming rows, and the path-walk window (kWinRows rows x
// kWinCols columns).
template <int kSeg>
struct Tb {
    using S = Seg<kSeg>;
    static constexpr int kRing = kSeg == 64 ? kTbWaveRing : kTbSegRing;  // LDS DP ring of the main pass
    static constexpr int kBWin = kSeg == 64 ? 512 : 256;
    static constexpr int kAWin = kSeg == 64 ? 256 : 128;
    static constexpr int kWinRows = 16, kWinCols = kSeg == 64 ? 32 : 16;
    static constexpr int kLanesPerRow = kSeg / kWinRows, kBytesPerLane = kWinCols / kLanesPerRow;
    static_assert(kBWin >= kRing + kSeg, "a row's band and its last step fit the letter window");
    struct Lds {
        uint8_t bwin[kBWin];
        uint8_t awin[kAWin];
        uint8_t win[kWinRows * kWinCols];
    };
    // A value every lane of the segment holds alike: for a whole wave moved to a scalar register
    // (loads through plain pointers land in vector registers, and control flow on them would be
    // compiled as divergent); a segment keeps it per lane.
    static __device__ __forceinline__ int U(int x) {
        if constexpr (kSeg == 64) return __builtin_amdgcn_readfirstlane(x);
        return x;
    }
    static __device__ __forceinline__ uint64_t U64(uint64_t x) {
        if constexpr (kSeg == 64) return first_lane64(x);
        return x;
    }
    // the segment's lane 0 value, in every lane of the segment
    static __device__ __forceinline__ int first(int x) { return S::first(x); }
    static __device__ __forceinline__ uint64_t first64(uint64_t x) {
        return ((uint64_t)(uint32_t)S::first((int)(x >> 32)) << 32) | (uint32_t)S::first((int)(uint32_t)x);
    }
};

// Column b's letter of B: B[N - b] on reversed prefixes, else B[b] (column 0 has none).
template <int kSeg>
__device__ __forceinline__ void load_bwin(typename Tb<kSeg>::Lds &L, const uint8_t *B, int N, bool reverse, int bw0) {
    for (int i = Seg<kSeg>::lane(); i < Tb<kSeg>::kBWin; i += kSeg) {
        const int b = bw0 + i;
        L.bwin[i] = (b >= 1 && b <= N) ? (reverse ? B[N - b] : B[b]) : 0;
    }
}
template <int kSeg>
__device__ __forceinline__ void load_awin(typename Tb<kSeg>::Lds &L, const uint8_t *A, int M, bool reverse, int aw0) {
    for (int i = Seg<kSeg>::lane(); i < Tb<kSeg>::kAWin; i += kSeg) {
        const int a = aw0 + i;
        L.awin[i] = (a >= 1 && a <= M) ? (reverse ? A[M - a] : A[a]) : 0;
    }
}

// ALIGN_EX of one direction by one segment (arguments alike across it).  Rows walk A (a = 1..M),
// the X-drop band walks B; row 0 is the gap run in A.  Row a's traceback bytes, columns
// [rows[a].x, rows[a].x + rows[a].z), lie at tb[rows[a].y ...].  Returns the best score, its cell
// in (*a_off, *b_off); sets *ovf when the band outgrows the ring or the rows outgrow tb_cap or
// rows_cap (nothing usable is left then).  The row table is staged in registers: the lane whose
// index is the row's index mod kSeg keeps the row, and a block of kSeg rows goes out in one
// store when it is full and at the end.
template <int kSeg>
__device__ int align_ex_seg(const int16_t (*mat)[kMatRow], typename Tb<kSeg>::Lds &L, int2 *ring, int wmask,
                            const uint8_t *A, const uint8_t *B, int M, int N, int X, int go, int ge, bool reverse,
                            uint8_t *tb, unsigned long long tb_cap, int4 *rows, int rows_cap, int *a_off, int *b_off,
                            bool *ovf, unsigned long long *cells) {
    using T = Tb<kSeg>;
    using S = Seg<kSeg>;
    const int sl = S::lane();
    const int goe = go + ge;
    const int lane_ge = sl * ge;
    *a_off = 0;
    *b_off = 0;
    if (X < goe) X = goe;
    if (N <= 0 || M <= 0) return 0;
    const int W = wmask + 1;
    const int k0 = (X - goe) / ge + 1;
    int size = (N < k0 ? N : k0) + 1;
    if (M + 1 > rows_cap || size > W || (unsigned long long)size > tb_cap) {
        *ovf = true;
        return 0;
    }
    for (int b = sl; b < size; b += kSeg) {
        const int h = b == 0 ? 0 : -goe - (b - 1) * ge;
        ring[b & wmask] = make_int2(h, h - goe);
        tb[b] = kTbGapA;
    }
    int4 rreg = make_int4(0, 0, size, 0);  // row 0 (this lane's staged row while sl == 0)
    // stages row a (first column, byte offset, length); a full block of kSeg rows goes out
    auto stage_row = [&](int a, int rs, int ro, int len) {
        const bool mine = sl == (a & (kSeg - 1));
        rreg.x = mine ? rs : rreg.x;
        rreg.y = mine ? ro : rreg.y;
        rreg.z = mine ? len : rreg.z;
        if ((a & (kSeg - 1)) == kSeg - 1) rows[a - (kSeg - 1) + sl] = rreg;
    };
    // the rows of the last, partial block [a & ~(kSeg - 1), a]
    auto flush_rows = [&](int a) {
        if ((a & (kSeg - 1)) != kSeg - 1 && sl <= (a & (kSeg - 1))) rows[(a & ~(kSeg - 1)) + sl] = rreg;
    };
    uint32_t used = (uint32_t)size;
    int bw0 = 0, aw0 = 1;
    wave_sync();
    load_bwin<kSeg>(L, B, N, reverse, bw0);
    load_awin<kSeg>(L, A, M, reverse, aw0);
    wave_sync();
    int best = 0, first = 0, ba = 0, bb = 0;
    unsigned long long nc = 0;
    for (int a = 1; a <= M; a++) {
        first = T::U(first);
        size = T::U(size);
        bw0 = T::U(bw0);
        best = T::U(best);
        if (a >= aw0 + T::kAWin) {
            aw0 = a;
            load_awin<kSeg>(L, A, M, reverse, aw0);
            wave_sync();
        }
        const int16_t *mrow = mat[L.awin[a - aw0]];
        const int rs = first;
        used = (uint32_t)T::U((int)used);
        if ((unsigned long long)used + (unsigned long long)(size - rs) > tb_cap) {
            *ovf = true;
            return 0;
        }
        uint8_t *trow = tb + ((long long)used - rs);  // trow[b] for b >= rs
        int carry_g = kMinIntDp;   // max of max(D,E)(b') + b' * extend over earlier steps
        int carry_h = best;        // running best over earlier steps
        int carry_hp = kMinIntDp;  // previous row's H left of this step (none for the first cell)
        int fa = -1, la = -1, rmax = kMinIntDp, rmax_b = -1;
        for (int c0 = first; c0 < size; c0 += kSeg) {
            const int c1 = min(c0 + kSeg, size);
            if (c0 < bw0 || c1 > bw0 + T::kBWin) {  // the window moves to the row's band, or this step
                bw0 = c1 - first <= T::kBWin ? first : c0;
                load_bwin<kSeg>(L, B, N, reverse, bw0);
                wave_sync();
            }
            const int b = c0 + sl;
            const bool valid = b < size;
            const int bge = c0 * ge + lane_ge;  // b * extend
            const int2 prev = valid ? ring[b & wmask] : make_int2(kMinIntDp, kMinIntDp);
            const int hp_left = S::shift_up1(prev.x, carry_hp);
            const int sb = (valid && b >= 1) ? mrow[L.bwin[b - bw0]] : 0;
            const int D = hp_left + sb;
            const int E = prev.y;
            const int de = valid ? max(D, E) : kMinIntDp;
            const int g = valid ? de + bge : kMinIntDp;
            const int gin = S::incl_max(g);
            const int gex = max(S::shift_up1(gin, kMinIntDp), carry_g);
            const int f = gex - go - bge;  // row gap entering the cell: gex - open - extend * (b - 1) - extend
            const int h = valid ? max(de, f) : kMinIntDp;
            const int hin = S::incl_max(h);
            const int hex = max(S::shift_up1(hin, kMinIntDp), carry_h);
            const bool alive = valid && hex - h <= X;
            const uint64_t am = S::ballot(alive);
            if (am) {
                if (fa < 0) fa = c0 + (int)__builtin_ctzll(am);
                la = c0 + 63 - (int)__builtin_clzll(am);
            }
            const int cm = S::last(hin);
            if (cm > carry_h) {
                rmax = cm;
                rmax_b = c0 + (int)__builtin_ctzll(S::ballot(valid && h == cm));
            }
            if (valid) {
                int t = f > de ? kTbGapA : (E > D ? kTbGapB : kTbSub);
                if (alive) {
                    if (E - ge >= h - goe) t |= kTbExtB;  // the column gap leaving the cell extends
                    if (f - ge >= h - goe) t |= kTbExtA;  // the row gap leaving the cell extends
                }
                trow[b] = (uint8_t)t;
                ring[b & wmask] = alive ? make_int2(h, max(h - goe, E - ge)) : make_int2(kMinIntDp, kMinIntDp);
            }
            carry_g = max(carry_g, S::last(gin));
            carry_h = max(carry_h, cm);
            carry_hp = S::last(prev.x);
        }
        nc += (unsigned long long)(size - first);
        int row_end = size;
        if (fa < 0) {
            stage_row(a, rs, (int)used, row_end - rs);
            flush_rows(a);
            break;
        }
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
                if (size + k - first > W || (unsigned long long)used + (unsigned long long)(size + k - rs) > tb_cap) {
                    *ovf = true;
                    return 0;
                }
                for (int i = sl; i < k; i += kSeg) {
                    const int hv = gr - i * ge;
                    ring[(size + i) & wmask] = make_int2(hv, hv - goe);
                    trow[size + i] = kTbGapA;
                }
                size += k;
                row_end = size;
            }
        }
        stage_row(a, rs, (int)used, row_end - rs);
        if (a == M) flush_rows(a);
        used += (uint32_t)(row_end - rs);
        if (size <= N) {
            if (size + 1 - first > W) {
                *ovf = true;
                return 0;
            }
            ring[size & wmask] = make_int2(kMinIntDp, kMinIntDp);  // every lane, the same value
            size++;
        }
        wave_sync();
    }
    *a_off = ba;
    *b_off = bb;
    *cells += nc;
    return best;
}

// Reads the path back from cell (a, b) to the origin (NCBI ALIGN_EX's traceback: a cell reached
// through a gap continues that gap while its extension bit for that direction is set, else takes
// the cell's own operation) and appends its operations, as runs op << 28 | length in walk order,
// to runs[*n, cap).  The rows come through the segment's LDS window of kWinRows x kWinCols cells,
// fetched by the segment around the current cell whenever the walk leaves it.  Operation codes:
// the byte's kTbGapA (a B letter against a gap in A) and kTbGapB (an A letter against a gap in B).
template <int kSeg>
__device__ int walk_path(typename Tb<kSeg>::Lds &L, const uint8_t *tb, const int4 *rows, int a, int b, uint32_t *runs,
                         int cap, int *n) {
    using T = Tb<kSeg>;
    const int sl = Seg<kSeg>::lane();
    int cur = kTbSub, run_op = -1, run_len = 0, nr = *n;
    int wa = -1, wb = -1;  // window: rows (wa - kWinRows, wa], columns (wb - kWinCols, wb]
    while (a > 0 || b > 0) {
        if (a > wa || a <= wa - T::kWinRows || b > wb || b <= wb - T::kWinCols) {
            wa = a;
            wb = b;
            wave_sync();
            const int wr = sl / T::kLanesPerRow, wc = (sl % T::kLanesPerRow) * T::kBytesPerLane;
            const int r = wa - wr;  // lane: one row, kBytesPerLane of its columns
            const int4 m = r >= 0 ? rows[r] : make_int4(0, 0, 0, 0);
            uint8_t *w = &L.win[wr * T::kWinCols + wc];
#pragma unroll
            for (int j = 0; j < T::kBytesPerLane; j++) {
                const int i = wb - wc - j - m.x;
                w[j] = (r >= 0 && i >= 0 && i < m.z) ? tb[(unsigned)m.y + (unsigned)i] : (uint8_t)kNoCell;
            }
            wave_sync();
        }
        const int next = T::U(L.win[(wa - a) * T::kWinCols + (wb - b)]);
        if (next == kNoCell) return kWalkBad;
        if (cur == kTbGapA)
            cur = (next & kTbExtA) ? kTbGapA : (next & kTbOpMask);
        else if (cur == kTbGapB)
            cur = (next & kTbExtB) ? kTbGapB : (next & kTbOpMask);
        else
            cur = next & kTbOpMask;
        if (cur == kTbGapA) {
            b--;
        } else if (cur == kTbGapB) {
            a--;
        } else {
            a--;
            b--;
        }
        if (a < 0 || b < 0 || cur == kTbOpMask) return kWalkBad;
        if (cur == run_op) {
            run_len++;
        } else {
            if (run_op >= 0) {
                if (nr >= cap) return kWalkFull;
                if (sl == 0) runs[nr] = ((uint32_t)run_op << 28) | (uint32_t)run_len;
                nr++;
            }
            run_op = cur;
            run_len = 1;
        }
    }
    if (run_op >= 0) {
        if (nr >= cap) return kWalkFull;
        if (sl == 0) runs[nr] = ((uint32_t)run_op << 28) | (uint32_t)run_len;
        nr++;
    }
    *n = nr;
    return kWalkOk;
}

// Run i of the HSP's edit script left to right: the left walk's runs wr[0, nl) as they are, then
// the right walk's wr[nl, nl + nr) backwards, the two runs at the junction merged when they hold
// the same operation.
__device__ __forceinline__ uint32_t final_run(const uint32_t *wr, int nl, int nr, bool merge, int i) {
    if (i < nl) {
        const uint32_t v = wr[i];
        return (merge && i == nl - 1) ? v + (wr[nl + nr - 1] & kRunLen) : v;
    }
    return wr[nl + nr - 1 - (i - nl + (merge ? 1 : 0))];
}

__device__ __forceinline__ bool contained(const Hsp &h, const Hsp &o) {
    return h.score <= o.score && o.q_off <= h.q_off && h.q_off <= o.q_end && o.s_off <= h.s_off &&
           h.s_off <= o.s_end && o.q_off <= h.q_end && h.q_end <= o.q_end && o.s_off <= h.s_end && h.s_end <= o.s_end;
}

// The traceback of one pair by one segment.  Traced HSPs are built in out[0, n) (the segment's
// lane 0 writes), purged and sorted there; returns the outcome and the kept count in *kept.
template <int kSeg>
__device__ int trace_pair(const TbArgs &a, const int16_t (*mat)[kMatRow], typename Tb<kSeg>::Lds &L, int2 *ring, int wmask,
                          uint8_t *tb, int4 *rows, uint32_t *wr, uint32_t pi, int *kept, unsigned long long *cells,
                          unsigned long long *dps) {
    using T = Tb<kSeg>;
    using S = Seg<kSeg>;
    const int sl = S::lane();
    const TbPair pv = a.pairs[pi];
    const int64_t sid = (int64_t)T::U64((uint64_t)pv.sid);
    const uint64_t off = T::U64(pv.off);
    const int qi = T::U(pv.qi), n_pre = T::U(pv.n);
    const TbQuery Qv = a.qs[qi];
    const uint8_t *qs = reinterpret_cast<const uint8_t *>(T::U64(reinterpret_cast<uint64_t>(Qv.q)));
    const int qlen = T::U(Qv.qlen), cutoff = T::U(Qv.cutoff);
    const uint64_t s0 = T::U64(a.offs[sid]);
    const uint8_t *s = a.res + s0;
    const int slen = T::U((int)(a.offs[sid + 1] - s0));
    TbHsp *out = a.out + off;
    const Hsp *pre = a.pre + off;
    int nt = 0;
    for (int k = 0; k < n_pre; k++) {
        const Hsp hv = pre[k];
        const Hsp h{T::U(hv.score), T::U(hv.q_off), T::U(hv.q_end), T::U(hv.s_off), T::U(hv.s_end), T::U(hv.q_gs),
                    T::U(hv.s_gs)};
        bool skip = false;
        for (int t0 = 0; t0 < nt && !skip; t0 += kSeg) {
            const int t = t0 + sl;
            skip = S::ballot(t < nt && contained(h, out[t].h)) != 0;
        }
        if (skip) continue;
        bool ovf = false;
        int la, lb, ra, rb;
        const int left = align_ex_seg<kSeg>(mat, L, ring, wmask, qs, s, h.q_gs + 1, h.s_gs + 1, a.x_final, a.gap_open,
                                            a.gap_extend, true, tb, a.tb_cap, rows, a.rows_cap, &la, &lb, &ovf, cells);
        if (ovf) return kPairBack;
        int nl = 0;
        int ws = walk_path<kSeg>(L, tb, rows, la, lb, wr, a.wruns_cap, &nl);
        if (ws != kWalkOk) return ws == kWalkFull ? kPairBack : kPairBad;
        const int right = align_ex_seg<kSeg>(mat, L, ring, wmask, qs + h.q_gs, s + h.s_gs, qlen - h.q_gs - 1,
                                             slen - h.s_gs - 1, a.x_final, a.gap_open, a.gap_extend, false, tb,
                                             a.tb_cap, rows, a.rows_cap, &ra, &rb, &ovf, cells);
        if (ovf) return kPairBack;
        *dps += 2;
        const int score = left + right;
        if (score < cutoff) continue;
        int nall = nl;
        ws = walk_path<kSeg>(L, tb, rows, ra, rb, wr, a.wruns_cap, &nall);
        if (ws != kWalkOk) return ws == kWalkFull ? kPairBack : kPairBad;
        wave_sync();  // the segment's lane 0 wrote wr[]
        const int nr = nall - nl;
        const bool merge = nl > 0 && nr > 0 && (wr[nl - 1] >> 28) == (wr[nl + nr - 1] >> 28);
        const int cnt = nl + nr - (merge ? 1 : 0);
        unsigned long long base = 0;
        if (sl == 0) base = atomicAdd(&a.ctr->runs_used, (unsigned long long)cnt);
        base = T::first64(base);
        const bool fits = base + (unsigned long long)cnt <= a.runs_cap;
        // columns, gaps and identities along the script; the script places the runs as it goes
        const int q_off = h.q_gs + 1 - la, s_off = h.s_gs + 1 - lb;
        int qp = q_off, sp = s_off, ident = 0, columns = 0, gaps = 0, opens = 0;
        for (int i = 0; i < cnt; i++) {
            const uint32_t v = (uint32_t)T::U((int)final_run(wr, nl, nr, merge, i));
            if (fits && sl == 0) a.runs[base + i] = v;
            const int op = (int)(v >> 28), len = (int)(v & kRunLen);
            columns += len;
            if (op == kTbSub) {
                for (int c0 = 0; c0 < len; c0 += kSeg) {
                    const int c = c0 + sl;
                    ident += __popcll(S::ballot(c < len && qs[qp + c] == s[sp + c]));
                }
                qp += len;
                sp += len;
            } else {
                gaps += len;
                opens++;
                if (op == kTbGapA)
                    sp += len;
                else
                    qp += len;
            }
        }
        if (qp != h.q_gs + 1 + ra || sp != h.s_gs + 1 + rb) return kPairBad;
        if (sl == 0) {
            TbHsp o;
            o.h = Hsp{score, q_off, h.q_gs + 1 + ra, s_off, h.s_gs + 1 + rb, h.q_gs, h.s_gs};
            o.num_ident = ident;
            o.align_len = columns;
            o.gaps = gaps;
            o.gap_opens = opens;
            o.op_first = (int32_t)(fits ? base : 0);
            o.op_count = cnt;
            out[nt] = o;
        }
        nt++;
        wave_sync();  // out[] is read by every lane of the segment in the containment test
    }
    // common-endpoint purge and score order (the segment's lane 0; lists are short)
    int m = 0;
    if (sl == 0 && nt > 0) {
        insertion_sort<cmp_start>(out, nt);
        for (int i = 0; i < nt; i++) {
            if (m > 0 && out[m - 1].h.q_off == out[i].h.q_off && out[m - 1].h.s_off == out[i].h.s_off) continue;
            out[m++] = out[i];
        }
        int m2 = 0;
        insertion_sort<cmp_end>(out, m);
        for (int i = 0; i < m; i++) {
            if (m2 > 0 && out[m2 - 1].h.q_end == out[i].h.q_end && out[m2 - 1].h.s_end == out[i].h.s_end) continue;
            out[m2++] = out[i];
        }
        m = m2;
        insertion_sort<cmp_score>(out, m);
    }
    *kept = T::first(m);
    return kPairDone;
}

// kSeg lanes per pair; kWide: the wide pass (one pair per wave, global DP rings sized by the
// host, pairs from a.list).  Each segment takes pairs from the work-stealing counter on its own.
template <int kSeg, bool kWide>
__global__ void __launch_bounds__(64) k_traceback(TbArgs a) {
    using T = Tb<kSeg>;
    using S = Seg<kSeg>;
    constexpr int kSegs = 64 / kSeg;
    __shared__ int16_t mat[28][kMatRow];
    __shared__ typename T::Lds lds[kSegs];
    __shared__ int2 lring[kWide ? 1 : kSegs * T::kRing];
    const int lane = lane_id();
    const int seg = lane / kSeg;
    const int sl = S::lane();
    for (int i = lane; i < 28 * kMatRow; i += 64) mat[i / kMatRow][i % kMatRow] = i % kMatRow < 28 ? kBlosum62Dev[i / kMatRow][i % kMatRow] : 0;
    wave_sync();
    const size_t w = (size_t)blockIdx.x * kSegs + seg;  // this segment's storage
    uint8_t *tb = a.tb + w * a.tb_cap;
    int4 *rows = a.rows + w * (size_t)a.rows_cap;
    uint32_t *wr = a.wruns + w * (size_t)a.wruns_cap;
    int2 *ring = kWide ? a.ring + w * ((size_t)a.ring_mask + 1) : lring + seg * T::kRing;
    const int wmask = kWide ? a.ring_mask : T::kRing - 1;
    const uint32_t n_items = a.list ? a.n_list : a.n_pairs;
    unsigned long long cells = 0, dps = 0;
    for (;;) {
        int k = 0;
        if (sl == 0) k = (int)atomicAdd(&a.ctr->next_pair, 1u);
        k = T::first(k);
        if ((uint32_t)k >= n_items) break;
        const uint32_t pi = a.list ? a.list[k] : (uint32_t)k;
        int kept = 0;
        const int r = trace_pair<kSeg>(a, mat, lds[seg], ring, wmask, tb, rows, wr, pi, &kept, &cells, &dps);
        if (sl == 0) {
            if (r == kPairDone) {
                a.out_n[pi] = kept;
            } else if (r == kPairBack) {  // the wide pass counts the pairs it could not take either
                a.out_n[pi] = -1;
                const uint32_t j = atomicAdd(&a.ctr->back_count, 1u);
                if (!kWide && j < a.back_cap) a.back[j] = pi;
            } else {
                a.out_n[pi] = -2;
                atomicAdd(&a.ctr->bad_walk, 1u);
            }
        }
        wave_sync();
    }
    if (sl == 0) {
        atomicAdd(&a.ctr->cells, cells);
        atomicAdd(&a.ctr->dps, dps);
    }
}

}  // namespace

hipError_t launch_traceback(const TbArgs &a, int blocks, int lanes, bool wide, hipStream_t st) {
    if (blocks <= 0) return hipSuccess;
    if (wide)
        hipLaunchKernelGGL((k_traceback<64, true>), dim3(blocks), dim3(64), 0, st, a);
    else if (lanes == 64)
        hipLaunchKernelGGL((k_traceback<64, false>), dim3(blocks), dim3(64), 0, st, a);
    else
        hipLaunchKernelGGL((k_traceback<16, false>), dim3(blocks), dim3(64), 0, st, a);
    return hipGetLastError();
}

int traceback_blocks_per_cu(int lanes, bool wide) {
    int n = 0;
    hipError_t e;
    if (wide)
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_traceback<64, true>, 64, 0);
    else if (lanes == 64)
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_traceback<64, false>, 64, 0);
    else
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_traceback<16, false>, 64, 0);
    return e == hipSuccess ? n : 0;
}

}  // namespace blz
