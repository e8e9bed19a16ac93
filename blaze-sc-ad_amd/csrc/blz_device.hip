// blz_device.hip -- CDNA4 (gfx950) kernels of the BLASTP prelim stage.
//
// Stage 1, k_seed: one 64-lane wavefront per subject at a time (subjects taken from a
//   device counter, so the launch is persistent and work-stealing).  A wave reads 64
//   consecutive subject words per step (coalesced), looks up their query hits, and runs the
//   two-hit state machine of NCBI's word finder on a per-wave diagonal ring in LDS.  Hits of
//   one step that fall on the same diagonal are applied in subject order (an LDS atomicMin
//   on the hit's rank elects the earliest).  Pairs are extended ungapped (X-drop) in the lane
//   that owns the hit; segments at or above the ungapped cutoff are staged and then written
//   to a pool with one WorkItem per subject.
//   Reference counterpart: gpuBLAZE_full, BLAZE_gpu/gpuBLAZE.cu:885-1252.
// Stage 2, k_gapped: one lane per subject of the work list runs the gapped stage of NCBI's
//   BLAST_GetGappedScore: init hits in score order, containment test against the HSPs
//   saved so far, gapped start, score-only semi-gapped X-drop DP left and right, then the
//   common-endpoint purge and the score sort.
//   Reference counterpart: gpuBLAZE_everything, BLAZE_gpu/gpuBLAZE.cu:1254-1397.
#include <hip/hip_runtime.h>

#include "blz_common.h"
#include "blz_device.h"
#include "blz_dev_common.h"

namespace blz {

// Diagonal ring entry: 0 = no hit yet; otherwise bit 31 = "extended" flag and the low bits
// hold (wave-relative position + 1) of the last hit, or of the end barrier when flagged.
constexpr uint32_t kFlag = 0x80000000u;

__global__ void __launch_bounds__(256) k_seed(SeedArgs a) {
    extern __shared__ uint32_t lds[];
    const int lane = lane_id();
    const int wave = threadIdx.x >> 6;
    const int nwaves = blockDim.x >> 6;
    const uint32_t M = a.ring_mask + 1;

    // per block: the matrix; per wave: ring[M], claim[M], excl[65], p0[64], counter
    int16_t(*mat)[28] = reinterpret_cast<int16_t(*)[28]>(lds);
    const uint32_t mat_words = (28 * 28 * 2 + 3) / 4;
    uint32_t *wbase = lds + mat_words + (size_t)wave * (2 * M + 64 + 65 + 1);
    uint32_t *ring = wbase;
    uint32_t *claim = ring + M;
    uint32_t *excl = claim + M;
    uint32_t *p0s = excl + 65;
    uint32_t *nih_lds = p0s + 64;

    for (int i = threadIdx.x; i < 28 * 28; i += blockDim.x)
        mat[i / 28][i % 28] = kBlosum62Dev[i / 28][i % 28];
    for (uint32_t i = lane; i < M; i += 64) {
        ring[i] = 0;
        claim[i] = 0xFFFFFFFFu;
    }
    if (lane == 0) *nih_lds = 0;
    __syncthreads();

    const uint64_t gwave = (uint64_t)blockIdx.x * nwaves + wave;
    InitHit *stage = a.stage + gwave * (uint64_t)a.stage_cap;
    const uint8_t *Q = a.query;
    const int qlen = a.qlen;
    uint32_t abs_base = 0;
    unsigned long long n_hits = 0, n_ext = 0;

    for (;;) {
        int64_t sid = 0;
        if (lane == 0) sid = (int64_t)atomicAdd(&a.ctr->next_subject, 1u) + a.s_begin;
        sid = __shfl(sid, 0, 64);
        if (sid >= a.s_end) break;
        const uint64_t o0 = a.offs[sid];
        const int slen = (int)(a.offs[sid + 1] - o0);
        const uint8_t *S = a.res + o0;
        const uint32_t abs0 = abs_base;
        abs_base += (uint32_t)slen + 128u;

        for (int c0 = 0; c0 + kWord <= slen; c0 += 64) {
            const int s = c0 + lane;
            const bool valid = s + kWord <= slen;
            uint32_t pa = 0, nh = 0;
            if (valid) {
                uint32_t w = S[s] * 784u + S[s + 1] * 28u + S[s + 2];
                pa = a.prefix[w];
                nh = a.prefix[w + 1] - pa;
            }
            const uint32_t inc = wave_inclusive_sum(nh);
            const uint32_t H = __shfl(inc, 63, 64);
            excl[lane] = inc - nh;
            p0s[lane] = pa;
            if (lane == 0) excl[64] = H;
            wave_sync();
            n_hits += nh;
            for (uint32_t r = 0; r < H; r += 64) {
                const uint32_t g = r + lane;
                bool pend = g < H;
                int src = 0, qo = 0;
                if (pend) {
                    int lo = 0, hi = 63;
                    while (lo < hi) {
                        int mid = (lo + hi + 1) >> 1;
                        if (excl[mid] <= g) lo = mid;
                        else hi = mid - 1;
                    }
                    src = lo;
                    qo = a.entries[p0s[src] + (g - excl[src])];
                }
                const int sp = c0 + src;
                const uint32_t slot = (uint32_t)(qo - sp) & a.ring_mask;
                while (__ballot(pend)) {
                    if (pend) atomicMin(&claim[slot], g);
                    wave_sync();
                    const bool win = pend && claim[slot] == g;
                    wave_sync();
                    if (win) {
                        claim[slot] = 0xFFFFFFFFu;
                        pend = false;
                        const uint32_t st = ring[slot];
                        const uint32_t sabs = abs0 + (uint32_t)sp;
                        if (st & kFlag) {
                            if (sabs >= (st & ~kFlag) - 1u) ring[slot] = sabs + 1u;
                        } else {
                            const uint32_t diff = st ? sabs - (st - 1u) : 0x7FFFFFFFu;
                            if (diff >= (uint32_t)a.window) {
                                ring[slot] = sabs + 1u;
                            } else if (diff >= (uint32_t)kWord) {
                                const int last_local = (int)(st - 1u - abs0);
                                int hq, hs, hl, rext, slast;
                                int score = two_hit(mat, Q, qlen, S, slen, last_local + kWord, sp,
                                                    qo, a.x_ungapped, &hq, &hs, &hl, &rext, &slast);
                                n_ext++;
                                if (score >= a.ungapped_cutoff) {
                                    uint32_t k = atomicAdd(nih_lds, 1u);
                                    if (k < (uint32_t)a.stage_cap) stage[k] = InitHit{hq, hs, hl, score};
                                }
                                ring[slot] = rext ? ((abs0 + (uint32_t)slast - 2u + 1u) | kFlag)
                                                  : (sabs + 1u);
                            }
                        }
                    }
                    wave_sync();
                }
            }
            wave_sync();
        }

        // hand the subject's init hits to the gapped stage
        wave_sync();
        const uint32_t nih = *nih_lds;
        wave_sync();
        if (nih > 0) {
            if (nih > (uint32_t)a.stage_cap) {
                if (lane == 0) {
                    uint32_t k = atomicAdd(&a.ctr->overflow_count, 1u);
                    if (k < a.ovf_cap) a.ovf[k] = sid;
                }
            } else {
                unsigned long long base = 0;
                if (lane == 0) base = atomicAdd(&a.ctr->pool_used, (unsigned long long)nih);
                base = __shfl(base, 0, 64);
                if (base + nih > a.pool_cap) {
                    if (lane == 0) {
                        uint32_t k = atomicAdd(&a.ctr->overflow_count, 1u);
                        if (k < a.ovf_cap) a.ovf[k] = sid;
                    }
                } else {
                    for (uint32_t k = lane; k < nih; k += 64) a.pool[base + k] = stage[k];
                    if (lane == 0) {
                        uint32_t wi = atomicAdd(&a.ctr->work_count, 1u);
                        a.work[wi] = WorkItem{sid, (uint64_t)base, (int32_t)nih, 0};
                        atomicAdd(&a.ctr->init_hits, (unsigned long long)nih);
                    }
                }
            }
            if (lane == 0) *nih_lds = 0;
            wave_sync();
        }
    }
    // lane-local counts -> device counters
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

// ---------------------------------------------------------------------------------------
// Score-only semi-gapped X-drop DP (NCBI Blast_SemiGappedAlign score-only path; the
// reference states the banded X-drop at gpuBLAZE.cu:356-677).  Lane-serial; the live band
// [first_b_index, b_size) lives in a ring of W (power of two) {best, best_gap} pairs.
// Sets *ovf when the band would not fit the ring.
// ---------------------------------------------------------------------------------------
__device__ int semi_gapped(const int16_t (*mat)[28], const uint8_t *A, const uint8_t *B, int M,
                           int N, int x_dropoff, int gap_open, int gap_extend, bool reverse,
                           int *a_off, int *b_off, int2 *ring, int wmask, bool *ovf,
                           unsigned long long *cells) {
    const int goe = gap_open + gap_extend;
    *a_off = 0;
    *b_off = 0;
    if (x_dropoff < goe) x_dropoff = goe;
    if (N <= 0 || M <= 0) return 0;
    const int W = wmask + 1;
    int score = -goe;
    ring[0] = make_int2(0, -goe);
    int i;
    for (i = 1; i <= N; i++) {
        if (score < -x_dropoff) break;
        if (i >= W) {
            *ovf = true;
            return 0;
        }
        ring[i & wmask] = make_int2(score, score - goe);
        score -= gap_extend;
    }
    int b_size = i;
    int best_score = 0;
    int first_b_index = 0;
    unsigned long long nc = 0;
    for (int a_index = 1; a_index <= M; a_index++) {
        const int16_t *mrow = mat[reverse ? A[M - a_index] : A[a_index]];
        int sc = kMinIntDp;
        int score_gap_row = kMinIntDp;
        int last_b_index = first_b_index;
        for (int b_index = first_b_index; b_index < b_size; b_index++) {
            int2 cell = ring[b_index & wmask];
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
                    ring[b_index & wmask].x = kMinIntDp;
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
                cell.y = open > score_gap_col ? open : score_gap_col;
                score_gap_row = open > score_gap_row ? open : score_gap_row;
                cell.x = sc;
                ring[b_index & wmask] = cell;
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
                ring[b_size & wmask] = make_int2(score_gap_row, score_gap_row - goe);
                score_gap_row -= gap_extend;
                b_size++;
            }
        }
        if (b_size <= N) {
            if (b_size - first_b_index >= W) {
                *ovf = true;
                return 0;
            }
            ring[b_size & wmask] = make_int2(kMinIntDp, kMinIntDp);
            b_size++;
        }
    }
    *cells += nc;
    return best_score;
}

__device__ __forceinline__ bool init_before(const InitHit &a, const InitHit &b) {
    if (a.score != b.score) return a.score > b.score;
    if (a.s_start != b.s_start) return a.s_start < b.s_start;
    if (a.len != b.len) return a.len > b.len;
    return a.q_start < b.q_start;
}

__device__ __forceinline__ int gapped_start(const int16_t (*mat)[28], const uint8_t *q,
                                            const uint8_t *s, int q_start, int s_start, int len) {
    if (len <= kHspMaxWindow) return q_start + len / 2;
    int score = 0;
    for (int i = 0; i < kHspMaxWindow; i++) score += mat[q[q_start + i]][s[s_start + i]];
    int max_score = score;
    int max_offset = q_start + kHspMaxWindow - 1;
    for (int i = q_start + kHspMaxWindow; i < q_start + len; i++) {
        const int k = i - q_start;
        score -= mat[q[i - kHspMaxWindow]][s[s_start + k - kHspMaxWindow]];
        score += mat[q[i]][s[s_start + k]];
        if (score > max_score) {
            max_score = score;
            max_offset = i;
        }
    }
    return max_score > 0 ? max_offset - kHspMaxWindow / 2 : q_start;
}

// orders of the purge and of the final list (same keys as the oracle's comparators)
__device__ __forceinline__ int cmp_start(const Hsp &a, const Hsp &b) {
    if (a.q_off != b.q_off) return a.q_off < b.q_off ? -1 : 1;
    if (a.s_off != b.s_off) return a.s_off < b.s_off ? -1 : 1;
    if (a.score != b.score) return a.score > b.score ? -1 : 1;
    if (a.q_end != b.q_end) return a.q_end > b.q_end ? -1 : 1;
    if (a.s_end != b.s_end) return a.s_end > b.s_end ? -1 : 1;
    if (a.q_gs != b.q_gs) return a.q_gs < b.q_gs ? -1 : 1;
    if (a.s_gs != b.s_gs) return a.s_gs < b.s_gs ? -1 : 1;
    return 0;
}
__device__ __forceinline__ int cmp_end(const Hsp &a, const Hsp &b) {
    if (a.q_end != b.q_end) return a.q_end < b.q_end ? -1 : 1;
    if (a.s_end != b.s_end) return a.s_end < b.s_end ? -1 : 1;
    if (a.score != b.score) return a.score > b.score ? -1 : 1;
    if (a.q_off != b.q_off) return a.q_off > b.q_off ? -1 : 1;
    if (a.s_off != b.s_off) return a.s_off > b.s_off ? -1 : 1;
    if (a.q_gs != b.q_gs) return a.q_gs < b.q_gs ? -1 : 1;
    if (a.s_gs != b.s_gs) return a.s_gs < b.s_gs ? -1 : 1;
    return 0;
}
__device__ __forceinline__ int cmp_score(const Hsp &a, const Hsp &b) {
    if (a.score != b.score) return a.score > b.score ? -1 : 1;
    if (a.s_off != b.s_off) return a.s_off < b.s_off ? -1 : 1;
    if (a.s_end != b.s_end) return a.s_end > b.s_end ? -1 : 1;
    if (a.q_off != b.q_off) return a.q_off < b.q_off ? -1 : 1;
    if (a.q_end != b.q_end) return a.q_end > b.q_end ? -1 : 1;
    if (a.q_gs != b.q_gs) return a.q_gs < b.q_gs ? -1 : 1;
    if (a.s_gs != b.s_gs) return a.s_gs < b.s_gs ? -1 : 1;
    return 0;
}

template <int (*CMP)(const Hsp &, const Hsp &)>
__device__ void insertion_sort(Hsp *h, int n) {
    for (int i = 1; i < n; i++) {
        Hsp x = h[i];
        int j = i - 1;
        while (j >= 0 && CMP(x, h[j]) < 0) {
            h[j + 1] = h[j];
            j--;
        }
        h[j + 1] = x;
    }
}

__global__ void __launch_bounds__(1024) k_gapped(GappedArgs a) {
    __shared__ int16_t mat[28][28];
    for (int i = threadIdx.x; i < 28 * 28; i += blockDim.x) mat[i / 28][i % 28] = kBlosum62Dev[i / 28][i % 28];
    __syncthreads();
    const uint64_t tid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t nthreads = (uint64_t)gridDim.x * blockDim.x;
    int2 *ring = a.dp_scratch + tid * (uint64_t)(a.dp_wmask + 1);
    Hsp *saved = a.hsp_scratch + tid * (uint64_t)a.saved_cap;
    const uint32_t nwork = a.ctr->work_count;
    unsigned long long cells = 0, gext = 0;
    for (uint64_t item = tid; item < nwork; item += nthreads) {
        const WorkItem wi = a.work[item];
        InitHit *ih = a.pool + wi.off;
        const int n = wi.n;
        for (int i = 1; i < n; i++) {  // score order of the init hits
            InitHit x = ih[i];
            int j = i - 1;
            while (j >= 0 && init_before(x, ih[j])) {
                ih[j + 1] = ih[j];
                j--;
            }
            ih[j + 1] = x;
        }
        const uint64_t o0 = a.offs[wi.sid];
        const int slen = (int)(a.offs[wi.sid + 1] - o0);
        const uint8_t *S = a.res + o0;
        const QueryDev &qd = a.qd[wi.qi];
        const uint8_t *Q = qd.query;
        const int qlen = qd.qlen;
        int ns = 0;
        bool ovf = false;
        for (int k = 0; k < n && !ovf; k++) {
            const InitHit h = ih[k];
            const int q0 = h.q_start, q1 = h.q_start + h.len - 1;
            const int s0 = h.s_start, s1 = h.s_start + h.len - 1;
            bool skip = false;
            for (int t = 0; t < ns && !skip; t++) {
                const Hsp &u = saved[t];
                skip = h.score <= u.score && u.q_off <= q0 && q0 <= u.q_end && u.s_off <= s0 &&
                       s0 <= u.s_end && u.q_off <= q1 && q1 <= u.q_end && u.s_off <= s1 &&
                       s1 <= u.s_end;
            }
            if (skip) continue;
            const int q_gs = gapped_start(mat, Q, S, h.q_start, h.s_start, h.len);
            const int s_gs = h.s_start + (q_gs - h.q_start);
            int la, lb, ra, rb;
            const int left = semi_gapped(mat, Q, S, q_gs + 1, s_gs + 1, qd.x_gapped, qd.gap_open,
                                         qd.gap_extend, true, &la, &lb, ring, a.dp_wmask, &ovf, &cells);
            const int right = semi_gapped(mat, Q + q_gs, S + s_gs, qlen - q_gs - 1, slen - s_gs - 1,
                                          qd.x_gapped, qd.gap_open, qd.gap_extend, false, &ra, &rb,
                                          ring, a.dp_wmask, &ovf, &cells);
            gext++;
            const int score = left + right;
            if (!ovf && score >= qd.gapped_cutoff) {
                if (ns >= a.saved_cap) {
                    ovf = true;
                    break;
                }
                saved[ns++] = Hsp{score, q_gs + 1 - la, q_gs + 1 + ra, s_gs + 1 - lb, s_gs + 1 + rb, q_gs, s_gs};
            }
        }
        if (ovf) {
            uint32_t k = atomicAdd(&a.ctr->overflow_count, 1u);
            if (k < a.ovf_cap) a.ovf[k] = ((int64_t)wi.qi << kOvfShift) | wi.sid;
            continue;
        }
        if (ns == 0) continue;
        // purge HSPs sharing a start, then HSPs sharing an end; then the score order
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
        const unsigned long long base = atomicAdd(&a.ctr->hsp_used, (unsigned long long)ns);
        if (base + ns > a.hsp_cap) {
            uint32_t k = atomicAdd(&a.ctr->overflow_count, 1u);
            if (k < a.ovf_cap) a.ovf[k] = ((int64_t)wi.qi << kOvfShift) | wi.sid;
            continue;
        }
        for (int i = 0; i < ns; i++) a.hsps[base + i] = saved[i];
        const uint32_t ri = atomicAdd(&a.ctr->result_count, 1u);
        a.results[ri] = ResultItem{wi.sid, (uint64_t)base, ns, wi.qi};
    }
    atomicAdd(&a.ctr->dp_cells, cells);
    atomicAdd(&a.ctr->gapped_ext, gext);
}

// ---------------------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------------------
size_t seed_lds_bytes(uint32_t ring_words, int waves_per_block) {
    const size_t mat_words = (28 * 28 * 2 + 3) / 4;
    return 4 * (mat_words + (size_t)waves_per_block * (2 * ring_words + 64 + 65 + 1));
}

hipError_t launch_seed(const SeedArgs &a, int blocks, int waves_per_block, hipStream_t st) {
    const size_t lds = seed_lds_bytes(a.ring_mask + 1, waves_per_block);
    hipLaunchKernelGGL(k_seed, dim3(blocks), dim3(64 * waves_per_block), lds, st, a);
    return hipGetLastError();
}

hipError_t launch_gapped(const GappedArgs &a, int blocks, int threads, hipStream_t st) {
    hipLaunchKernelGGL(k_gapped, dim3(blocks), dim3(threads), 0, st, a);
    return hipGetLastError();
}

}  // namespace blz
