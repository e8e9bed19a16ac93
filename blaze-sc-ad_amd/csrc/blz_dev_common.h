// blz_dev_common.h -- device helpers shared by the kernels: wave utilities and the ungapped
// X-drop walks.  Included by each .hip translation unit.
#pragma once
#include <hip/hip_runtime.h>

#include "blz_common.h"

namespace blz {

static __constant__ int16_t kBlosum62Dev[28][28] = {
#include "blz_blosum62.inc"
};

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ int lane_id() { return (int)(threadIdx.x & 63); }

// Inclusive prefix sum over the 64 lanes with DPP: row_shr 1/2/4/8 inside each 16-lane row,
// then row_bcast:15 (rows 1 and 3 add the last lane of rows 0 and 2) and row_bcast:31 (rows 2
// and 3 add lane 31).  Each step is one v_add_u32_dpp whose destination is its own second
// operand: a lane whose DPP source lies outside the row, or whose row the row mask leaves out,
// is not written and keeps its sum.  (The compiler does not fuse the builtin form into the add
// here, which costs two extra moves per step.)  Each DPP read follows an s_nop: a VALU write of
// the register needs two wait states before DPP reads it, a VALU write of EXEC five (the first
// one, whose predecessor is not known here).  All 64 lanes must be active.
__device__ __forceinline__ uint32_t wave_inclusive_sum(uint32_t v) {
    uint32_t x = v;
    asm volatile(
        "s_nop 4\n\t"
        "v_add_u32_dpp %0, %0, %0 row_shr:1 row_mask:0xf bank_mask:0xf\n\t"
        "s_nop 1\n\t"
        "v_add_u32_dpp %0, %0, %0 row_shr:2 row_mask:0xf bank_mask:0xf\n\t"
        "s_nop 1\n\t"
        "v_add_u32_dpp %0, %0, %0 row_shr:4 row_mask:0xf bank_mask:0xf\n\t"
        "s_nop 1\n\t"
        "v_add_u32_dpp %0, %0, %0 row_shr:8 row_mask:0xf bank_mask:0xf\n\t"
        "s_nop 1\n\t"
        "v_add_u32_dpp %0, %0, %0 row_bcast:15 row_mask:0xa bank_mask:0xf\n\t"
        "s_nop 1\n\t"
        "v_add_u32_dpp %0, %0, %0 row_bcast:31 row_mask:0xc bank_mask:0xf"
        : "+v"(x));
    return x;
}

// Inclusive prefix max over the 64 lanes with DPP (same lane pattern as wave_inclusive_sum);
// lanes whose source is outside the row read INT_MIN.
__device__ __forceinline__ int wave_inclusive_max(int x) {
    constexpr int kMin = -2147483647 - 1;
    x = max(x, __builtin_amdgcn_update_dpp(kMin, x, 0x111, 0xf, 0xf, false));
    x = max(x, __builtin_amdgcn_update_dpp(kMin, x, 0x112, 0xf, 0xf, false));
    x = max(x, __builtin_amdgcn_update_dpp(kMin, x, 0x114, 0xf, 0xf, false));
    x = max(x, __builtin_amdgcn_update_dpp(kMin, x, 0x118, 0xf, 0xf, false));
    x = max(x, __builtin_amdgcn_update_dpp(kMin, x, 0x142, 0xa, 0xf, false));
    x = max(x, __builtin_amdgcn_update_dpp(kMin, x, 0x143, 0xc, 0xf, false));
    return x;
}

// Each lane receives the value of the lane below it (DPP wave_shr:1); lane 0 receives `in`.
__device__ __forceinline__ int wave_shift_up1(int x, int in) {
    return __builtin_amdgcn_update_dpp(in, x, 0x138, 0xf, 0xf, false);
}

// value of lane 63 / lane 0, read into a scalar register
__device__ __forceinline__ uint32_t last_lane(uint32_t v) { return (uint32_t)__builtin_amdgcn_readlane((int)v, 63); }
__device__ __forceinline__ uint32_t first_lane(uint32_t v) { return (uint32_t)__builtin_amdgcn_readfirstlane((int)v); }
__device__ __forceinline__ uint64_t first_lane64(uint64_t v) {
    return ((uint64_t)first_lane((uint32_t)(v >> 32)) << 32) | first_lane((uint32_t)v);
}

// Order of a subject's init hits before the gapped stage (NCBI BLAST_GetGappedScore's sort:
// score descending, subject start, length descending, query start).
__device__ __forceinline__ bool init_before(const InitHit &a, const InitHit &b) {
    if (a.score != b.score) return a.score > b.score;
    if (a.s_start != b.s_start) return a.s_start < b.s_start;
    if (a.len != b.len) return a.len > b.len;
    return a.q_start < b.q_start;
}

// ---------------------------------------------------------------------------------------
// Ungapped X-drop walks (NCBI s_BlastAaExtendLeft/Right; the reference restates the stop
// rules at gpuBLAZE.cu:680-883).  Lane-serial; the matrix row comes from LDS.
// ---------------------------------------------------------------------------------------
__device__ __forceinline__ int extend_left(const int16_t (*mat)[28], const uint8_t *q,
                                           const uint8_t *s, int q_off, int s_off, int dropoff,
                                           int maxscore, int *length) {
    int n = s_off < q_off ? s_off : q_off;
    int best_i = n + 1;
    int score = maxscore;
    for (int i = n; i >= 0; i--) {
        score += mat[q[q_off - n + i]][s[s_off - n + i]];
        if (score > maxscore) {
            maxscore = score;
            best_i = i;
        }
        if (maxscore - score >= dropoff) break;
    }
    *length = n - best_i + 1;
    return maxscore;
}

__device__ __forceinline__ int extend_right(const int16_t (*mat)[28], const uint8_t *q, int qlen,
                                            const uint8_t *s, int slen, int q_off, int s_off,
                                            int dropoff, int maxscore, int *length,
                                            int *s_last_off) {
    int n = slen - s_off < qlen - q_off ? slen - s_off : qlen - q_off;
    int best_i = -1;
    int score = maxscore;
    int i;
    for (i = 0; i < n; i++) {
        score += mat[q[q_off + i]][s[s_off + i]];
        if (score > maxscore) {
            maxscore = score;
            best_i = i;
        }
        if (score <= 0 || maxscore - score >= dropoff) break;
    }
    *length = best_i + 1;
    *s_last_off = s_off + i;
    return maxscore;
}

// Two-hit extension from the second hit (q, s) back towards the first hit, whose word ends
// at s_left_off (NCBI s_BlastAaExtendTwoHit; reference gpuBLAZE.cu:1116-1154).
__device__ __forceinline__ int two_hit(const int16_t (*mat)[28], const uint8_t *q, int qlen,
                                       const uint8_t *s, int slen, int s_left_off, int s_right_off,
                                       int q_right_off, int dropoff, int *hsp_q, int *hsp_s,
                                       int *hsp_len, int *right_extend, int *s_last_off) {
    int left_d = 0, right_d = 0, left_score = 0, right_score = 0, score = 0;
#pragma unroll
    for (int i = 0; i < kWord; i++) {
        score += mat[q[q_right_off + i]][s[s_right_off + i]];
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
    left_score = extend_left(mat, q, s, q_right_off - 1, s_right_off - 1, dropoff, 0, &left_d);
    if (left_d >= s_right_off - s_left_off) {
        *right_extend = 1;
        right_score = extend_right(mat, q, qlen, s, slen, q_right_off, s_right_off, dropoff,
                                   left_score, &right_d, s_last_off);
    }
    *hsp_q = q_right_off - left_d;
    *hsp_s = s_right_off - left_d;
    *hsp_len = left_d + right_d;
    return left_score > right_score ? left_score : right_score;
}

// ---------------------------------------------------------------------------------------
// purge orders and the final order (NCBI s_QueryOffsetCompareHSPs, s_QueryEndCompareHSPs,
// ScoreCompareHSPs; the gapped start breaks remaining ties)
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

// records sorted by their HSP fields: Hsp itself, or a record holding one as member h
__device__ __forceinline__ const Hsp &hsp_key(const Hsp &h) { return h; }
template <class R>
__device__ __forceinline__ const Hsp &hsp_key(const R &r) { return r.h; }

template <int (*CMP)(const Hsp &, const Hsp &), class R>
__device__ void insertion_sort(R *h, int n) {
    for (int i = 1; i < n; i++) {
        R x = h[i];
        int j = i - 1;
        while (j >= 0 && CMP(hsp_key(x), hsp_key(h[j])) < 0) {
            h[j + 1] = h[j];
            j--;
        }
        h[j + 1] = x;
    }
}

// ---------------------------------------------------------------------------------------
// Cross-lane steps within a segment of kSeg lanes: the whole wave (64), or one DPP row (16), so
// that a wave runs four independent DPs side by side.  Values are per lane and uniform within a
// segment.
template <int kSeg>
struct Seg;
template <>
struct Seg<64> {
    static __device__ __forceinline__ int lane() { return lane_id(); }
    static __device__ __forceinline__ int incl_max(int x) { return wave_inclusive_max(x); }
    static __device__ __forceinline__ int shift_up1(int x, int in) { return wave_shift_up1(x, in); }
    static __device__ __forceinline__ int last(int x) { return (int)last_lane((uint32_t)x); }
    static __device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p); }
    static __device__ __forceinline__ int first(int x) { return (int)first_lane((uint32_t)x); }
};
template <>
struct Seg<16> {
    static __device__ __forceinline__ int lane() { return lane_id() & 15; }
    static __device__ __forceinline__ int incl_max(int x) {  // row_shr 1/2/4/8 inside the row
        constexpr int kMin = -2147483647 - 1;
        x = max(x, __builtin_amdgcn_update_dpp(kMin, x, 0x111, 0xf, 0xf, false));
        x = max(x, __builtin_amdgcn_update_dpp(kMin, x, 0x112, 0xf, 0xf, false));
        x = max(x, __builtin_amdgcn_update_dpp(kMin, x, 0x114, 0xf, 0xf, false));
        x = max(x, __builtin_amdgcn_update_dpp(kMin, x, 0x118, 0xf, 0xf, false));
        return x;
    }
    static __device__ __forceinline__ int shift_up1(int x, int in) {  // row_shr:1, the row's lane 0 gets in
        return __builtin_amdgcn_update_dpp(in, x, 0x111, 0xf, 0xf, false);
    }
    // DPP row_share:n (gfx90a and later): every lane of the row reads the row's lane n.  last()
    // folds into a following max (v_max_i32_dpp); bcast() writes every lane and needs no old
    // value (update_dpp's would cost a move per use).
    static __device__ __forceinline__ int last(int x) { return __builtin_amdgcn_update_dpp(0, x, 0x15F, 0xf, 0xf, false); }
    static __device__ __forceinline__ int bcast(int x) { return __builtin_amdgcn_mov_dpp(x, 0x15F, 0xf, 0xf, false); }
    static __device__ __forceinline__ uint64_t ballot(bool p) { return (__ballot(p) >> (lane_id() & 48)) & 0xFFFFull; }
    static __device__ __forceinline__ int first(int x) { return __builtin_amdgcn_update_dpp(0, x, 0x150, 0xf, 0xf, false); }
    static __device__ __forceinline__ int reduce_max(int x) { return bcast(incl_max(x)); }
};

// ---------------------------------------------------------------------------------------
// Pairs of signed 16-bit values in one 32-bit register (low half first), for the packed DP:
// v_pk_* instructions act on both halves; adds and subtracts saturate at the int16 range.
typedef short s16x2 __attribute__((ext_vector_type(2)));
constexpr uint32_t kNeg16x2 = 0x80008000u;  // (-32768, -32768): minus infinity in both halves
__device__ __forceinline__ s16x2 as_s16x2(uint32_t x) { return __builtin_bit_cast(s16x2, x); }
__device__ __forceinline__ uint32_t as_u32(s16x2 v) { return __builtin_bit_cast(uint32_t, v); }
__device__ __forceinline__ s16x2 pk_splat(int x) { return s16x2{(short)x, (short)x}; }
__device__ __forceinline__ s16x2 pk_max(s16x2 a, s16x2 b) { return __builtin_elementwise_max(a, b); }
__device__ __forceinline__ s16x2 pk_add(s16x2 a, s16x2 b) { return __builtin_elementwise_add_sat(a, b); }
__device__ __forceinline__ s16x2 pk_sub(s16x2 a, s16x2 b) { return __builtin_elementwise_sub_sat(a, b); }
// bits of a where mask is set, else bits of b
__device__ __forceinline__ uint32_t pk_select(uint32_t mask, uint32_t a, uint32_t b) { return (a & mask) | (b & ~mask); }
// Inclusive prefix max over a 16-lane DPP row of a lane's two columns (low, high).  The lane's
// own maximum goes to the high half, the row's prefix max of those runs on 32-bit values with
// fused v_max_i32_dpp steps (a lane without a source keeps its own value, which a max ignores),
// and the low half takes the maximum of the lanes before it.
__device__ __forceinline__ s16x2 pk_incl_max16(s16x2 v) {
    constexpr int kMin = -2147483647 - 1;
    const int lo = v.x;
    int x = max((int)v.y, lo);
    x = max(x, __builtin_amdgcn_update_dpp(kMin, x, 0x111, 0xf, 0xf, false));
    x = max(x, __builtin_amdgcn_update_dpp(kMin, x, 0x112, 0xf, 0xf, false));
    x = max(x, __builtin_amdgcn_update_dpp(kMin, x, 0x114, 0xf, 0xf, false));
    x = max(x, __builtin_amdgcn_update_dpp(kMin, x, 0x118, 0xf, 0xf, false));
    const int before = __builtin_amdgcn_update_dpp(kMin, x, 0x111, 0xf, 0xf, false);  // lanes before
    return s16x2{(short)max(lo, before), (short)x};
}

}  // namespace blz
