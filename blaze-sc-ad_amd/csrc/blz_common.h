// blz_common.h -- constants and record layouts shared by the host runtime and the kernels.
#pragma once
#include <stdint.h>

#ifndef __host__
#define __host__
#define __device__
#endif

namespace blz {

constexpr int kAlpha = 28;           // NCBIstdaa letters
constexpr int kWords = 28 * 28 * 28; // word index a*784 + b*28 + c (gpuBLAZE.cu:991)
constexpr int kWord = 3;             // word size
constexpr int kRelBlocks = kWords / 64 + 1;  // 64-word blocks of the compact lookup prefix
// Longest subject a database may hold: the seed kernels' 32-bit diagonal rings keep 30-bit positions, and
// an init hit keeps its length in 24 bits.
constexpr uint64_t kMaxSubjectLen = (1u << 24) - 1;
// Longest subject the first seed kernel (seed variant 1) takes: it keeps 16-bit positions.
constexpr uint64_t kSeed1MaxSubjectLen = 65535;
constexpr int kHspMaxWindow = 11;    // gapped-start window (gpuBLAZE.cu:64)
constexpr int kMinIntDp = -1073741824; // INT_MIN/2, the DP's "minus infinity"

// BLOSUM62 in NCBIstdaa order (- A B C D E F G H I K L M N P Q R S T V W X Y Z U * O J),
// the reference's table (BLAZE_gpu/util.hpp:30-57); letter 0 (gap) scores -32768.
extern const int16_t kBlosum62[28][28];

// Ungapped HSP that reached the ungapped cutoff ("init hit"): the segment and its score.
struct InitHit {
    int32_t q_start, s_start, len, score;
};

// Subject that reached the gapped stage: its init hits are pool[off, off+n).
struct WorkItem {
    int64_t sid;
    uint64_t off;
    int32_t n;
    int32_t qi;      // query of the batch
    int32_t bucket;  // expected gapped-stage cost class, 0 = heaviest (from the best init hit)
    int32_t pad;
};

constexpr int kBuckets = 8;
// Expected gapped-stage cost class of a subject, 0 = heaviest, from its init hit count, best
// ungapped score and length.  Each init hit outside the HSPs found before it is a DP of its own,
// and a chance hit's DP runs about as long as a homolog's (X-drop 38 in a random walk), so the
// count comes first: 5 or more, 4, 3, 2.  Single hits: a score of 150 or more first (a long
// homologous alignment), then by subject length (400, 200 letters), which bounds the DP's
// length: a low-scoring hit on a long homolog runs as long as the alignment, and ordered by
// score it was taken among the last.
__host__ __device__ inline int cost_bucket(int best_ungapped, int n_init, int slen) {
    if (n_init >= 5) return 0;
    if (n_init >= 2) return 5 - n_init;
    if (best_ungapped >= 150) return 4;
    return slen >= 400 ? 5 : slen >= 200 ? 6 : 7;
}

// HSP record, same layout as blz_hsp in blz_prelim.h.
struct Hsp {
    int32_t score, q_off, q_end, s_off, s_end, q_gs, s_gs;
};

// Subject's HSP list on the device: hsps[off, off+n).
struct ResultItem {
    int64_t sid;
    uint64_t off;
    int32_t n;
    int32_t qi;
};

// Overflow record: query index in bits 40..61, subject id in the low 40 bits; bit 62 marks a
// pair the gapped stage gave up on (its init hits were already counted by the seed stage).
constexpr int kOvfShift = 40;
constexpr unsigned long long kOvfGapped = 1ull << 62;
constexpr long long kOvfQiMask = (1ll << 22) - 1;
constexpr long long kOvfSidMask = (1ll << kOvfShift) - 1;

// Device counters of one run (atomically accumulated by the kernels).
struct DevCounters {
    unsigned long long word_hits, ungapped_ext, init_hits, gapped_ext, dp_cells;
    unsigned long long pool_used, hsp_used;
    unsigned int work_count, result_count, overflow_count, next_subject, next_work, next_item;
    unsigned int bucket_count[kBuckets], bucket_cursor[kBuckets];
    unsigned int redo_count, next_redo, next_heavy, retry_count;
    unsigned int next_retry;
    unsigned int ovf_lost;  // overflow records that did not fit the overflow list
};

}  // namespace blz
