// blz_device.h -- kernel argument blocks and launchers (host side of blz_device.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "blz_common.h"

namespace blz {

// One query of a batch as the kernels see it.
struct QueryDev {
    const uint8_t *query;
    const uint32_t *prefix;    // lookup prefix [kWords+1]
    const uint16_t *entries;   // query offsets per word, ascending
    const uint16_t *rel16;     // compact prefix: offset of word w inside its 64-word block
    const uint32_t *blockbase; // compact prefix: first entry of each 64-word block
    // word records, one 16-byte load per word: x = entry count (255: at least 255, or first
    // entry past 2^24; read the prefix) | first entry << 8; y, z, w = the first kWordRecInline
    // entries (u16 pairs, low half first)
    const uint4 *rec;
    uint32_t n_entries;
    int32_t qlen;
    uint32_t ring_mask, claim_mask;
    // seed kernel 2: diagonal (q - s) sits in ring slot (q + ring_base - s) mod ring_len, with
    // ring_base a multiple of ring_len above every subject offset; ring_magic = ceil(2^32 /
    // ring_len) makes the modulo a multiply-high (exact below 2^32 / ring_len)
    uint32_t ring_len, ring_base, ring_magic;
    int32_t window, x_ungapped, ungapped_cutoff;
    int32_t x_gapped, gap_open, gap_extend, gapped_cutoff;
    int32_t reap_score;        // HSPs scoring below it are dropped after the purge and sort (0: none)
    // a pack of queries scanned together (seed stage): query offsets of the concatenation
    // belong to member m when member_start[m] <= q < member_start[m+1]; member_idx[m] is the
    // member's index in the run's query table.  A single query is a pack of one.
    const uint32_t *member_start;
    const uint32_t *member_idx;
    int32_t n_members;
    // subject groups (seed stage): virtual distance between the subjects of a group, 0 when a
    // wave scans one subject at a time
    int32_t group_gap;
};
constexpr int kMaxMembers = 32;
constexpr int kWordRecInline = 6;

// Longest subject the seed kernel's LDS diagonal ring (16-bit entries) takes; longer subjects
// use SeedArgs::long_ring.
constexpr int kSeedNarrowMax = 0x3FFF;

struct SeedArgs {
    const uint8_t *res;        // concatenated subject residues (device)
    const uint64_t *offs;      // subject offsets [nsubj+1] (device)
    int64_t s_begin, s_end;    // subject range of this launch
    const uint8_t *query;      // query residues (device)
    int32_t qlen;
    const uint32_t *prefix;    // lookup prefix [kWords+1]
    const uint16_t *entries;   // query offsets per word, ascending
    int32_t window, x_ungapped, ungapped_cutoff;
    uint32_t ring_mask;        // seed kernel 1: diagonal ring size - 1 (power of two >= qlen + 64 + window)
    uint32_t ring_len;         // seed kernel 2: largest diagonal ring of the launch's packs
    uint32_t claim_mask;       // same-diagonal arbitration table size - 1 (power of two)
    const uint16_t *rel16;     // compact prefix: offset of word w inside its 64-word block
    const uint32_t *blockbase; // compact prefix: first entry of each 64-word block
    uint32_t n_entries;        // lookup entries
    InitHit *stage;            // per-wave staging, stage_cap records per wave
    int32_t stage_cap;
    InitHit *pool;             // init hits of all subjects
    unsigned long long pool_cap;
    WorkItem *work;            // one item per subject with init hits
    uint32_t work_cap;
    int64_t *ovf;              // subjects past a cap
    uint32_t ovf_cap;
    DevCounters *ctr;
    // batched form (seed kernel 2): items are (query, block of subjects), query-major
    const QueryDev *qd;
    int32_t nq;                // packs of this launch
    uint32_t *item_ctr;        // this launch's work-item counter
    uint32_t *long_ring;       // long-subject launch: diagonal rings, ring_len words per wave
    const int64_t *long_ids;   // long-subject launch: the subjects longer than kSeedNarrowMax
    const int64_t *items;      // work items, three values each: pack, first subject, end subject
    uint32_t n_items;
    // retry launch: the (query << kOvfShift | subject) pairs the main launches could not stage,
    // rescanned with qd = the run's per-query table and stage_cap records per wave
    const int64_t *retry;
};

struct GappedArgs {
    const uint8_t *res;
    const uint64_t *offs;
    const QueryDev *qd;        // queries of the batch, indexed by WorkItem::qi
    InitHit *pool;
    const WorkItem *work;
    const uint32_t *order;     // work indices to process (cost order, or the redo list)
    const uint32_t *count;     // how many of them (device counter) ...
    uint32_t count_cap;        // ... at most this many
    uint32_t *next;            // work-stealing cursor over order[] (lane-serial passes)
    uint32_t *next_heavy;      // work-stealing cursor of the wave pass
    int32_t heavy_classes;     // cost classes [0, heavy_classes) go to the wave pass
    uint32_t *redo;            // work indices whose DP band outgrew the LDS ring
    uint32_t redo_cap;
    int2 *dp_scratch;          // wide variant: per-lane DP ring, dp_wmask+1 entries per lane
    int32_t dp_wmask;
    Hsp *hsp_scratch;          // per-lane saved HSPs, saved_cap per lane
    int32_t saved_cap;
    Hsp *hsps;                 // output HSPs
    unsigned long long hsp_cap;
    ResultItem *results;       // one per subject with HSPs
    int64_t *ovf;
    uint32_t ovf_cap;
    DevCounters *ctr;
};

size_t seed_lds_bytes(uint32_t ring_words, int waves_per_block);
size_t seed2_lds_bytes(const SeedArgs &a, int table, int waves);
hipError_t launch_seed2(const SeedArgs &a, int table, int blocks, int waves, bool long_subjects, hipStream_t st);
// Second chance for the subjects the seed launches recorded as past a cap: moves the overflow
// list to `retry` (and clears it), then rescans those pairs with a larger stage.
hipError_t launch_seed_retry(const SeedArgs &a, int64_t *retry, int blocks, int waves, hipStream_t st);
size_t seed_retry_lds_bytes(const SeedArgs &a, int waves);
// The retry scan alone over a.retry[0, ctr->retry_count) (the overflow queue's pass; the caller
// sets retry_count and clears the rest of a.ctr).
hipError_t launch_seed_pairs(const SeedArgs &a, int blocks, int waves, hipStream_t st);
// Resident workgroups per CU of that launch shape (registers and LDS), 0 if it cannot run.
int seed2_blocks_per_cu(const SeedArgs &a, int table, int waves, bool long_subjects);
hipError_t launch_seed(const SeedArgs &a, int blocks, int waves_per_block, hipStream_t st);
hipError_t launch_fill_u32(uint32_t *p, size_t n, uint32_t v, hipStream_t st);
// cost-ordered work list: order[] receives the work indices, heaviest class first
hipError_t launch_order(const WorkItem *work, DevCounters *ctr, uint32_t *order, uint32_t work_cap, int blocks,
                        hipStream_t st);
// gapped stage with the DP band in LDS (16-bit scores relative to the row's best); items whose
// band outgrows the ring go to a.redo
hipError_t launch_gapped_lds(const GappedArgs &a, int blocks, hipStream_t st);
// gapped stage with the DP band in global memory (wide rings), for the redo list
hipError_t launch_gapped_wide(const GappedArgs &a, int blocks, int threads, hipStream_t st);
constexpr int kLdsRing = 128;   // LDS DP ring entries per lane (lane-serial pass)
constexpr int kWaveRing = 512;  // LDS DP ring entries per wave (wave pass); wider bands are redone
constexpr int kSegRing = 128;   // LDS DP ring entries per 16-lane segment (segment pass); measured bands stay below 112
// gapped stage with one wave per subject of the heavy cost classes
hipError_t launch_gapped_wave(const GappedArgs &a, int blocks, hipStream_t st);
// Resident k_gapped_wave workgroups (one wave each) per CU.
int gapped_wave_blocks_per_cu();
// Four subjects per wave (16-lane segments); saved-HSP scratch of 4 x saved_cap per workgroup.
// packed: the 16-bit DP (two columns per lane, pairs of int16); else 32-bit, one column per lane
hipError_t launch_gapped_seg(const GappedArgs &a, int blocks, bool packed, hipStream_t st);
int gapped_seg_blocks_per_cu(bool packed);

// ---- traceback stage ----------------------------------------------------------------------
// Traced HSP, same layout as blz_tb_hsp in blz_prelim.h: the HSP fields, then identities,
// alignment columns, gap letters, gap openings and the edit script runs[op_first, +op_count).
struct TbHsp {
    Hsp h;
    int32_t num_ident, align_len, gaps, gap_opens, op_first, op_count;
};
// One (query, subject) pair of a traceback call: prelim HSPs pre[off, off + n) in score order;
// its traced HSPs go to out[off, off + n) (out_n[pair] of them, -1 if the pass gave it back).
struct TbPair {
    int64_t sid;
    uint64_t off;
    int32_t qi, n;
};
struct TbQuery {
    const uint8_t *q;
    int32_t qlen;
    int32_t cutoff;  // traced HSPs scoring below it are dropped
};
struct TbCounters {
    unsigned long long runs_used;  // edit-script runs reserved (may pass runs_cap: then the call runs again)
    unsigned long long cells;      // DP cells inside the X-drop bands
    unsigned long long dps;        // one-direction alignments
    unsigned int next_pair;        // work-stealing cursor
    unsigned int back_count;       // pairs handed to the wide pass (back[])
    unsigned int bad_walk;         // paths that left the stored rows (never expected)
    unsigned int pad;
};
struct TbArgs {
    const uint8_t *res;
    const uint64_t *offs;
    const TbQuery *qs;
    const TbPair *pairs;
    uint32_t n_pairs;
    const uint32_t *list;      // wide pass: pair indices back[0, n_list); main pass: null
    uint32_t n_list;
    const Hsp *pre;
    TbHsp *out;
    int32_t *out_n;
    uint32_t *runs;
    unsigned long long runs_cap;
    uint32_t *back;            // main pass: pairs whose DP outgrew a per-wave buffer
    uint32_t back_cap;
    TbCounters *ctr;
    // per-wave storage (wave w at w * cap)
    uint8_t *tb;               // traceback bytes of one DP
    unsigned long long tb_cap;
    int4 *rows;                // per row: first column, byte offset, length
    int32_t rows_cap;
    uint32_t *wruns;           // runs of the current HSP before they are placed
    int32_t wruns_cap;
    int2 *ring;                // wide pass: DP band ring, ring_mask + 1 entries per wave
    int32_t ring_mask;
    int32_t x_final, gap_open, gap_extend;
};
// main pass: DP band ring in LDS per pair (a 64-lane wave, or a 16-lane segment of four per
// wave); wider bands go to the wide pass
constexpr int kTbWaveRing = 256, kTbSegRing = 128;
// lanes: 16 (four pairs per wave) or 64 (one pair per wave) in the main pass; the wide pass
// runs one pair per wave.  Each workgroup is one wave; per-pair storage areas are indexed by
// workgroup x pairs per wave + segment.
hipError_t launch_traceback(const TbArgs &a, int blocks, int lanes, bool wide, hipStream_t st);
int traceback_blocks_per_cu(int lanes, bool wide);

}  // namespace blz
