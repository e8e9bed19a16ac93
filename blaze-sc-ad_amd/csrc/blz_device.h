// blz_device.h -- kernel argument blocks and launchers (host side of blz_device.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "blz_common.h"

namespace blz {

struct SeedArgs {
    const uint8_t *res;        // concatenated subject residues (device)
    const uint64_t *offs;      // subject offsets [nsubj+1] (device)
    int64_t s_begin, s_end;    // subject range of this launch
    const uint8_t *query;      // query residues (device)
    int32_t qlen;
    const uint32_t *prefix;    // lookup prefix [kWords+1]
    const uint16_t *entries;   // query offsets per word, ascending
    int32_t window, x_ungapped, ungapped_cutoff;
    uint32_t ring_mask;        // diagonal ring size - 1 (power of two >= qlen + 64 + window)
    InitHit *stage;            // per-wave staging, stage_cap records per wave
    int32_t stage_cap;
    InitHit *pool;             // init hits of all subjects
    unsigned long long pool_cap;
    WorkItem *work;            // one item per subject with init hits
    int64_t *ovf;              // subjects past a cap
    uint32_t ovf_cap;
    DevCounters *ctr;
};

struct GappedArgs {
    const uint8_t *res;
    const uint64_t *offs;
    const uint8_t *query;
    int32_t qlen;
    int32_t x_gapped, gap_open, gap_extend, gapped_cutoff;
    InitHit *pool;
    const WorkItem *work;
    int2 *dp_scratch;          // per-lane DP ring, dp_wmask+1 entries per lane
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
hipError_t launch_seed(const SeedArgs &a, int blocks, int waves_per_block, hipStream_t st);
hipError_t launch_gapped(const GappedArgs &a, int blocks, hipStream_t st);

}  // namespace blz
