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
// (The gapped stage is in blz_gapped.hip; the batched seed kernel in blz_seed.hip.)
#include <hip/hip_runtime.h>

#include <algorithm>

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
                    int best = 0;
                    for (uint32_t k = lane; k < nih; k += 64) {
                            const InitHit h = stage[k];
                            a.pool[base + k] = h;
                            best = h.score > best ? h.score : best;
                        }
                        #pragma unroll
                        for (int o = 32; o > 0; o >>= 1) {
                            const int t = __shfl_xor(best, o, 64);
                            best = t > best ? t : best;
                        }
                    if (lane == 0) {
                        const uint32_t wi = atomicAdd(&a.ctr->work_count, 1u);
                        if (wi < a.work_cap) {
                            const int bucket = cost_bucket(best, (int)nih, slen);
                            a.work[wi] = WorkItem{sid, (uint64_t)base, (int32_t)nih, 0, bucket, 0};
                            atomicAdd(&a.ctr->bucket_count[bucket], 1u);
                            atomicAdd(&a.ctr->init_hits, (unsigned long long)nih);
                        } else {
                            const uint32_t k = atomicAdd(&a.ctr->overflow_count, 1u);
                            if (k < a.ovf_cap) a.ovf[k] = sid;
                        }
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

// Clears counters with a kernel rather than hipMemsetAsync, so the same sequence can be captured
// into a hipGraph.
__global__ void __launch_bounds__(256) k_fill_u32(uint32_t *p, uint32_t n, uint32_t v) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = v;
}

hipError_t launch_fill_u32(uint32_t *p, size_t n, uint32_t v, hipStream_t st) {
    if (n == 0) return hipSuccess;
    const int blocks = (int)std::min<size_t>((n + 255) / 256, 64);
    hipLaunchKernelGGL(k_fill_u32, dim3(blocks), dim3(256), 0, st, p, (uint32_t)n, v);
    return hipGetLastError();
}


}  // namespace blz
