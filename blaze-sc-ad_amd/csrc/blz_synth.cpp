// blz_synth.cpp -- deterministic synthetic protein databases in NCBIstdaa.
//
// Residues are drawn i.i.d. from the Robinson-Robinson background; subject lengths are
// log-normal and clipped; a chosen fraction of subjects are planted homologs of the given
// queries (BLOSUM62 target-frequency substitutions to a drawn identity, geometric indels);
// optional low-complexity repeat blocks.  Every subject has its own random stream keyed by
// (seed, index), so filling is parallel and the result does not depend on the thread count.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <thread>
#include <vector>

#include "blz_common.h"
#include "blz_prelim.h"

namespace {

// NCBIstdaa codes of the 20 standard residues and their Robinson & Robinson (1991)
// background frequencies (per 1000).
const int kStd[20] = {1, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 17, 18, 19, 20, 22};
//                    A  C  D  E  F  G  H  I  K   L   M   N   P   Q   R   S   T   V   W   Y
const double kRobinson[20] = {78.05, 19.25, 53.64, 62.95, 38.56, 73.77, 21.99,
                              51.42, 57.44, 90.19, 22.43, 44.87, 52.03, 42.64,
                              51.29, 71.20, 58.41, 64.41, 13.30, 32.16};

inline uint64_t splitmix64(uint64_t &x) {
    uint64_t z = (x += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

struct Rng {  // xoshiro256**
    uint64_t s[4];
    explicit Rng(uint64_t seed) {
        uint64_t x = seed;
        for (auto &v : s) v = splitmix64(x);
    }
    static uint64_t rotl(uint64_t v, int k) { return (v << k) | (v >> (64 - k)); }
    uint64_t next() {
        uint64_t r = rotl(s[1] * 5, 7) * 9, t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return r;
    }
    double uniform() { return (next() >> 11) * (1.0 / 9007199254740992.0); }
    double normal() {  // Box-Muller
        double u1 = uniform(), u2 = uniform();
        if (u1 < 1e-300) u1 = 1e-300;
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    }
    int geometric(double mean) {  // support 1,2,...; mean >= 1
        double p = 1.0 / mean;
        int k = 1;
        while (uniform() > p && k < 1000) k++;
        return k;
    }
};

uint64_t stream_key(uint64_t seed, uint64_t i, uint64_t tag) {
    uint64_t x = seed ^ (i * 0xD1B54A32D192ED03ULL) ^ (tag * 0x8CB92BA72F3D8DD7ULL);
    return splitmix64(x);
}

struct Tables {
    uint8_t bg[65536];            // background letter by 16 random bits
    uint8_t sub[20][65536];       // substitute (different letter) by 16 random bits
    int idx_of[32];               // NCBIstdaa code -> 0..19, or -1
    Tables() {
        double tot = std::accumulate(kRobinson, kRobinson + 20, 0.0);
        fill(bg, kRobinson, tot);
        for (int i = 0; i < 20; i++) {
            double w[20], t = 0;
            for (int j = 0; j < 20; j++) {
                w[j] = (j == i) ? 0.0 : kRobinson[j] * std::pow(2.0, blz::kBlosum62[kStd[i]][kStd[j]] / 2.0);
                t += w[j];
            }
            fill(sub[i], w, t);
        }
        for (int &v : idx_of) v = -1;
        for (int i = 0; i < 20; i++) idx_of[kStd[i]] = i;
    }
    static void fill(uint8_t *out, const double *w, double tot) {
        double c = 0;
        int j = 0;
        double edge = w[0] / tot;
        for (int u = 0; u < 65536; u++) {
            double x = (u + 0.5) / 65536.0;
            while (x > edge && j < 19) {
                j++;
                c = edge;
                edge = c + w[j] / tot;
            }
            out[u] = (uint8_t)kStd[j];
        }
    }
};

const Tables &tables() {
    static Tables t;
    return t;
}

inline uint8_t draw_bg(Rng &r) { return tables().bg[r.next() >> 48]; }

int draw_length(const blz_synth_spec &sp, uint64_t i) {
    Rng r(stream_key(sp.seed, i, 1));
    double len = sp.len_median * std::exp(sp.len_sigma * r.normal());
    long v = std::lround(len);
    if (v < sp.len_min) v = sp.len_min;
    if (v > sp.len_max) v = sp.len_max;
    return (int)v;
}

bool is_homolog(const blz_synth_spec &sp, uint64_t i) {
    Rng r(stream_key(sp.seed, i, 2));
    return r.uniform() < sp.homolog_frac;
}

// Mutated copy of q[a, a+n): substitutions to identity `ident`, indels at `rate`.
void mutate(const uint8_t *q, int n, double ident, double rate, double mean, Rng &r,
            std::vector<uint8_t> &out) {
    const Tables &t = tables();
    out.clear();
    for (int k = 0; k < n; k++) {
        if (r.uniform() < rate) {
            int len = r.geometric(mean);
            if (r.uniform() < 0.5) {
                for (int x = 0; x < len; x++) out.push_back(draw_bg(r));
            } else {
                k += len - 1;
                continue;
            }
        }
        uint8_t c = q[k];
        int ci = c < 32 ? t.idx_of[c] : -1;
        if (ci < 0 || r.uniform() < ident)
            out.push_back(c);
        else
            out.push_back(t.sub[ci][r.next() >> 48]);
    }
}

void fill_subject(const blz_synth_spec &sp, uint64_t gen_index, int len, int hq,
                  const uint8_t *const *queries, const int32_t *qlens, uint8_t *dst) {
    Rng r(stream_key(sp.seed, gen_index, 3));
    for (int k = 0; k < len; k++) dst[k] = draw_bg(r);
    if (hq >= 0) {
        const uint8_t *q = queries[hq];
        int ql = qlens[hq];
        int seg = std::min(ql, len);
        int a = ql > seg ? (int)(r.next() % (uint64_t)(ql - seg + 1)) : 0;
        double ident = sp.ident_min + (sp.ident_max - sp.ident_min) * r.uniform();
        std::vector<uint8_t> core;
        mutate(q + a, seg, ident, sp.indel_rate, sp.indel_mean, r, core);
        int cl = (int)core.size();
        int take = std::min(cl, len);
        int from = cl > take ? (int)(r.next() % (uint64_t)(cl - take + 1)) : 0;
        int at = len > take ? (int)(r.next() % (uint64_t)(len - take + 1)) : 0;
        std::memcpy(dst + at, core.data() + from, take);
    }
    if (sp.repeat_frac > 0 && len >= 20) {
        long target = std::lround(sp.repeat_frac * len);
        long covered = 0;
        while (covered < target) {
            int period = sp.repeat_period_min +
                         (int)(r.next() % (uint64_t)(sp.repeat_period_max - sp.repeat_period_min + 1));
            int blk = std::min<long>(len, std::max<long>(2 * period, std::min<long>(target - covered, 60)));
            int at = (int)(r.next() % (uint64_t)(len - blk + 1));
            uint8_t unit[64];
            for (int k = 0; k < period; k++) unit[k] = draw_bg(r);
            for (int k = 0; k < blk; k++) dst[at + k] = unit[k % period];
            covered += blk;
        }
    }
}

}  // namespace

extern "C" {

void blz_synth_default(blz_synth_spec *sp) {
    std::memset(sp, 0, sizeof(*sp));
    sp->seed = 0;
    sp->n_subjects = 10000;
    sp->len_median = 300;
    sp->len_sigma = 0.8;
    sp->len_min = 30;
    sp->len_max = 5000;
    sp->homolog_frac = 0.02;
    sp->ident_min = 0.5;
    sp->ident_max = 0.5;
    sp->indel_rate = 0.01;
    sp->indel_mean = 3.0;
    sp->repeat_frac = 0.0;
    sp->repeat_period_min = 3;
    sp->repeat_period_max = 10;
    sp->sort_by_length = 1;
}

void blz_synth_query(uint64_t seed, int32_t qlen, uint8_t *out) {
    Rng r(stream_key(seed, 0x51, 4));
    for (int k = 0; k < qlen; k++) out[k] = draw_bg(r);
}

blz_status blz_synth_db_size(const blz_synth_spec *sp, int64_t *n_subjects, uint64_t *n_residues) {
    if (!sp || sp->n_subjects < 0 || sp->len_min < 1 || sp->len_max < sp->len_min) return BLZ_EINVAL;
    uint64_t tot = 0;
    for (int64_t i = 0; i < sp->n_subjects; i++) tot += (uint64_t)draw_length(*sp, (uint64_t)i);
    if (n_subjects) *n_subjects = sp->n_subjects;
    if (n_residues) *n_residues = tot;
    return BLZ_OK;
}

blz_status blz_synth_db_fill(const blz_synth_spec *sp, const uint8_t *const *queries,
                             const int32_t *qlens, int32_t nq, uint8_t *residues,
                             uint64_t *offsets, int32_t *homolog_of) {
    if (!sp || !residues || !offsets || sp->n_subjects < 0) return BLZ_EINVAL;
    if (sp->homolog_frac > 0 && (nq <= 0 || !queries || !qlens)) return BLZ_EINVAL;
    if (sp->repeat_period_min < 1 || sp->repeat_period_max > 64 ||
        sp->repeat_period_max < sp->repeat_period_min)
        return BLZ_EINVAL;
    const int64_t n = sp->n_subjects;
    std::vector<int> len(n);
    std::vector<int> hq(n, -1);
    int64_t nh = 0;
    for (int64_t i = 0; i < n; i++) {
        len[i] = draw_length(*sp, (uint64_t)i);
        if (sp->homolog_frac > 0 && is_homolog(*sp, (uint64_t)i)) hq[i] = (int)(nh++ % nq);
    }
    std::vector<int64_t> order(n);
    std::iota(order.begin(), order.end(), 0);
    if (sp->sort_by_length)
        std::stable_sort(order.begin(), order.end(),
                         [&](int64_t a, int64_t b) { return len[a] < len[b]; });
    offsets[0] = 0;
    for (int64_t k = 0; k < n; k++) offsets[k + 1] = offsets[k] + (uint64_t)len[order[k]];
    unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<std::thread> th;
    for (unsigned t = 0; t < nt; t++) {
        th.emplace_back([&, t]() {
            for (int64_t k = t; k < n; k += nt) {
                int64_t g = order[k];
                fill_subject(*sp, (uint64_t)g, len[g], hq[g], queries, qlens, residues + offsets[k]);
            }
        });
    }
    for (auto &x : th) x.join();
    if (homolog_of)
        for (int64_t k = 0; k < n; k++) homolog_of[k] = hq[order[k]];
    return BLZ_OK;
}

}  // extern "C"
