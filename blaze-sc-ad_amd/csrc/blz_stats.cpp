// blz_stats.cpp -- search statistics of the prelim stage: Karlin-Altschul parameters, effective
// lengths, cutoffs, X-drops from bits, the prelim e-value reap threshold and the per-query hit list.
//
// The reference takes all of these from NCBI BLAST+ 2.13.0 at run time (the cutoffs it passes
// to its hook are computed at BLAZE_ncbi/blast_engine.c:2971-2981; the e-values and the prelim
// reap are blast_engine.c:1436-1457 and :896-930; the hit list is built from the HSP stream in
// BLAZE_ncbi/prelim_stage.cpp:1462-1498).  NCBI's routines (blast_stat.c, blast_parameters.c,
// blast_hits.c) are not vendored by the reference; this file restates their published methods:
//   - ungapped lambda, H and K: Karlin & Altschul (1990), PNAS 87:2264, lattice case;
//   - gapped lambda, K, H, alpha, beta: NCBI's published table for BLOSUM62 (no closed form);
//   - length adjustment: the largest integer l with l <= alpha/lambda (ln K + ln((m - l)(n - N l)))
//     + beta, found by NCBI's bracketing iteration (BLAST_ComputeLengthAdjustment);
//   - cutoff from an e-value: S = ceil(ln(K * searchsp / E) / lambda) (BlastKarlinEtoS_simple);
//   - e-value of a score: searchsp * exp(-lambda * S + ln K) (BLAST_KarlinStoE_simple).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <vector>

#include "blz_common.h"
#include "blz_prelim.h"
#include "blz_stats.h"

using namespace blz;

namespace {

// Robinson & Robinson (1991) background frequencies of the 20 standard residues, the background
// NCBI uses for protein statistics; ambiguity letters have none.  Indexed by NCBIstdaa code.
struct Background {
    double p[kAlpha] = {};
    Background() {
        const struct { int code; double f; } rr[] = {
            {1, 78.05}, {3, 19.25}, {4, 53.64}, {5, 62.95}, {6, 38.56}, {7, 73.77}, {8, 21.99},
            {9, 51.42}, {10, 57.44}, {11, 90.19}, {12, 22.43}, {13, 44.87}, {14, 52.03}, {15, 42.64},
            {16, 51.29}, {17, 71.20}, {18, 58.41}, {19, 64.41}, {20, 13.30}, {22, 32.16}};
        double sum = 0;
        for (const auto &x : rr) sum += x.f;
        for (const auto &x : rr) p[x.code] = x.f / sum;
    }
};
const Background kRobinson;

// Published Karlin-Altschul parameters of BLOSUM62 with affine gap costs (NCBI blast_stat.c's
// BLOSUM62 table; gapped statistics have no closed form and were fitted by simulation).
struct GappedRow {
    int open, extend;
    double lambda, K, H, alpha, beta;
};
const GappedRow kBlosum62Gapped[] = {
    {11, 2, 0.297, 0.082, 0.27, 1.1, -10}, {10, 2, 0.291, 0.075, 0.23, 1.3, -15},
    {9, 2, 0.279, 0.058, 0.19, 1.5, -19},  {8, 2, 0.264, 0.045, 0.15, 1.8, -26},
    {7, 2, 0.239, 0.027, 0.10, 2.5, -46},  {6, 2, 0.201, 0.012, 0.061, 3.3, -58},
    {13, 1, 0.292, 0.071, 0.23, 1.2, -11}, {12, 1, 0.283, 0.059, 0.19, 1.5, -19},
    {11, 1, 0.267, 0.041, 0.14, 1.9, -30}, {10, 1, 0.243, 0.024, 0.10, 2.5, -44},
    {9, 1, 0.206, 0.010, 0.052, 4.0, -87},
};
constexpr double kLn2 = 0.69314718055994530942;

// Lambda: the positive root of sum_s P(s) exp(lambda s) = 1, which exists when the expected score
// is negative and some score is positive.  Bisection on a bracket grown until the sum exceeds 1
// (the sum is convex in lambda, below 1 just above 0 and unbounded as lambda grows).
bool solve_lambda(const std::vector<double> &P, int lo, double &lambda) {
    auto f = [&](double l) {
        double s = 0;
        for (size_t k = 0; k < P.size(); k++)
            if (P[k] > 0) s += P[k] * std::exp(l * (double)((int)k + lo));
        return s - 1.0;
    };
    double a = 0, b = 0.5;
    while (f(b) <= 0) {
        b *= 2;
        if (b > 1e3) return false;
    }
    // the left end: just above 0, where f < 0 because the mean score is negative
    a = b;
    while (a > 1e-12 && f(a) > 0) a /= 2;
    if (f(a) > 0) return false;
    for (int it = 0; it < 200; it++) {
        const double m = 0.5 * (a + b);
        (f(m) > 0 ? b : a) = m;
    }
    lambda = 0.5 * (a + b);
    return true;
}

}  // namespace

namespace blz {

// Karlin-Altschul parameters of a score distribution P over [lo, lo + P.size()) (lattice case):
//   H = lambda * sum_s s P(s) exp(lambda s);
//   K = lambda d exp(-2 sigma) / (H (1 - exp(-lambda d))), d = gcd of the scores with P > 0,
//   sigma = sum_{k>=1} (1/k) [ E(exp(lambda S_k); S_k < 0) + P(S_k >= 0) ],
// S_k the sum of k scores (the terms fall geometrically; summed until they stop adding).
bool karlin_from_distribution(const std::vector<double> &P, int lo, blz_karlin &out) {
    double mean = 0;
    int hi = lo - 1, d = 0;
    for (size_t k = 0; k < P.size(); k++)
        if (P[k] > 0) {
            const int s = (int)k + lo;
            mean += P[k] * s;
            hi = s;
            d = std::gcd(d, std::abs(s));
        }
    if (mean >= 0 || hi <= 0 || d == 0) return false;
    double lambda;
    if (!solve_lambda(P, lo, lambda)) return false;
    double H = 0;
    for (size_t k = 0; k < P.size(); k++) {
        const double s = (double)((int)k + lo);
        H += P[k] * s * std::exp(lambda * s);
    }
    H *= lambda;
    double sigma = 0;
    std::vector<double> Pk = P, next;
    int klo = lo;
    for (int k = 1; k <= 2000; k++) {
        double term = 0;
        for (size_t i = 0; i < Pk.size(); i++) {
            const int s = (int)i + klo;
            term += s < 0 ? Pk[i] * std::exp(lambda * s) : Pk[i];
        }
        sigma += term / k;
        if (term / k < 1e-16 * sigma) break;
        next.assign(Pk.size() + P.size() - 1, 0.0);  // distribution of S_{k+1}
        for (size_t i = 0; i < Pk.size(); i++)
            if (Pk[i] > 0)
                for (size_t j = 0; j < P.size(); j++) next[i + j] += Pk[i] * P[j];
        Pk.swap(next);
        klo += lo;
    }
    out.lambda = lambda;
    out.H = H;
    out.K = lambda * d * std::exp(-2 * sigma) / (H * (1 - std::exp(-lambda * d)));
    out.logK = std::log(out.K);
    return true;
}

// BLAST_ComputeLengthAdjustment restated: the largest integer l in [0, l_max] with
// f(l) = alpha/lambda (ln K + ln((m - l)(n - N l))) + beta >= l, where l_max is the root of
// (m - l)(n - N l) = max(m, n) / K.  Bracketing iteration: l_min is the largest value seen with
// f(l) >= l, l_max the smallest with f(l) < l; a proposal f(l) inside the bracket is taken, else
// the bracket's midpoint (its top on the first step).  After convergence floor(l_min) is the
// answer unless ceil(l_min) also satisfies the inequality.
int32_t length_adjustment(double K, double logK, double alpha_d_lambda, double beta, int64_t query_len,
                          int64_t db_len, int64_t db_seqs) {
    const double m = (double)query_len, n = (double)db_len, N = (double)db_seqs;
    const double c = n * m - std::max(m, n) / K;
    if (c < 0) return 0;
    const double a = N, b = N * m + n;
    double ell_max = 2 * c / (b + std::sqrt(b * b - 4 * a * c));
    double ell_min = 0, ell_next = 0;
    bool converged = false;
    for (int i = 1; i <= 20; i++) {
        const double ell = ell_next;
        const double ss = (m - ell) * (n - N * ell);
        const double ell_bar = alpha_d_lambda * (logK + std::log(ss)) + beta;
        if (ell_bar >= ell) {
            ell_min = ell;
            if (ell_bar - ell_min <= 1.0) {
                converged = true;
                break;
            }
            if (ell_min == ell_max) break;
        } else {
            ell_max = ell;
        }
        if (ell_min <= ell_bar && ell_bar <= ell_max)
            ell_next = ell_bar;
        else
            ell_next = i == 1 ? ell_max : 0.5 * (ell_min + ell_max);
    }
    int32_t adj = (int32_t)ell_min;
    if (converged) {
        const double ell = std::ceil(ell_min);
        if (ell <= ell_max) {
            const double ss = (m - ell) * (n - N * ell);
            if (alpha_d_lambda * (logK + std::log(ss)) + beta >= ell) adj = (int32_t)ell;
        }
    }
    return adj;
}

}  // namespace blz

extern "C" {

void blz_stats_default_options(blz_stats_options *o) {
    std::memset(o, 0, sizeof(*o));
    o->expect = 10.0;
    o->prelim_evalue = 0.0;
    o->x_ungapped_bits = 7.0;
    o->x_gapped_bits = 15.0;
    o->x_final_bits = 25.0;
    o->gap_trigger_bits = 22.0;
    o->gap_open = 11;
    o->gap_extend = 1;
    o->word_threshold = 11;
    o->window = 40;
    o->query_composition = 0;
    o->max_target_seqs = 500;
    o->composition_based_stats = 2;
}

blz_status blz_karlin_ungapped(const double *freq, blz_karlin *out) {
    if (!out) return BLZ_EINVAL;
    const double *p = freq ? freq : kRobinson.p;
    double sum = 0;
    for (int i = 0; i < kAlpha; i++) {
        if (!(p[i] >= 0)) return BLZ_EINVAL;
        sum += p[i];
    }
    if (!(sum > 0)) return BLZ_EINVAL;
    // scores of letter pairs with both frequencies, weighted by the product (against the
    // Robinson background on the database side, as NCBI scores a query's composition)
    constexpr int lo = -4, hi = 11;  // BLOSUM62's range over letters 1..27
    std::vector<double> P(hi - lo + 1, 0.0);
    for (int i = 1; i < kAlpha; i++)
        for (int j = 1; j < kAlpha; j++) {
            const double w = p[i] / sum * kRobinson.p[j];
            if (w > 0) P[kBlosum62[i][j] - lo] += w;
        }
    return karlin_from_distribution(P, lo, *out) ? BLZ_OK : BLZ_EINVAL;
}

blz_status blz_karlin_gapped(int32_t gap_open, int32_t gap_extend, blz_karlin *out, double *alpha, double *beta) {
    if (!out) return BLZ_EINVAL;
    for (const GappedRow &r : kBlosum62Gapped)
        if (r.open == gap_open && r.extend == gap_extend) {
            out->lambda = r.lambda;
            out->K = r.K;
            out->logK = std::log(r.K);
            out->H = r.H;
            if (alpha) *alpha = r.alpha;
            if (beta) *beta = r.beta;
            return BLZ_OK;
        }
    return BLZ_EINVAL;
}

double blz_evalue(const blz_search_stats *s, int32_t score) {
    if (!s) return -1.0;
    return s->eff_searchsp * std::exp(-s->gapped.lambda * (double)score + s->gapped.logK);
}

blz_status blz_search_setup(const blz_stats_options *o, const uint8_t *query, int32_t qlen, uint64_t db_residues,
                            int64_t db_seqs, blz_params *params, blz_search_stats *st) {
    if (!o || !st || qlen < 1 || db_seqs < 1 || db_residues < 1 || !(o->expect > 0) || o->prelim_evalue < 0 ||
        (o->query_composition && !query))
        return BLZ_EINVAL;
    std::memset(st, 0, sizeof(*st));
    // ungapped block: the background on both sides, or the query's composition on one side
    if (o->query_composition) {
        double f[kAlpha] = {};
        for (int i = 0; i < qlen; i++) {
            if (query[i] >= kAlpha) return BLZ_EINVAL;
            if (query[i] > 0) f[query[i]] += 1;
        }
        if (blz_karlin_ungapped(f, &st->ungapped) != BLZ_OK) return BLZ_EINVAL;
    } else if (blz_karlin_ungapped(nullptr, &st->ungapped) != BLZ_OK) {
        return BLZ_EINVAL;
    }
    if (blz_karlin_gapped(o->gap_open, o->gap_extend, &st->gapped, &st->alpha, &st->beta) != BLZ_OK) return BLZ_EINVAL;
    // effective lengths (BLAST_CalcEffLengths with the gapped block)
    st->length_adjustment = length_adjustment(st->gapped.K, st->gapped.logK, st->alpha / st->gapped.lambda, st->beta,
                                              qlen, (int64_t)db_residues, db_seqs);
    st->eff_query_len = qlen - st->length_adjustment;
    if (st->eff_query_len < 1) st->eff_query_len = 1;
    st->eff_db_len = (int64_t)db_residues - db_seqs * (int64_t)st->length_adjustment;
    if (st->eff_db_len < 1) st->eff_db_len = 1;
    st->eff_searchsp = (double)st->eff_db_len * (double)st->eff_query_len;
    // gapped cutoff from the expect value (BLAST_Cutoffs with no user score: S from E)
    const double E = std::max(o->expect, 1e-7);
    const int32_t gapped_cutoff =
        (int32_t)std::ceil(std::log(st->gapped.K * st->eff_searchsp / E) / st->gapped.lambda);
    // gap trigger in bits -> raw with the ungapped block; the ungapped cutoff is the smaller of it
    // and the gapped cutoff
    st->gap_trigger = (int32_t)((o->gap_trigger_bits * kLn2 + st->ungapped.logK) / st->ungapped.lambda);
    const int32_t ungapped_cutoff = std::min(st->gap_trigger, gapped_cutoff);
    // X-drops: ungapped rounded up with the ungapped block, gapped truncated with the gapped one
    const int32_t x_ungapped = (int32_t)std::ceil(o->x_ungapped_bits * kLn2 / st->ungapped.lambda);
    const int32_t x_gapped = (int32_t)(o->x_gapped_bits * kLn2 / st->gapped.lambda);
    st->x_final = (int32_t)(o->x_final_bits * kLn2 / st->gapped.lambda);
    // prelim reap: HSPs whose e-value exceeds prelim_evalue are discarded.  The e-value falls as
    // the score rises, so the reap is "score >= reap_score", reap_score the lowest score whose
    // e-value (the same double expression as blz_evalue) is within the threshold.
    st->prelim_evalue = o->prelim_evalue > 0 ? o->prelim_evalue : o->expect;
    int32_t s = std::max<int32_t>(1, (int32_t)std::floor(std::log(st->gapped.K * st->eff_searchsp / st->prelim_evalue) /
                                                        st->gapped.lambda) - 2);
    while (blz_evalue(st, s) > st->prelim_evalue) s++;
    while (s > 1 && blz_evalue(st, s - 1) <= st->prelim_evalue) s--;
    st->reap_score = s;
    // prelim hit-list size (SBlastHitsParametersNew): composition-based statistics keep 2x + 50
    // lists for the traceback's rescoring, plain gapped searches min(2x, x + 50)
    const int32_t hl = std::max(1, o->max_target_seqs);
    st->hitlist_size = o->composition_based_stats ? 2 * hl + 50 : std::min(2 * hl, hl + 50);
    if (params) {
        blz_default_params(params);
        params->word_threshold = o->word_threshold;
        params->window = o->window;
        params->x_ungapped = x_ungapped;
        params->x_gapped = x_gapped;
        params->gap_open = o->gap_open;
        params->gap_extend = o->gap_extend;
        params->ungapped_cutoff = ungapped_cutoff;
        params->gapped_cutoff = gapped_cutoff;
        params->reap_score = st->reap_score;
    }
    return BLZ_OK;
}

}  // extern "C"
