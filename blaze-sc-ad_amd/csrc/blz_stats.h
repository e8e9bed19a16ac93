// blz_stats.h -- statistics routines shared by the C ABI (blz_stats.cpp) and its tests.
#pragma once
#include <cstdint>
#include <vector>

#include "blz_prelim.h"

namespace blz {

// Karlin-Altschul lambda, K, H of a lattice score distribution P over scores [lo, lo + P.size());
// false unless the mean score is negative and some score is positive.
bool karlin_from_distribution(const std::vector<double> &P, int lo, blz_karlin &out);

// NCBI's length adjustment (BLAST_ComputeLengthAdjustment): the largest integer l with
// alpha/lambda (ln K + ln((m - l)(n - N l))) + beta >= l.
int32_t length_adjustment(double K, double logK, double alpha_d_lambda, double beta, int64_t query_len,
                          int64_t db_len, int64_t db_seqs);

}  // namespace blz
