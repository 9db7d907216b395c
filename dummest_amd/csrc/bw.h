// Device-side layout of the profile builder's Baum-Welch E-step (bw.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace dm {

// Baum-Welch E-step of the profile builder (dummest-build.cc:561-736), FP64.
// Per sequence of n letters the six (L+2) x (n+2) matrices X, Y, Z (forward:
// match, delete, insert) and Wb, Yb, Zb (backward), row-major with pitch n+2
// (bwCells(L, n) doubles each), at cellOff[s] from each base pointer; `raw` holds per sequence and profile
// row i-1 (i = 1..L+1) the sums gamma, etap, betap, beta, epsilonp, epsilon,
// emis[0..A) of calculateTransitionCounts (kBwRaw + A values).
constexpr int kBwRaw = 6;
inline int64_t bwCells(int L, int n) { return (int64_t)(L + 2) * (n + 2); }
struct BWArgs {
    int L, A;
    int seqBase;               // first sequence of this launch
    const uint8_t *letters;    // letter codes (0..A), all sequences back to back, one pad after
    const int64_t *letterOff;  // per sequence
    const int *len;            // per sequence
    const int64_t *cellOff;    // per sequence, relative to the matrix bases
    const double *rates;       // (L+1) x 4 per profile row: a' = ins (1 - ext), ext, d', e'
    const double *ratio;       // (L+1) x (A+1): match likelihood ratio per row and letter (0 for A)
    double *X, *Y, *Z, *Wb, *Yb, *Zb;
    const double *weights;  // per sequence
    double *total;          // out: per sequence forward total v
    double *wt;             // out: per sequence weight / (v 2^-480)
    double *raw;            // per sequence (L+1) x (kBwRaw + A)
};
struct BWReduceArgs {
    int L, A, nseq;
    const double *wt;  // as in BWArgs
    const double *raw;
    double *counts;  // out: (L+1) x (7 + A) expected counts (delete starts left 0)
};
// Letters of one sequence sit in LDS during forward and backward: maxLen (the
// longest sequence of the launch) must stay within kBwMaxLen.
constexpr int kBwMaxLen = 40000;
hipError_t launch_bw_forward(const BWArgs &a, int nseq, int maxLen, hipStream_t s);
hipError_t launch_bw_backward(const BWArgs &a, int nseq, int maxLen, hipStream_t s);
hipError_t launch_bw_counts(const BWArgs &a, int nseq, hipStream_t s);
hipError_t launch_bw_reduce(const BWReduceArgs &a, hipStream_t s);

}  // namespace dm
