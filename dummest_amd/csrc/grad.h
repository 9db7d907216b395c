// Device-side layout of the summed-likelihood gradient (grad.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "dm_layout.h"

namespace dm {

// Per pair (profile row i = 0..M, DNA column j = 0..n-1) six FP64 matrices:
// the forward w0, its delete chain Y and insert chain Z0, and the backward W1,
// Yb, Zb (dummest-dp-forward.hh:9-250, dummest-dp-backward.hh:444-615). Both
// passes run the rows in 64-row strips (strip s = rows 64s..64s+63), lane k of
// a wavefront one row, one column behind its neighbour, and store each step's
// 64 values contiguously: matrix[s][t][lane] with t = 0..n+62 (forward: row
// 64s+lane, column t-lane; backward: row 64s+63-lane, column n-1-t+lane),
// zeros where the cell lies outside the profile or the sequence. Values are
// stored scaled by a power of two per strip and 64-column block (expF / expB:
// true value = stored x 2^exponent), so any ln Z fits. Expected counts per
// row: the 11 parameters of the row's Params (alpha'[3], beta'[3], delta'[3],
// eps', enter) and the emission by codon (64 codons, 64 = a codon with a
// non-ACGT letter).
//
// Checkpoints: every strip keeps the edge row it hands to its neighbour strip
// (n values per strip and quantity, 1/64 of a matrix strip), so any run of
// strips can be computed again from its edges alone, with the same values. A
// pair whose matrices do not fit the budget runs both passes once without
// storing matrices (the edges only), then per segment of strips the two
// passes again with the matrices and the count pass: O((M/64) n) edge memory
// plus one segment's matrices in place of O(M n).
constexpr int kGradParams = 11;
constexpr int kGradCodons = 65;
constexpr int kGradRaw = kGradParams + kGradCodons;
// longest sequence: one 64-row strip of its matrices takes 3 KiB per letter
constexpr int kGradMaxN = 1 << 28;

inline int64_t gradCells(int M, int n) { return (int64_t)(M + 1) * n; }
inline int gradStrips(int M) { return (M + 64) / 64; }  // 64-row strips covering rows 0..M
// stored values per matrix and pair: strips x steps x lanes
inline int64_t gradSlots(int M, int n) { return (int64_t)gradStrips(M) * (n + 63) * 64; }
inline int gradBlocks(int n) { return (n + 63) / 64; }  // 64-column blocks
// exponents of one pair: forward and backward each strips x blocks
inline int64_t gradExps(int M, int n) { return (int64_t)gradStrips(M) * gradBlocks(n); }

struct GradArgs {
    const ProfInfo *prof;  // the profile (one per call)
    const float *params;   // (M+2) x 11 at prof->parOff
    const float *emis;     // (M+1) x kEmStride at prof->emOff
    const float *bg;       // 32 background values of the profile
    const PairInfo *pairs;
    const Rec *rec;
    const uint8_t *codon;  // per pair: codon index of each column (at pairs[p].seqOff)
    int pairBase;          // first pair of this launch
    const int64_t *cellOff;  // per pair, relative to the matrix bases
    // strips stripLo..stripHi of each pair, strip s at cellOff[p] + (s -
    // stripLo) x (n+63) x 64; the forward and backward launches run those
    // strips and store them when store != 0, the count launch reads them
    double *w0, *Y, *Z0, *W1, *Yb, *Zb;
    int stripLo, stripHi, store;
    // per pair and strip (n values at bndOff[p] + s x n), the edge row a strip
    // hands on: forward (row 64s+63, to strip s+1) the inflow into the next row
    // (fwdA), Y (fwdB) and the exponent each column is stored at (fwdE);
    // backward (row 64s, to strip s-1) W1 (bwdA), Yb (bwdB), the column sum of
    // W1 over rows 64s..M (bwdC) and the exponent (bwdE). Strip 0's bwdC holds
    // the column sums over all rows.
    double *fwdA, *fwdB, *bwdA, *bwdB, *bwdC;
    int *fwdE, *bwdE;
    const int64_t *bndOff;
    // per pair (strips x blocks at expOff), strip s = rows 64s..64s+63 in both
    // passes; forward block c = columns 64c..64c+63, backward block c =
    // columns n-1-64c down to n-64-64c
    int *expF, *expB;
    const int64_t *expOff;
    double *zInfo;      // per slot: Z = zInfo[0] x 2^zInfo[1]
    const int *rawIdx;  // per pair: its slot in raw and zInfo (the caller's order)
    double *raw;        // per slot (M+1) x kGradRaw: flows / 2^zInfo[1]
};

struct GradReduceArgs {
    int M, npairs;
    const double *raw;
    const double *zInfo;
    const double *weight;  // per pair
    double *lnZ;           // out: per pair
    double *counts;        // out: (M+1) x kGradRaw, weighted sums of the expected counts
};

hipError_t launch_grad_forward(const GradArgs &a, int npairs, int maxN, hipStream_t s);
hipError_t launch_grad_backward(const GradArgs &a, int npairs, int maxN, hipStream_t s);
hipError_t launch_grad_z(const GradArgs &a, int npairs, hipStream_t s);
hipError_t launch_grad_counts(const GradArgs &a, int npairs, int M, int maxN, hipStream_t s);
// factor: npairs doubles of scratch (weight / Z mantissa per pair)
hipError_t launch_grad_reduce(const GradReduceArgs &a, double *factor, hipStream_t s);

}  // namespace dm
