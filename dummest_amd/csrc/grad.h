// Device-side layout of the summed-likelihood gradient (grad.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "dm_layout.h"

namespace dm {

// Per pair (profile row i = 0..M, DNA column j = 0..n-1) six FP64 matrices,
// row-major with pitch n: the forward w0, its delete chain Y and insert chain
// Z0, and the backward W1, Yb, Zb (dummest-dp-forward.hh:9-250,
// dummest-dp-backward.hh:444-615, here without rescaling: FP64 holds every
// value of a pair whose ln Z stays below ~700 nats). Expected counts per row:
// the 11 parameters of the row's Params (alpha'[3], beta'[3], delta'[3], eps',
// enter) and the emission by codon (64 codons, 64 = a codon with a non-ACGT
// letter), then the row's share of Z = sum_j prefixSeed(j) W1(i, j).
constexpr int kGradParams = 11;
constexpr int kGradCodons = 65;
constexpr int kGradRaw = kGradParams + kGradCodons + 1;
// the pair's per-column codes and seeds sit in LDS: sequences up to kGradMaxN
constexpr int kGradMaxN = 4096;

inline int64_t gradCells(int M, int n) { return (int64_t)(M + 1) * n; }

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
    double *w0, *Y, *Z0, *W1, *Yb, *Zb;
    double *bndA, *bndB;  // per pair (n each, at the pair's bndOff): strip hand-over of two values per column
    const int64_t *bndOff;
    const int *rawIdx;  // per pair: its slot in raw (the caller's order)
    double *raw;        // per slot (M+1) x kGradRaw
};

struct GradReduceArgs {
    int M, npairs;
    const double *raw;
    const double *weight;  // per pair
    double *lnZ;           // out: per pair
    double *counts;        // out: (M+1) x (kGradRaw - 1), weighted sums of the expected counts
};

hipError_t launch_grad_forward(const GradArgs &a, int npairs, int maxN, hipStream_t s);
hipError_t launch_grad_backward(const GradArgs &a, int npairs, int maxN, hipStream_t s);
hipError_t launch_grad_counts(const GradArgs &a, int npairs, int M, hipStream_t s);
// factor: npairs doubles of scratch (weight / Z per pair)
hipError_t launch_grad_reduce(const GradReduceArgs &a, double *factor, hipStream_t s);

}  // namespace dm
