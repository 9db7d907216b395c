// Device-side layout of the batched post-gate path (postgate.hip): for every
// gated pair of a batch, from the dump-mode DP matrices to hits and segments.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "dm_layout.h"

namespace dm {

// Per-column arrays of a pair (kPostCols arrays of n+8 ExpScores each).
enum PostCol { kColBest = 0, kFwdSide, kRevSide, kMid, kSr, kUni, kRevEv, kFwdEv, kFa, kWe, kPostCols };

struct PostPair {
    int M, n;
    int G, R, T;    // dump geometry: lanes per pair, rows per lane, steps per strip
    int prof;       // profile index
    int pairIdx;    // the pair's record in the dump chunk (PairInfo: seeds, scales)
    float minRatio; // minProbRatio (negative: the best anchor only, no alignment)
    int64_t wOff;   // dump: float offset of the pair's w0 and W1 values
    int64_t eOff;   // dump: int offset of its per-step exponents
    int64_t matOff; // ExpScore offset of X, pre, suf: (M+2) rows of n+7 each, column j at j+3
    int64_t colOff; // ExpScore offset of the column arrays
    int64_t intOff; // int offset: colRow (n+8), rowMaxExp (M+2), aligned (n+8)
    int64_t segOff; // Seg offset of the output (segCap)
    int64_t hitOff; // PostHit offset (ncand)
    int64_t candOff;  // offset of the sorted candidates (filled by the host)
    int ncand;        // candidates at or above the threshold (filled by the host)
    int segCap;
};

struct PostRes {
    unsigned long long gkey;  // the largest w0 W1 as an ordered key (exponent, mantissa)
    int revMax, fwdMax;       // largest exponents of the flank totals
    int live;                 // passed the gate with minRatio >= 0: tables built
    int nhits, nsegs;
    int overflow;             // a walk needed more segments than its buffer holds
};

struct PostSeg {
    int start1, start2, length;
};
struct PostCand {
    float m;
    int e;
    int a1, a2;
};
struct PostHit {
    float m;
    int e;
    int a1, a2, seg0, nseg;
};

struct PostArgs {
    int npairs;
    const PostPair *pp;
    PostRes *res;
    const ProfInfo *profs;
    const float *params, *emis, *bg;  // bg: 32 per profile
    const PairInfo *pairs;
    const Rec *rec;
    const uint8_t *codes;
    const float *w0buf, *wbuf;
    const int *fbuf, *ebuf;
    ES *mat, *col;
    int *ints;
    PostSeg *segs;
    PostHit *hits;
    PostCand *cands;  // per pair n: each column's best cell, then the ranked anchors
    float bgfs1, bgfs2;
};

// zero the tables, then the per-column pass (anchors, column totals, X), the
// flank totals, the MEA tables, all before the host ranks the candidates
hipError_t launch_post_tables(const PostArgs &a, int maxN, int maxM, int64_t matCells, hipStream_t s);
// rows per lane the per-column pass is built for
bool post_rows_supported(int R);
// the tracebacks of the ranked candidates
hipError_t launch_post_walk(const PostArgs &a, hipStream_t s);

}  // namespace dm
