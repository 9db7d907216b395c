// Host-side pieces of the drop-in: profile reading/finalizing, sequence
// decoding, and the post-gate processing of hit pairs (flank totals, MEA
// tables, traceback) over the DP matrices the gfx950 kernels produce.
#pragma once
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "dm_layout.h"

namespace dm {

struct HostProfile {
    std::string name;
    int M = 0;
    std::vector<float> values;  // (M+1) x 29, finalized
    std::vector<float> params;  // (M+2) x 11
    std::vector<float> bg;      // 29
    std::vector<char> consensus;  // letter index per position (dummest-profile.hh:310-312)
};

// readProfiles + finalizeProfile (dummest-profile.hh:105-329), -m 0.
bool readHmmFile(const char *path, int bgType, const double rates[6], std::vector<HostProfile> &out,
                 std::string &msg);
bool readHmmText(const std::string &text, int bgType, const double rates[6], std::vector<HostProfile> &out,
                 std::string &msg);

// decodeSequence (dummest-sequence.hh:181-204) on raw DNA letters.
void decodeDna(const char *dna, int64_t n, uint8_t *codes);

struct Seg {
    int start1, start2, length;
};
struct Hit {
    ES pair;
    double lnratio;
    int anchor1, anchor2;
    std::vector<Seg> segs;
};

// DP matrices of one pair as produced by the dump-mode kernel.
struct PostInput {
    int M, n, G, R;
    const float *w0, *W1;  // [(n+G-1) * G*R], step-major
    const int *ef, *eb;    // per-step exponents
    const uint8_t *codes;
    std::vector<float> prefixSeed;
    float s1, s2, sC;
    float bgfs1, bgfs2;
};

// forwardPass epilogue for one pair (dummest-dp-forward.hh:252-545): gate,
// flank totals, MEA tables, anchors, traceback, overlap rejection.
void postProcess(const HostProfile &P, const PostInput &in, float minProbRatio, std::vector<Hit> &out);

}  // namespace dm
