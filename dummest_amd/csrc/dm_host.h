// Host-side pieces of the drop-in: profile reading/finalizing, sequence
// decoding, and the post-gate processing of hit pairs (flank totals, MEA
// tables, traceback) over the DP matrices the gfx950 kernels produce.
#pragma once
#include <cmath>
#include <cstdint>
#include <cstring>
#include <iosfwd>
#include <stdexcept>
#include <string>
#include <vector>

#include "dm_layout.h"

namespace dm {

// Worker threads for the library's host-side pools: the CPUs this process may
// run on (its affinity mask), at most 16.
int hostThreads();

struct HostProfile {
    std::string name;
    int M = 0;
    std::vector<float> values;  // (M+1) x 29, finalized
    std::vector<float> params;  // (M+2) x 11
    std::vector<float> bg;      // 29
    std::vector<char> consensus;  // letter index per position (dummest-profile.hh:310-312)
    int masked = 0;               // positions masked by tantan (profile masking)
};

// readProfiles + finalizeProfile (dummest-profile.hh:105-329). maskThr >= 0
// turns on profile masking (-m 1) at that tantan threshold; negative = off.
// filterSd > 0 applies the Gaussian letter filter of -d (keepNonvarying: -D)
// instead of masking (dummest-profile.hh:64-103, 119-121).
bool readHmmFile(const char *path, int bgType, const double rates[6], std::vector<HostProfile> &out,
                 std::string &msg, float maskThr = -1.f, double filterSd = 0, bool keepNonvarying = false);
bool readHmmText(const std::string &text, int bgType, const double rates[6], std::vector<HostProfile> &out,
                 std::string &msg, float maskThr = -1.f, double filterSd = 0, bool keepNonvarying = false);

// finalizeProfile on values as read (probabilities, (M+1) x 29, reader order).
bool finalizeRaw(HostProfile &p, int bgType, const double rates[6], std::string &msg, float maskThr = -1.f,
                 double filterSd = 0, bool keepNonvarying = false);

// tantan as DUMMEST calls it (tantan-wrapper.hh:13-64, gap-free HMM): the
// likelihood ratios (BLOSUM62 at lambda 0.324 for proteins, +1/-1 for DNA
// letters 0..3), the repeat-offset limit K (50 / 100) and the transition
// probabilities.
struct TantanParams {
    int K;
    double b2b, f2b, f2f0;
    double b2f[100];
    double ratio[22][22];
};
TantanParams tantanParams(bool isProtein);
// Posterior repeat probability of every letter (small-number alphabet).
void tantanRepeatProbs(const uint8_t *seq, int n, bool isProtein, float *out);

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

// DP matrices of one pair as exact ExpScores, (M+1) x n, row-major.
struct PostInput {
    int M = 0, n = 0;
    std::vector<ES> w0, W1;
    const uint8_t *codes = nullptr;
    std::vector<float> prefixSeed;
    float s1 = 0, s2 = 0, sC = 0;
    float bgfs1 = 0, bgfs2 = 0;
};

// An alignment anchor: its best anchored score and cell (profile row a1,
// column a2).
struct Anchor {
    ES pair;
    int a1, a2;
};

// The anchors of a pair in the order the traceback tries them
// (dummest-dp-forward.hh:459-500): every column's best cell, ranked highest
// first by std::sort over all columns (whose order of tied anchors the
// results depend on), those at or above minProbRatio kept. For minProbRatio <
// 0, only the first of the best anchors.
void rankAnchors(const ES *colBest, const int *colRow, int n, float minProbRatio, std::vector<Anchor> &out);

// A stored FP32 DP value at a power-of-two offset as an ExpScore (0 for
// non-positive values, +inf saturates as promoteRescaledScore does).
ES esFromStored(float v, int off);

// forwardPass epilogue for one pair (dummest-dp-forward.hh:252-545): gate,
// flank totals, MEA tables, anchors, traceback, overlap rejection.
void postProcess(const HostProfile &P, const PostInput &in, float minProbRatio, std::vector<Hit> &out);

// One calibration trial's random frameshifting codon DNA (L + border letters)
// from the profile's background letter frequencies (estimateK,
// dummest-calibration.hh:363-376, 436-482).
void calibSequence(const float *letterFreqs20, float bgStop, float bgfs1, float bgfs2, int L, int border, int trial,
                   char *out);

// estimateGumbel (dummest-calibration.hh:7-104) on one score sample:
// out = MM lambda, MM K, MM K (lambda 1), ML lambda, ML K, ML K (lambda 1),
// L-moment lambda, L-moment K.
void gumbelFits(std::vector<double> scores, double seqLen, double out[8]);

// Profile builder (dummest-build, dummest-build.cc).
struct BuildGapPrior {  // match, insStart, delStart, insEnd, insExtend, delEnd, delExtend pseudocounts
    double match = 0, insStart = 0, delStart = 0, insEnd = 0, insExt = 0, delEnd = 0, delExt = 0;
};
struct BuildOptions {
    bool hand = false;       // --hand: RF annotation defines the positions
    bool counts = false;     // --counts: print weighted counts
    bool countOnly = false;  // --countonly: no Baum-Welch
    bool pnone = false;      // --pnone: no priors
    double symfrac = 0.5;    // --symfrac
    double ere = 0;          // --ere (0: 0.59 protein, 0.62 DNA)
    double esigma = 45.0;    // --esigma (1e37: --enone)
    double maxIter = 1000;   // --maxiter
    double maxDiff = 1e-6;   // --maxdiff
    int verbosity = 0;
    std::vector<double> mixture;  // --dmix: per component the coefficient then the alphas
    int mixtureComponents = 0;
    bool hasGapPrior = false;  // --gapprior
    BuildGapPrior gapPrior;
    std::string versionText = "DUMMEST-AMD 1";  // in the header's brackets
    std::string bmText;                          // the BM line after "BM   "
    int device = 0;                              // HIP device for Baum-Welch
};
// A malformed alignment (fatal: the reference prints "FATAL: ..." and aborts)
// or a Baum-Welch overflow (the reference prints the message and exits 1).
struct BuildError : std::runtime_error {
    explicit BuildError(const char *m, bool fatalError = true) : std::runtime_error(m), fatal(fatalError) {}
    bool fatal;
};
bool readGapPrior(std::istream &in, BuildGapPrior &gp);
bool readMixture(std::istream &in, std::vector<double> &params, int &components);
// Builds a profile per alignment of `in`, writes them to out and progress to
// log; returns 0 or 1 (after an error message on log); throws BuildError.
int buildProfiles(std::istream &in, const BuildOptions &o, std::ostream &out, std::ostream &log);

// Calibration cache (ProfileCache, dummest-calibration.hh:106-336).
struct Hash128 {
    uint64_t lo, hi;
};
// XXH3 128-bit hash, seed 0, default secret.
Hash128 xxh3_128(const void *data, size_t len);
// high 64 bits then low 64 bits, 16 lowercase hex digits each
std::string hashHex(const Hash128 &h);
// Hash of the running executable and of this library.
std::string binaryHash();
// The 27 doubles of a cache entry, in the reference's CacheEntry order:
// lambda, K and K (lambda = 1) of the end-, start- and mid-anchored samples for
// the method of moments, maximum likelihood (lambda, K, K1) and L-moments
// (lambda, K), then the filter's method-of-moments mid lambda, K and K1.
struct CacheEntry {
    double v[27];
};
// computeCacheKey (dummest-calibration.hh:201-226) for a finalized profile.
std::string calibrationKey(const std::string &binHash, const HostProfile &p, int seqLen, int border, int trials,
                           const double rates[6], float tantanThr);
// cache.bin in the first usable cache directory (getCacheFilePath, :233-265).
std::string calibrationCachePath();
// cereal BinaryArchive layout of unordered_map<string, CacheEntry>. A missing or
// empty file reads as no entries; a malformed one returns false with err.
bool readCalibrationCache(const std::string &path, std::vector<std::pair<std::string, CacheEntry>> &out,
                          std::string &err);
bool writeCalibrationCache(const std::string &path, const std::vector<std::pair<std::string, CacheEntry>> &entries);

}  // namespace dm
