// ORACLE -- test infrastructure only, never part of the product.
//
// A CPU restatement of DUMMEST's frameshift-aware DNA-vs-profile-HMM search
// path, used as the checker for the HIP implementation. Only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg load this library.
//
// Written in C++ (not C) because two pieces of the reference's observable
// behaviour are libstdc++ semantics: the unstable std::sort tie order used when
// ranking anchors (dummest-dp-forward.hh:500) and the <random> distributions
// that generate calibration sequences (dummest-calibration.hh:436-477).
//
// Every function cites the reference file:line it restates. Arithmetic follows
// the reference's operation order left to right, with no FMA contraction
// (built with -ffp-contract=off) except where the reference calls fma
// explicitly. The reference build may contract mul+add pairs on its own; that
// choice is compiler-dependent and is not reproduced here.
//
// Precision: every DP entry point exists for float (the dummestl_core build)
// and for double (the dummest_core build, -DDOUBLE).
//
// Pinning (how each part is checked independently of this file's formulas):
//   null model      brute-force enumeration of all 1/2/3-nt tilings (tests)
//   DP recurrences  brute-force path enumeration on tiny inputs, total-mass
//                   equality, lane-batch invariance, power-of-two rescale
//                   invariance (tests)
//   ExpScore ops    frexp/ldexp contracts of tests/test_batch_pairs.cc
//   Gumbel fits     closed forms on synthetic Gumbel samples
//   profile parse   closed forms on the committed HMM fixtures
//   traceback       the herpes/medium span fingerprints need tantan masking
//                   (-m 3), which this oracle does not restate: parity of
//                   span coordinates against the reference is unpinned.

#include <algorithm>
#include <cassert>
#include <cfloat>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <functional>
#include <random>
#include <string>
#include <unordered_map>
#include <numeric>
#include <vector>
#include <xmmintrin.h>
#include <pmmintrin.h>

namespace orc {

// ----------------------------------------------------------------------------
// ExpScore: value = m * 2^e, m in [0.5,1) or m == 0.  dummest-score.hh:29-33
// ----------------------------------------------------------------------------
struct ES {
    float m = 0.0f;
    int32_t e = 0;
};

// dummest-core.hh:143-146: the reference runs every worker thread with FTZ +
// DAZ. Entry points set it for their duration and restore the caller's mode.
struct FtzScope {
    unsigned old;
    FtzScope() : old(_mm_getcsr()) {
        _MM_SET_FLUSH_ZERO_MODE(_MM_FLUSH_ZERO_ON);
        _MM_SET_DENORMALS_ZERO_MODE(_MM_DENORMALS_ZERO_ON);
    }
    ~FtzScope() { _mm_setcsr(old); }
};

// Planted mistakes for the tests that show a check can fail (0 = none):
// 1 backward 1-nt/2-nt insertion scales swapped, 2 forward match emission read
// one column late, 3 filter match step scaled as two letters, 4 backward codon
// insertion background read one column late. Set only by or_set_mutation.
static int g_mutation = 0;

// dummest-score.hh:62-67
static inline ES normalizeScore(double m, int e) {
    if (m == 0.0) return ES{0.0f, 0};
    int ex = 0;
    m = std::frexp(m, &ex);
    return ES{(float)m, e + ex};
}

// dummest-score.hh:70-76
static inline bool scoreLess(const ES &a, const ES &b) {
    if (a.m == 0.0f) return b.m != 0.0f;
    if (b.m == 0.0f) return false;
    if (a.e != b.e) return a.e < b.e;
    return a.m < b.m;
}

static const int RESCALE_BITS = 64;  // dummest-score.hh:83

static inline int clampRescaleCount(int cum) {  // dummest-score.hh:98-103
    if (cum > 1000000) return 1000000;
    if (cum < -1000000) return -1000000;
    return cum;
}

// dummest-score.hh:107-116 (scalar promote)
template <class F> static inline ES promoteRescaledScore(F v, int cum) {
    if (!(v > 0)) return ES{0.0f, 0};
    cum = clampRescaleCount(cum);
    if (!(v < (F)INFINITY)) return ES{std::nextafterf(1.0f, 0.0f), 1023 + RESCALE_BITS * cum};
    return normalizeScore((double)v, RESCALE_BITS * cum);
}

// promoteRescaledScore with an arbitrary rescale quantum (AUDIT_FORCE_RESCALE
// builds use 8, dummest-score.hh:80-94).
template <class F> static inline ES promoteBits(F v, int cum, int bits) {
    if (!(v > 0)) return ES{0.0f, 0};
    cum = clampRescaleCount(cum);
    if (!(v < (F)INFINITY)) return ES{std::nextafterf(1.0f, 0.0f), 1023 + bits * cum};
    return normalizeScore((double)v, bits * cum);
}

// dummest-batch.hh:146-211 lane-wise promote. For normal inputs the AVX2 path
// equals the scalar path; subnormal inputs flush to zero.
template <class F> static inline ES batchPromote(F v, int cum) {
    if (std::fpclassify(v) == FP_SUBNORMAL) return ES{0.0f, 0};
    return promoteRescaledScore<F>(v, cum);
}

// dummest-score.hh:155-160
static inline ES scaleScore(const ES &a, double f) {
    if (a.m == 0.0f) return ES{0.0f, 0};
    if (!(f > 0.0) || !std::isfinite(f)) return ES{0.0f, 0};
    return normalizeScore((double)a.m * f, a.e);
}

// dummest-score.hh:169-181
static inline ES addScores(ES a, ES b) {
    if (a.m == 0.0f) return b;
    if (b.m == 0.0f) return a;
    if (a.e >= b.e) {
        int d = a.e - b.e;
        if (d > 60) return a;
        return normalizeScore((double)a.m + std::ldexp((double)b.m, -d), a.e);
    }
    int d = b.e - a.e;
    if (d > 60) return b;
    return normalizeScore((double)b.m + std::ldexp((double)a.m, -d), b.e);
}

// dummest-score.hh:184-188
static inline ES mulScores(const ES &a, const ES &b) {
    if (a.m == 0.0f || b.m == 0.0f) return ES{0.0f, 0};
    return normalizeScore((double)a.m * (double)b.m, a.e + b.e);
}

// dummest-score.hh:197-199
static inline ES pow2Score(int e2) { return ES{0.5f, e2 + 1}; }

// dummest-score.hh:202-214
static inline double scoreToLog(const ES &a) {
    if (a.m == 0.0) return -INFINITY;
    static const double LN2 = 0.6931471805599453;
    return std::log(a.m) + (double)a.e * LN2;
}
static inline double scoreToLog2(const ES &a) {
    if (a.m == 0.0) return -INFINITY;
    return std::log2((double)a.m) + (double)a.e;
}

// dummest-batch.hh:42-53
static inline void splitNormalizedDouble(double x, float &mOut, int &eAdd) {
    if (x == 0.0) { mOut = 0.0f; eAdd = 0; return; }
    uint64_t u;
    std::memcpy(&u, &x, 8);
    int ef = (int)((u >> 52) & 0x7FFULL);
    if (ef == 0x7FF) { mOut = (float)x; eAdd = 0; return; }
    if (ef == 0) { mOut = 0.0f; eAdd = 0; return; }
    eAdd = ef - 1022;
    uint64_t v = (u & 0xFFFFFFFFFFFFFULL) | (1022ULL << 52);
    double d;
    std::memcpy(&d, &v, 8);
    mOut = (float)d;
}

// dummest-batch.hh:58-72 (one lane)
static inline ES batchMul(const ES &A, const ES &B, int eFold) {
    const bool nz = (A.m != 0.0f) & (B.m != 0.0f);
    const double p = (double)A.m * (double)B.m;
    const int e = A.e + B.e + eFold;
    const bool lo = (p < 0.5);
    const double ps = lo ? p * 2.0 : p;
    const int es = lo ? e - 1 : e;
    return nz ? ES{(float)ps, es} : ES{0.0f, 0};
}

// dummest-batch.hh:78-108 (one lane)
static inline ES batchAdd(const ES &A, const ES &B) {
    const float am = A.m, bm = B.m;
    const int ae = A.e, be = B.e;
    const bool a0 = (am == 0.0f), b0 = (bm == 0.0f);
    const bool agebe = (ae >= be);
    const float mh = agebe ? am : bm;
    const float ml = agebe ? bm : am;
    const int eh = agebe ? ae : be;
    const int el = agebe ? be : ae;
    int d = eh - el;
    d = d > 60 ? 60 : d;
    double s = (double)mh + (double)ml * (1.0 / (double)(1ULL << (unsigned)d));
    int e = eh;
    const bool hi = (s >= 1.0);
    s = hi ? s * 0.5 : s;
    e = hi ? e + 1 : e;
    const ES arith{(float)s, e};
    return a0 ? B : (b0 ? A : arith);
}

// dummest-batch.hh:114-130 (one lane)
static inline ES batchScale(const ES &A, double fv) {
    const bool a0 = (A.m == 0.0f);
    const bool fok = (fv > 0.0) && std::isfinite(fv);
    float m;
    int ex;
    splitNormalizedDouble(a0 ? 0.0 : (double)A.m * (fok ? fv : 1.0), m, ex);
    const ES arith{m, A.e + ex};
    return (a0 || !fok) ? ES{0.0f, 0} : arith;
}

// dummest-score.hh:224-230 with EXP2_HI = 120 (float) / 1023 (double)
template <class F> static inline F exp2Hi() { return sizeof(F) == 4 ? (F)120 : (F)1023; }
template <class F> static inline F scoreToFloatClamped(const ES &a) {
    if (a.m == 0.0f) return (F)0.0;
    int e = a.e;
    if (e > (int)exp2Hi<F>()) e = (int)exp2Hi<F>();
    return std::ldexp((F)a.m, e);
}
template <class F> static inline F rescaleThresh() { return sizeof(F) == 4 ? (F)1e30 : (F)1e150; }

// dummest-scalerow.hh:26-38
static inline float scoreToRowScale(float m, int e, int E) {
    uint32_t u;
    std::memcpy(&u, &m, 4);
    int d = e - E;
    d = d > 0 ? 0 : (d < -1000 ? -1000 : d);
    int newExp = 126 + d;
    uint32_t mb = (u & 0x807FFFFFu) | ((uint32_t)(newExp > 0 ? newExp : 0) << 23);
    float s;
    std::memcpy(&s, &mb, 4);
    return (m == 0.0f || newExp <= 0) ? 0.0f : s;
}

// dummest-scalerow.hh:43-55
static inline ES scoreFromRowScale(float s, int E) {
    uint32_t u;
    std::memcpy(&u, &s, 4);
    uint32_t ef = (u >> 23) & 0xFFu;
    const bool live = (s > 0.0f) & (ef != 0);
    uint32_t mb = (u & 0x807FFFFFu) | (126u << 23);
    float m;
    std::memcpy(&m, &mb, 4);
    ES o{m, (int)ef - 126 + E};
    return live ? o : ES{0.0f, 0};
}

// ----------------------------------------------------------------------------
// Options (dummest-core.hh:115-124): stored as Float, the background branch
// rates are Float sums.
// ----------------------------------------------------------------------------
template <class F> struct Rates {
    F insert1, insert2, delete1, delete2, stop, bgstop;
    F bgfs1, bgfs2;  // BACKGROUND_FRAMESHIFT_RATE(_2)
    void init(const double *o) {
        insert1 = (F)o[0]; insert2 = (F)o[1]; delete1 = (F)o[2]; delete2 = (F)o[3];
        stop = (F)o[4]; bgstop = (F)o[5];
        bgfs1 = insert1 + delete2;  // dummest-core.hh:128-131
        bgfs2 = insert2 + delete1;
    }
};

static const char *kAlphabet = "ACDEFGHIKLMNPQRSTVWYUO?*";
static const int kWidth = 29;      // 20 letters + nonLetterWidth 9
static const int kZeroIdx = 24;    // dummest-dp.hh:92-94

static int codonCount(char aa) {  // dummest-sequence.hh:103-124 (sizes)
    switch (aa) {
    case 'A': case 'G': case 'P': case 'T': case 'V': return 4;
    case 'R': case 'L': case 'S': return 6;
    case 'I': return 3;
    case 'M': case 'W': return 1;
    case '*': return 3;
    default: return 2;  // N D C U Q E H K O F Y
    }
}

// ----------------------------------------------------------------------------
// Profile transform.  dummest-profile.hh:30-62 (means), 105-233 (finalize),
// without tantan (-m 0): every position counts towards the background.
// values: (M+1) rows x 29, as read (reader order dummest-profile.hh:264-314).
// params out: (M+2) rows x 11 = alpha'[3] beta'[3] delta'[3] eps' enter.
// bg out: 29 entries (bg_probs, dummest-profile.hh:135-148).
// ----------------------------------------------------------------------------
template <class F>
static double myMean(const F *values, int length, int step, int meanType, const float *tantan = nullptr,
                     F thr = 0) {
    double mean = 0;
    int n = 0;
    std::vector<F> med;
    for (int i = 0; i < length; ++i) {
        if (tantan && tantan[i] >= thr) continue;  // dummest-profile.hh:41-43
        double v = values[i * step];
        if (meanType == 'G') mean += std::log(std::max(v, 1e-6));
        if (meanType == 'A') mean += v;
        if (meanType == 'M') med.push_back((F)v);
        ++n;
    }
    if (meanType == 'G') return std::exp(mean / n);
    if (meanType == 'A') return mean / n;
    std::sort(med.begin(), med.end());
    return med[n / 2];
}

// filterLetterProbabilities, dummest-profile.hh:64-103 (with normalize and
// meanOfLogs, :20-35, and mean, dummest-core.hh:335-340), for the 20 letters
// of `length` rows `step` apart. Doubles, results narrowed to F.
template <class F>
static void filterLetterProbabilities(F *letterProbs, int length, int step, double stdDev, bool keepNonvaryingTerm) {
    const double sqrt2pi = 2.5066282746310005;
    const double inv2var = 0.5 / (stdDev * stdDev);
    const int gaussianLimit = (int)std::ceil(stdDev * 8);
    const int alphabetSize = 20;
    std::vector<double> meanLogProbs(length), values(length);
    for (int i = 0; i < length; ++i) {
        double m = 1;
        for (int k = 0; k < alphabetSize; ++k) m *= letterProbs[i * step + k];
        meanLogProbs[i] = std::log(m) / alphabetSize;
    }
    for (int k = 0; k < alphabetSize; ++k) {
        for (int i = 0; i < length; ++i) values[i] = std::log((double)letterProbs[i * step + k]) - meanLogProbs[i];
        double addItBack = 0.0;
        if (keepNonvaryingTerm) {
            double t = 0;
            for (int i = 0; i < length; ++i) t += values[i];
            addItBack = t / length;
        }
        for (int i = 0; i < length; ++i) {
            double sum = 0;
            for (int j = -gaussianLimit; j <= gaussianLimit; ++j) {
                int x = (i + j) % length;  // circular profile
                if (x < 0) x += length;
                sum += values[x] * std::exp(-1.0 * j * j * inv2var);
            }
            sum /= stdDev * sqrt2pi;
            letterProbs[i * step + k] = (F)std::exp(values[i] - sum + addItBack);
        }
    }
    for (int i = 0; i < length; ++i) {
        F *x = letterProbs + i * step;
        double total = 0;
        for (int k = 0; k < alphabetSize; ++k) total += x[k];
        for (int k = 0; k < alphabetSize; ++k) x[k] /= total;
    }
}

template <class F>
static int finalizeProfile(int M, F *values, F *params, F *bgOut, const Rates<F> &R, int bgType,
                           const float *tantan = nullptr, F thr = 0, double filterSd = 0, bool keepNonvarying = false) {
    // tantan: repeat probabilities of the consensus (profile masking, -m 1),
    // computed after the final epsilon (dummest-profile.hh:116-126); positions
    // at or above thr leave the background means and emit like the background
    // (dummest-profile.hh:209-211, 225-226).
    const int W = kWidth;
    F *end = values + W * M;
    if (end[3] <= 0) end[3] = (F)myMean<F>(values + W + 3, M - 1, W, 'G');
    // dummest-profile.hh:119-126: the filter, when on, replaces profile masking
    if (filterSd > 0) {
        filterLetterProbabilities<F>(values + 4, M, W, filterSd, keepNonvarying);
        tantan = nullptr;
    }
    double sumOfMeans = 0;
    for (int k = 4; k < 24; ++k) {
        double mean = myMean<F>(values + k, M, W, bgType, tantan, thr);
        end[k] = (F)mean;
        sumOfMeans += mean;
    }
    for (int k = 0; k < 4; ++k) bgOut[k] = 0;
    for (int k = 4; k < 24; ++k) {
        end[k] = (F)((double)end[k] / sumOfMeans);
        end[k] *= (1 - R.bgstop) / (F)codonCount(kAlphabet[k - 4]);
        bgOut[k] = end[k];
    }
    bgOut[24] = 0;
    bgOut[25] = 0;
    bgOut[26] = (F)(1.0 / 64.0);
    bgOut[27] = (F)(R.bgstop / 3.0);
    bgOut[28] = 0;
    std::memset(params, 0, sizeof(F) * 11 * (M + 2));
    const F fsAll = 1 - R.insert1 - R.insert2 - R.delete1 - R.delete2;  // Float arithmetic
    for (int i = 0;; ++i) {
        F *pp = params + 11 * i;
        F *probs = values + i * W;
        double alpha = probs[0] * fsAll;
        double beta = probs[1];
        double alphaFS1 = R.insert1, alphaFS2 = R.insert2;
        pp[0] = (F)(alpha * (1 - beta));
        pp[1] = (F)alphaFS1;
        pp[2] = (F)alphaFS2;
        pp[3] = (F)beta;
        pp[4] = 0;
        pp[5] = 0;
        double delta = probs[2] * fsAll;
        double epsilon = probs[3];
        if (i == M) break;
        double epsilon1 = probs[W + 3];
        double deltaFS1 = R.delete1, deltaFS2 = R.delete2;
        pp[6] = (F)(delta * (1 - epsilon1));
        pp[7] = (F)deltaFS1;
        pp[8] = (F)deltaFS2;
        pp[9] = (F)(epsilon * (1 - epsilon1) / (1 - epsilon));
        pp[10] = (F)(1 - alpha - alphaFS1 - alphaFS2 - delta - deltaFS1 - deltaFS2);
        if (epsilon >= 1) return 0;
        probs[2] = (F)delta;
        probs[3] = (F)epsilon;
        const bool masked = tantan && tantan[i] >= thr;
        for (int k = 4; k < 24; ++k) {
            if (masked) {
                probs[k] = end[k];
                continue;
            }
            double p = probs[k];
            int cc = codonCount(kAlphabet[k - 4]);
            probs[k] = (F)((1 - R.stop) * p / cc);
        }
        probs[24] = probs[5];
        probs[25] = probs[12];
        probs[26] = (F)(1.0 / 64.0);
        probs[27] = (F)(R.stop / 3.0);
        probs[28] = 0;
    }
    return 1;
}

// ----------------------------------------------------------------------------
// tantan repeat probabilities, as DUMMEST calls them: calcTantanProbabilities
// (tantan-wrapper.hh:13-64) -> tantan::getProbabilities -> Tantan::calcRepeatProbs
// (tantan/tantan.cc:40-130 set-up, 222-330 the gap-free forward and backward
// steps, 332-392 rescaling and the driver) with firstGapProb = otherGapProb = 0.
// The reference is built with -O3 -march=x86-64-v3, where the sums over repeat
// offsets run in four interleaved partial sums (the AVX2 lanes of
// tantan/mcf_simd.hh:74) reduced as (s0+s2)+(s1+s3) (mcf_simd.hh:175-179), and
// GCC fuses these multiply-adds; the restatement spells both out (pinned
// bit-exactly against oracle/_ref, the reference's own tantan built from its
// sources):
//   forward   fp = fma(b, b2f[i], f * f2f0) * r        bg = fma(b, b2b, sum * f2b)
//   backward  f = fp * r   s = fma(b2f[i], f, s)   fp = fma(f, f2f0, f2b * bg)
//             bg = fma(b2b, bg, s)
//   total     z = fma(sum(fp), f2b, bg * b2b)
// ----------------------------------------------------------------------------
static const signed char kBlosum62[20][20] = {
    // A   C   D   E   F   G   H   I   K   L   M   N   P   Q   R   S   T   V   W   Y
    {4, 0, -2, -1, -2, 0, -2, -1, -1, -1, -1, -2, -1, -1, -1, 1, 0, 0, -3, -2},      // A
    {0, 9, -3, -4, -2, -3, -3, -1, -3, -1, -1, -3, -3, -3, -3, -1, -1, -1, -2, -2},  // C
    {-2, -3, 6, 2, -3, -1, -1, -3, -1, -4, -3, 1, -1, 0, -2, 0, -1, -3, -4, -3},     // D
    {-1, -4, 2, 5, -3, -2, 0, -3, 1, -3, -2, 0, -1, 2, 0, 0, -1, -2, -3, -2},        // E
    {-2, -2, -3, -3, 6, -3, -1, 0, -3, 0, 0, -3, -4, -3, -3, -2, -2, -1, 1, 3},      // F
    {0, -3, -1, -2, -3, 6, -2, -4, -2, -4, -3, 0, -2, -2, -2, 0, -2, -3, -2, -3},    // G
    {-2, -3, -1, 0, -1, -2, 8, -3, -1, -3, -2, 1, -2, 0, 0, -1, -2, -3, -2, 2},      // H
    {-1, -1, -3, -3, 0, -4, -3, 4, -3, 2, 1, -3, -3, -3, -3, -2, -1, 3, -3, -1},     // I
    {-1, -3, -1, 1, -3, -2, -1, -3, 5, -2, -1, 0, -1, 1, 2, 0, -1, -2, -3, -2},      // K
    {-1, -1, -4, -3, 0, -4, -3, 2, -2, 4, 2, -3, -3, -2, -2, -2, -1, 1, -2, -1},     // L
    {-1, -1, -3, -2, 0, -3, -2, 1, -1, 2, 5, -2, -2, 0, -1, -1, -1, 1, -1, -1},      // M
    {-2, -3, 1, 0, -3, 0, 1, -3, 0, -3, -2, 6, -2, 0, 0, 1, 0, -3, -4, -2},          // N
    {-1, -3, -1, -1, -4, -2, -2, -3, -1, -3, -2, -2, 7, -1, -2, -1, -1, -2, -4, -3}, // P
    {-1, -3, 0, 2, -3, -2, 0, -3, 1, -2, 0, 0, -1, 5, 1, 0, -1, -2, -2, -1},         // Q
    {-1, -3, -2, 0, -3, -2, 0, -3, 2, -2, -1, 0, -2, 1, 5, -1, -1, -3, -3, -2},      // R
    {1, -1, 0, 0, -2, 0, -1, -2, 0, -2, -1, 1, -1, 0, -1, 4, 1, -2, -3, -2},         // S
    {0, -1, -1, -1, -2, -2, -2, -1, -1, -1, -1, 0, -1, -1, -1, 1, 5, 0, -2, -2},     // T
    {0, -1, -3, -2, -1, -3, -3, 3, -2, 1, 1, -3, -2, -2, -3, -2, 0, 4, -3, -1},      // V
    {-3, -2, -4, -3, 1, -2, -2, -3, -3, -2, -1, -4, -4, -2, -3, -3, -2, -3, 11, 2},  // W
    {-2, -2, -3, -2, 3, -3, 2, -1, -2, -1, -1, -2, -3, -1, -2, -2, -2, -1, 2, 7},    // Y
};

static void tantanProbs(const uint8_t *seq, int n, bool isProtein, float *out) {
    // likelihood ratios (tantan-wrapper.hh:43-56): BLOSUM62 at lambda
    // 0.32403216734804408 (letters >= 20 score -4), or +1/-1 for DNA
    double ratio[22][22];
    for (int i = 0; i < 22; ++i)
        for (int j = 0; j < 22; ++j) {
            if (isProtein) {
                const int sc = (i < 20 && j < 20) ? kBlosum62[i][j] : -4;
                ratio[i][j] = std::exp(0.32403216734804408 * sc);
            } else {
                ratio[i][j] = (i < 4 && i == j) ? 3.0 : 1.0 / 3.0;
            }
        }
    const int K = isProtein ? 50 : 100;
    const double repeatProb = 0.005, repeatEndProb = 0.05, decay = 0.9;
    const double b2b = 1 - repeatProb, f2b = repeatEndProb, f2f0 = 1 - repeatEndProb;
    auto firstOffsetProb = [&](double pm) {  // tantan.cc:25-30
        if (pm < 1 || pm > 1) return (1 - pm) / (1 - std::pow(pm, K));
        return 1.0 / K;
    };
    const double b2fFirst = repeatProb * firstOffsetProb(decay);
    std::vector<double> b2f(K), fp(K, 0.0), scales(n / 16);
    double p = b2fFirst;
    for (int i = 0; i < K; ++i) {
        b2f[i] = p;
        p *= decay;
    }
    // forward (tantan.cc:283-330, 375-381)
    double bg = 1.0;
    for (int pos = 0; pos < n; ++pos) {
        const double b = bg;
        const double *lr = ratio[seq[pos]];
        const int m = std::min(pos, K);
        double s4[4] = {0, 0, 0, 0};
        int i = 0;
        for (; i + 4 <= m; i += 4)
            for (int l = 0; l < 4; ++l) {
                const double f = fp[i + l];
                s4[l] += f;
                fp[i + l] = std::fma(b, b2f[i + l], f * f2f0) * lr[seq[pos - i - l - 1]];
            }
        double sum = (s4[0] + s4[2]) + (s4[1] + s4[3]);
        for (; i < m; ++i) {
            const double f = fp[i];
            sum += f;
            fp[i] = std::fma(b, b2f[i], f * f2f0) * lr[seq[pos - i - 1]];
        }
        bg = std::fma(b, b2b, sum * f2b);
        if (pos % 16 == 15) {  // tantan.cc:364-371
            const double scale = 1 / bg;
            scales[pos / 16] = scale;
            bg *= scale;
            for (double &f : fp) f *= scale;
        }
        out[pos] = (float)bg;
    }
    double all = 0.0;
    for (double f : fp) all += f;
    const double z = std::fma(all, f2b, bg * b2b);  // forwardTotal, tantan.cc:135-141
    // backward (tantan.cc:147-151, 332-361, 384-397)
    bg = b2b;
    std::fill(fp.begin(), fp.end(), f2b);
    for (int pos = n - 1; pos >= 0; --pos) {
        const double nonRepeat = (double)out[pos] * bg / z;
        out[pos] = 1 - (float)nonRepeat;
        if (pos % 16 == 15) {
            const double scale = scales[pos / 16];
            bg *= scale;
            for (double &f : fp) f *= scale;
        }
        const double toBg = f2b * bg;
        const double *lr = ratio[seq[pos]];
        const int m = std::min(pos, K);
        double s4[4] = {0, 0, 0, 0};
        int i = 0;
        for (; i + 4 <= m; i += 4)
            for (int l = 0; l < 4; ++l) {
                const double f = fp[i + l] * lr[seq[pos - i - l - 1]];
                s4[l] = std::fma(b2f[i + l], f, s4[l]);
                fp[i + l] = std::fma(f, f2f0, toBg);
            }
        double tf = (s4[0] + s4[2]) + (s4[1] + s4[3]);
        for (; i < m; ++i) {
            const double f = fp[i] * lr[seq[pos - i - 1]];
            tf = std::fma(b2f[i], f, tf);
            fp[i] = std::fma(f, f2f0, toBg);
        }
        bg = std::fma(b2b, bg, tf);
    }
}

// DNA masking (makeMaskedSequence, dummest-sequence.hh:206-236): tantan runs on
// A,C,G,T = 0..3 (every other letter counts as A); positions whose repeat
// probability reaches thr become 'N' (the reference's mask value decodes like
// any non-ACGT letter: the codons over it are '?').
static void maskDna(const char *dna, int n, double thr, char *outDna, float *probs) {
    std::vector<uint8_t> x(n);
    for (int i = 0; i < n; ++i) {
        switch (std::toupper((unsigned char)dna[i])) {
        case 'C': x[i] = 1; break;
        case 'G': x[i] = 2; break;
        case 'T': x[i] = 3; break;
        default: x[i] = 0;
        }
    }
    tantanProbs(x.data(), n, false, probs);
    for (int i = 0; i < n; ++i) outDna[i] = (probs[i] < (float)thr) ? dna[i] : 'N';
}

// ----------------------------------------------------------------------------
// Sequence decode. dummest-sequence.hh:134-204, charToNumber dummest.cc:445-456.
// dna: raw letters. codes out: n bytes (last two are 0, the init value).
// ----------------------------------------------------------------------------
static char translateCodon(const char *c) {
    static const char *bases = "TCAG";
    // Standard code in TCAG order.
    static const char *aas = "FFLLSSSSYY**CC*WLLLLPPPPHHQQRRRRIIIMTTTTNNKKSSRRVVVVAAAADDEEGGGG";
    int idx = 0;
    for (int k = 0; k < 3; ++k) {
        char u = (char)std::toupper((unsigned char)c[k]);
        const char *p = std::strchr(bases, u);
        if (!p || !*p) return '?';
        idx = idx * 4 + (int)(p - bases);
    }
    return aas[idx];
}

static void decodeDna(const char *dna, int n, uint8_t *codes) {
    // dummest.cc:445-456 maps A/C/G/T to protein-alphabet numbers and IUPAC
    // codes to 22 ('?'); decodeSequence translates through an ACGT-only codon
    // table, so every other letter behaves like '?' in a codon.
    std::vector<char> dec(n);
    for (int i = 0; i < n; ++i) {
        char u = (char)std::toupper((unsigned char)dna[i]);
        dec[i] = (u == 'A' || u == 'C' || u == 'G' || u == 'T') ? u : '?';
    }
    for (int i = 0; i < n; ++i) codes[i] = 0;  // (uint8_t)INT_MIN, dummest-sequence.hh:188
    for (int i = 0; i + 2 < n; ++i) {
        char aa;
        if (dec[i] == '?' || dec[i + 1] == '?' || dec[i + 2] == '?') aa = '?';
        else aa = translateCodon(&dec[i]);
        codes[i] = (uint8_t)(std::strchr(kAlphabet, aa) - kAlphabet);
    }
}

// ----------------------------------------------------------------------------
// Null model for one sequence (lane).  dummest-dp-backward.hh:6-188.
// code(j) reads the lane-transposed decoded array with zero_idx padding
// (dummest-dp.hh:67-77): positions outside [0, n) read zero_idx.
// ----------------------------------------------------------------------------
template <class F> struct Null {
    int n = 0, W = 0;
    std::vector<ES> prefix, suffix;   // W + 4 entries
    ES perBase, perBaseInv, s1, s2, sC;
    std::vector<F> prefixSeed, suffixSeed;  // W + 4 entries (packed)
    F s1f = 0, s2f = 0, sCf = 0;
};

template <class F>
static void buildNull(Null<F> &N, const F *bgPerBase, const uint8_t *codes, int n, int W, const Rates<F> &R) {
    N.n = n;
    N.W = W;
    auto code = [&](int j) -> int { return (j >= 0 && j < n) ? codes[j] : kZeroIdx; };
    const double single = R.bgfs1, dbl = R.bgfs2;
    const double codonKeep = 1 - single - dbl;
    const double singleEmit = single * 0.25;
    const double doubleEmit = dbl * 0.0625;
    const double frame = 1.0 / 3.0;
    N.prefix.assign(W + 4, ES{});
    N.suffix.assign(W + 4, ES{});
    // buildNullSuffix, dummest-dp-backward.hh:6-55
    {
        float tm;
        int te;
        splitNormalizedDouble(0.5 * codonKeep, tm, te);
        N.suffix[n] = pow2Score(0);
        for (int i = W - 1; i >= 0; i--) {
            F bgc = bgPerBase[code(i)];
            double codonScale = (double)bgc * codonKeep;
            ES singleIn = (i < n) ? N.suffix[i + 1] : ES{0.0f, 0};
            ES doubleIn = (i + 1 < n) ? N.suffix[i + 2] : ((i + 1 == n) ? pow2Score(2) : ES{0.0f, 0});
            ES codonRaw = batchScale(N.suffix[i + 3], codonScale);
            ES singleB = batchScale(singleIn, singleEmit);
            ES doubleB = batchScale(doubleIn, doubleEmit);
            ES codonB = (i + 2 < n) ? codonRaw : ES{tm, te - 2 * (n - i) + 1};
            ES out = batchAdd(codonB, batchAdd(singleB, doubleB));
            if (i >= n) out = (i == n) ? pow2Score(0) : ES{0.0f, 0};
            N.suffix[i] = out;
        }
    }
    // buildNullPrefixFirstCodon, dummest-dp-backward.hh:58-100
    for (int i = 0; i < W; i++) {
        F bgc = bgPerBase[code(i - 2)];
        double codonScale = (double)bgc * codonKeep;
        ES ps = (i > 0) ? N.prefix[i - 1] : ES{0.0f, 0};
        ES pd = (i >= 2) ? N.prefix[i - 2] : ES{0.0f, 0};
        ES codonB;
        if (i >= 3) codonB = batchScale(N.prefix[i - 3], codonScale);
        else if (i == 2) codonB = scaleScore(normalizeScore((double)bgc, 0), frame);
        else if (i == 1) codonB = scaleScore(pow2Score(-4), frame);
        else codonB = scaleScore(pow2Score(-2), frame);
        ES sb = batchScale(ps, singleEmit), db = batchScale(pd, doubleEmit);
        ES out = batchAdd(codonB, batchAdd(sb, db));
        N.prefix[i] = (i < n) ? out : ES{0.0f, 0};
    }
    // buildNullTotalAndScales, dummest-dp-backward.hh:103-129
    {
        const double keepSingle = 1.0 - single;
        const double keepBoth = keepSingle - dbl;
        ES ts = scaleScore(N.prefix[n - 2], keepSingle * 0.25);
        ES td = scaleScore(N.prefix[n - 3], keepBoth * 0.0625);
        ES total = addScores(addScores(N.prefix[n - 1], ts), td);
        F nullLog = (F)scoreToLog2(total);
        double perBaseLog = -(double)nullLog / (double)(F)n;
        if (!(perBaseLog < (double)exp2Hi<F>())) perBaseLog = (double)exp2Hi<F>();
        double raw = std::exp2(perBaseLog);
        N.perBase = normalizeScore(raw, 0);
        N.perBaseInv = normalizeScore(1.0 / raw, 0);
        ES quarter = normalizeScore(0.25, 0), sixteenth = normalizeScore(0.0625, 0);
        N.s1 = mulScores(N.perBase, quarter);
        N.s2 = mulScores(mulScores(N.perBase, N.perBase), sixteenth);
        N.sC = mulScores(mulScores(N.perBase, N.perBase), N.perBase);
    }
    // tail zeroing, dummest-dp-backward.hh:229-242
    for (int j = W; j < W + 4; j++)
        if (j != n) { N.prefix[j] = ES{}; N.suffix[j] = ES{}; }
    // buildNullSeedModels, dummest-dp-backward.hh:133-165
    N.prefixSeed.assign(W + 4, (F)0);
    N.suffixSeed.assign(W + 4, (F)0);
    {
        ES running = pow2Score(0);
        for (int t = 0; t < n - 1; t++) running = batchMul(running, N.perBase, 0);
        for (int j = 0; j < W; j++) {
            N.suffixSeed[j] = scoreToFloatClamped<F>(batchMul(N.suffix[j + 1], running, 0));
            running = batchMul(running, N.perBaseInv, 0);
        }
        running = N.perBase;
        for (int j = 0; j < W; j++) {
            N.prefixSeed[j] = scoreToFloatClamped<F>(batchMul(N.prefix[j], running, 0));
            running = batchMul(running, N.perBase, 0);
        }
    }
    N.s1f = scoreToFloatClamped<F>(N.s1);
    N.s2f = scoreToFloatClamped<F>(N.s2);
    N.sCf = scoreToFloatClamped<F>(N.sC);
}

// ----------------------------------------------------------------------------
// Profile view used by the DP.
// ----------------------------------------------------------------------------
template <class F> struct Prof {
    int M;
    const F *values;  // (M+1) x 29, finalized
    const F *params;  // (M+2) x 11
    const F *bg;      // 29
    const F *emis(int i) const { return values + i * kWidth + 4; }
    const F *par(int i) const { return params + 11 * i; }
};

// ----------------------------------------------------------------------------
// Score-only filter DP.  dummest-dp-backward.hh:300-438 (one lane).
// ----------------------------------------------------------------------------
template <class F>
static F filterScore(const Prof<F> &P, const uint8_t *codes, int n, const Rates<F> &R, Null<F> *nullOut = nullptr) {
    Null<F> Nloc;
    Null<F> &N = nullOut ? *nullOut : Nloc;
    const int W = n;
    buildNull<F>(N, P.bg + 4, codes, n, W, R);
    auto code = [&](int j) -> int { return (j >= 0 && j < n) ? codes[j] : kZeroIdx; };
    const F s1 = N.s1f, s2 = N.s2f, sC = N.sCf;
    std::vector<F> row[2];
    row[0].assign(W + 4, (F)0);
    row[1].assign(W + 4, (F)0);
    std::vector<F> Ynext(W + 4, (F)0), Ycur(W + 4, (F)0);
    std::vector<F> bgc(W + 8);
    for (int j = -4; j < W + 4; j++) bgc[j + 4] = P.bg[4 + code(j)];
    F best = 0;
    const F scale = 1.0;
    for (int i = P.M; i >= 0; i--) {
        const F *pr = P.par(i);
        const F *em = P.emis(i);
        const F enterC = pr[10] * (g_mutation == 3 ? s2 : sC), d0 = pr[6], d1 = pr[7] * s2, d2 = pr[8] * s1;
        const F a0 = pr[0] * sC, a1 = pr[1] * s1, a2 = pr[2] * s2, b0 = pr[3] * sC;
        const F eps = pr[9], rsc = scale;
        F Z[4] = {0, 0, 0, 0};
        std::vector<F> &nx = row[(i + 1) & 1];
        std::vector<F> &cu = row[i & 1];
        F n3 = nx[W + 3], n2 = nx[W + 2], n1 = nx[W + 1];
        F c2 = cu[W + 2], c1 = cu[W + 1];
        for (int j = W - 1; j >= 0; j--) {
            int r0 = j & 3, r3 = (j + 3) & 3;
            F e = em[code(j + 1)];
            F g = bgc[j + 1 + 4];
            F nAt = nx[j];
            F w = n3 * e * enterC + Ynext[j] * d0 + n2 * d1 + n1 * d2 + Z[r3] * g * a0 + c2 * a2 + c1 * a1 +
                  N.suffixSeed[j] * rsc;
            w = w * (F)(j < n ? 1 : 0);
            F mid = w * N.prefixSeed[j];
            best = std::max(best, mid);  // Kokkos::max(best, mid)
            cu[j] = w;
            Ycur[j] = std::fma(eps, Ynext[j], w) * (F)(j < n ? 1 : 0);
            F zf = Z[r3] * g;
            Z[r0] = std::fma(b0, zf, w);
            n3 = n2; n2 = n1; n1 = nAt;
            c2 = c1; c1 = w;
        }
        std::swap(Ycur, Ynext);
    }
    return best;
}

// ----------------------------------------------------------------------------
// Full search DP for one lane: backwardPass + forwardPass + traceback.
// dummest-dp-backward.hh:191-297,444-615; dummest-dp-forward.hh:9-547;
// dummest-dp.hh:137-268; dummest-traceback.hh:6-118.
// ----------------------------------------------------------------------------
struct Seg { int start1, start2, length; };
struct Sim {
    ES pair;
    double logProbRatio;
    int anchor1, anchor2;
    std::vector<Seg> alignment;
    bool operator<(const Sim &o) const { return scoreLess(pair, o.pair); }
    bool operator>(const Sim &o) const { return scoreLess(o.pair, pair); }
};

struct ExpMat {  // dummest-padded.hh:191-222 for one lane
    int pre = 0, cols = 0;
    std::vector<ES> d;
    void assign(int nrows, int ncols, int prePad) { pre = prePad; cols = ncols; d.assign((size_t)nrows * (ncols + pre), ES{}); }
    ES &at(int i, int j) { return d[(size_t)i * (cols + pre) + (j + pre)]; }
    ES get(int i, int j) const { return d[(size_t)i * (cols + pre) + (j + pre)]; }
};

template <class F> struct Work {
    Null<F> N;
    std::vector<F> W1;  // (M+2) x (W+4)
    std::vector<int> bcnt, fcnt;
    std::vector<int> rowMaxExp;
    std::vector<F> revSide, fwdSide;  // revSide has 3 front pads
    ExpMat X, pre, suf;
    std::vector<ES> revEv, fwdEv;  // revEv has 3 front pads
    std::vector<ES> colBest;
    std::vector<int> colRow;
    ES globalBest, bwdBest;
    int W = 0;
    bool capture = false;
    std::vector<ES> w0es, W1es;  // (M+1) x n when capture is set
    F *w1(int i, int j) { return &W1[(size_t)i * (W + 4) + j]; }
};

// Flank totals, dummest-dp.hh:137-268 (one hit lane).
template <class F>
static void buildFlankEVs(Work<F> &Wk, const Prof<F> &P, const uint8_t *codes, int n, const Rates<F> &R,
                          int rescaleBits) {
    const int dpWidth = Wk.W;
    Null<F> &N = Wk.N;
    auto code = [&](int j) -> int { return (j >= 0 && j < n) ? codes[j] : kZeroIdx; };
    const double singleShift = R.bgfs1, doubleShift = R.bgfs2;
    const double codonKeep = 1.0 - singleShift - doubleShift;
    const double frame = 1.0 / 3.0, third = 1.0 / 3.0;
    const ES zero{0.0f, 0};
    const ES aS = normalizeScore((double)N.s1f, 0);
    const ES aD = normalizeScore((double)N.s2f, 0);
    const ES aC = normalizeScore((double)N.sCf, 0);
    const ES bS = scaleScore(aS, 4.0), bD = scaleScore(aD, 16.0);
    const int L = n;
    auto bgcAt = [&](int pos) { return normalizeScore((double)P.bg[4 + code(pos)], 0); };
    const int fcM = Wk.fcnt[P.M], bc0 = Wk.bcnt[0];
    std::vector<ES> ePM(dpWidth, zero);
    for (int p = 0; p < dpWidth; p++) ePM[p] = normalizeScore((double)N.prefixSeed[p], 0);
    std::vector<ES> eS(dpWidth, zero), eD(dpWidth, zero), eC(dpWidth, zero);
    if (dpWidth > 0) eS[0] = ePM[0];
    if (dpWidth > 1) eD[1] = scaleScore(aD, frame);
    if (dpWidth > 2) eC[2] = mulScores(scaleScore(aC, frame), bgcAt(0));
    for (int p = 1; p < dpWidth; p++) eS[p] = mulScores(scaleScore(ePM[p - 1], singleShift), aS);
    for (int p = 2; p < dpWidth; p++) eD[p] = mulScores(scaleScore(ePM[p - 2], doubleShift), aD);
    for (int p = 3; p < dpWidth; p++)
        eC[p] = mulScores(scaleScore(mulScores(ePM[p - 3], aC), codonKeep), bgcAt(p - 2));
    std::vector<ES> mid((size_t)dpWidth + 3, zero);
    {
        const int last = std::min(dpWidth, L);
        for (int p = last - 1; p >= 0; p--) {
            ES c = mulScores(scaleScore(mid[p + 1], singleShift), aS);
            c = addScores(c, mulScores(scaleScore(mid[p + 2], doubleShift), aD));
            c = addScores(c, mulScores(scaleScore(mulScores(mid[p + 3], aC), codonKeep), bgcAt(p + 1)));
            mid[p] = addScores(promoteBits<F>(Wk.revSide[p + 3], bc0, rescaleBits), c);
        }
    }
    std::vector<ES> sr((size_t)dpWidth + 4, zero);
    if (L >= 0) sr[L] = pow2Score(0);
    for (int p = L - 1; p >= 0; p--) {
        ES st = mulScores(scaleScore(sr[p + 1], singleShift), aS);
        ES dt = (p + 1 <= L - 1) ? mulScores(scaleScore(sr[p + 2], doubleShift), aD) : scaleScore(bS, doubleShift * 0.25);
        ES ct;
        if (p + 2 <= L - 1) ct = mulScores(scaleScore(mulScores(sr[p + 3], aC), codonKeep), bgcAt(p));
        else if (p + 1 <= L - 1) ct = scaleScore(bD, codonKeep * 0.0625);
        else ct = scaleScore(bS, codonKeep * 0.25);
        sr[p] = addScores(st, addScores(dt, ct));
    }
    std::vector<ES> fa(dpWidth, zero);
    for (int p = 0; p < dpWidth; p++) {
        ES a = promoteBits<F>(Wk.fwdSide[p], fcM, rescaleBits);
        if (p - 1 >= 0) a = addScores(a, mulScores(scaleScore(fa[p - 1], singleShift), aS));
        if (p - 2 >= 0) a = addScores(a, mulScores(scaleScore(fa[p - 2], doubleShift), aD));
        if (p - 3 >= 0) a = addScores(a, mulScores(scaleScore(mulScores(fa[p - 3], aC), codonKeep), bgcAt(p - 2)));
        fa[p] = a;
    }
    std::vector<ES> uni(dpWidth, zero);
    for (int p = 0; p < dpWidth; p++) {
        ES we = addScores(eS[p], addScores(scaleScore(eD[p], 2.0), scaleScore(eC[p], 3.0)));
        ES tot = mulScores(mid[p], we);
        if (p >= 1 && p < L) {
            ES x = mulScores(scaleScore(sr[p + 1], singleShift), aS);
            tot = addScores(tot, mulScores(fa[p - 1], x));
        }
        if (p >= 2 && p < L) {
            ES x = mulScores(scaleScore(sr[p + 1], doubleShift), aD);
            tot = addScores(tot, scaleScore(mulScores(fa[p - 2], x), 2.0));
        }
        if (p >= 3 && p < L) {
            ES x = mulScores(scaleScore(mulScores(sr[p + 1], aC), codonKeep), bgcAt(p - 2));
            tot = addScores(tot, scaleScore(mulScores(fa[p - 3], x), 3.0));
        }
        if (p == L - 1) {
            ES td = scaleScore(bS, doubleShift * 0.25);
            tot = addScores(tot, scaleScore(mulScores(fa[p - 1], td), 2.0));
            ES tc1 = scaleScore(bS, codonKeep * 0.25);
            tot = addScores(tot, scaleScore(mulScores(fa[p - 1], tc1), 3.0));
            if (p - 2 >= 0) {
                ES tc2 = scaleScore(bD, codonKeep * 0.0625);
                tot = addScores(tot, scaleScore(mulScores(fa[p - 2], tc2), 3.0));
            }
        }
        uni[p] = scaleScore(tot, third);
    }
    // entryFlankSums = reverseSideEv (3 front pads), exitFlankSums = forwardSideEv.
    Wk.revEv.assign(dpWidth + 4 + 3, zero);
    Wk.fwdEv.assign(dpWidth + 4, zero);
    ES cs = zero;
    for (int p = 0; p < dpWidth; p++) { cs = addScores(cs, uni[p]); Wk.revEv[p + 3] = cs; }
    cs = zero;
    for (int p = dpWidth - 1; p >= 0; p--) { cs = addScores(cs, uni[p]); Wk.fwdEv[p] = cs; }
}

// dummest-traceback.hh:6-32
static void addForwardMatch(std::vector<Seg> &al, int p1, int p2) {
    if (!al.empty()) {
        Seg &x = al.back();
        if (x.start1 + x.length == p1 && x.start2 + x.length == p2) ++x.length;
        if (x.start1 + x.length > p1 || x.start2 + x.length > p2) return;
    }
    al.push_back(Seg{p1, p2, 1});
}
static void addReverseMatch(std::vector<Seg> &al, int p1, int p2) {
    if (!al.empty()) {
        Seg &x = al.back();
        if (x.start1 - 1 == p1 && x.start2 - 1 == p2) { --x.start1; --x.start2; ++x.length; }
        if (x.start1 <= p1 || x.start2 <= p2) return;
    }
    al.push_back(Seg{p1, p2, 1});
}

template <class F>
static void forwardAlignment(Work<F> &Wk, int M, std::vector<Seg> &al, int iBeg, int jBeg) {
    // dummest-traceback.hh:35-66
    const int cols = Wk.suf.cols;
    const int bound = cols - 4;
    int i = iBeg, p = jBeg;
    while (i <= M && p < bound) {
        ES best{0.0f, 0};
        int bi = INT_MAX, bj = INT_MAX;
        bool em = false;
        auto consider = [&](const ES &c, int ni, int nj, bool e) {
            if (scoreLess(best, c)) { best = c; bi = ni; bj = nj; em = e; }
        };
        if (p + 3 < cols) consider(addScores(Wk.X.get(i, p + 3), Wk.suf.get(i + 1, p + 3)), i + 1, p + 3, true);
        consider(Wk.suf.get(i + 1, p), i + 1, p, false);
        consider(Wk.suf.get(i, p + 1), i, p + 1, false);
        consider(Wk.fwdEv[p], INT_MAX, INT_MAX, false);
        if (em) addForwardMatch(al, i, p + 1);
        i = bi;
        p = bj;
    }
}

template <class F>
static void reverseAlignment(Work<F> &Wk, std::vector<Seg> &al, int iEnd, int jEnd) {
    // dummest-traceback.hh:68-95
    int i = iEnd - 1, p = jEnd;
    while (i >= 0 && p >= 0) {
        ES best{0.0f, 0};
        int bi = -1, bj = -1;
        bool em = false;
        auto consider = [&](const ES &c, int ni, int nj, bool e) {
            if (scoreLess(best, c)) { best = c; bi = ni; bj = nj; em = e; }
        };
        if (p >= 3) {
            ES os = (i >= 1) ? Wk.pre.get(i - 1, p - 3) : ES{0.0f, 0};
            consider(addScores(Wk.X.get(i, p), os), i - 1, p - 3, true);
        }
        if (i >= 1) consider(Wk.pre.get(i - 1, p), i - 1, p, false);
        if (p > 0) consider(Wk.pre.get(i, p - 1), i, p - 1, false);
        consider(Wk.revEv[p + 3], -1, -1, false);
        if (em) addReverseMatch(al, i, p - 2);
        i = bi;
        p = bj;
    }
}

// rescale factor for one lane, dummest-score.hh:124-152
template <class F> static inline F rowRescale(int cnt) { return std::ldexp((F)1.0, -RESCALE_BITS * clampRescaleCount(cnt)); }

template <class F>
static void search(Work<F> &Wk, std::vector<Sim> &out, const Prof<F> &P, const uint8_t *codes, int n, F minProbRatio,
                   bool needAlign, bool wantBwdBest, const Rates<F> &R, F thresh, int rescaleBits) {
    const int M = P.M;
    const int W = n;
    Wk.W = W;
    Null<F> &N = Wk.N;
    buildNull<F>(N, P.bg + 4, codes, n, W, R);
    auto code = [&](int j) -> int { return (j >= 0 && j < n) ? codes[j] : kZeroIdx; };
    const F s1 = N.s1f, s2 = N.s2f, sC = N.sCf;
    const F scale = 1.0;
    const F step = std::ldexp((F)1.0, -rescaleBits);
    auto rowScaleOf = [&](int cnt) { return std::ldexp((F)1.0, -rescaleBits * clampRescaleCount(cnt)); };
    auto promote = [&](F v, int cnt) -> ES {
        if (std::fpclassify(v) == FP_SUBNORMAL) return ES{0.0f, 0};
        if (!(v > 0)) return ES{0.0f, 0};
        cnt = clampRescaleCount(cnt);
        if (!(v < (F)INFINITY)) return ES{std::nextafterf(1.0f, 0.0f), 1023 + rescaleBits * cnt};
        return normalizeScore((double)v, rescaleBits * cnt);
    };
    std::vector<F> bgc(W + 8);
    for (int j = -4; j < W + 4; j++) bgc[j + 4] = P.bg[4 + code(j)];
    auto bgAt = [&](int j) { return bgc[j + 4]; };
    Wk.W1.assign((size_t)(M + 2) * (W + 4), (F)0);
    Wk.bcnt.assign(M + 2, 0);
    Wk.fcnt.assign(M + 2, 0);
    Wk.rowMaxExp.assign(M + 2, INT_MIN);
    Wk.revSide.assign(W + 4 + 3, (F)0);
    std::vector<F> Ynext(W + 4, (F)0), Ycur(W + 4, (F)0);
    // ---- backward, dummest-dp-backward.hh:473-575
    for (int i = M; i >= 0; i--) {
        if (i < M) Wk.bcnt[i] = Wk.bcnt[i + 1];
        const F *pr = P.par(i);
        const F *em = P.emis(i);
        const F enterC = pr[10] * sC, d0 = pr[6], d1 = pr[7] * s2, d2 = pr[8] * s1;
        const F a0 = pr[0] * sC, a1 = pr[1] * (g_mutation == 1 ? s2 : s1), a2 = pr[2] * (g_mutation == 1 ? s1 : s2),
                b0 = pr[3] * sC;
        const int goff = (g_mutation == 4) ? 1 : 0;
        F Z[4] = {0, 0, 0, 0};
        const F eps = pr[9];
        F rsc = scale * rowScaleOf(Wk.bcnt[i]);
        F rowMax = 0;
        // shift registers start at (i+1, hi+3..hi+1) and (i, hi+2..hi+1), hi = W-1
        F n3 = *Wk.w1(i + 1, W + 2), n2 = *Wk.w1(i + 1, W + 1), n1 = *Wk.w1(i + 1, W);
        F c2 = *Wk.w1(i, W + 1), c1 = *Wk.w1(i, W);
        for (int j = W - 1; j >= 0; j--) {
            int r0 = j & 3, r3 = (j + 3) & 3;
            F e = em[code(j + 1)];
            F g = bgAt(j + 1 + goff);
            F nAt = *Wk.w1(i + 1, j);
            F w = n3 * e * enterC + Ynext[j] * d0 + n2 * d1 + n1 * d2 + Z[r3] * g * a0 + c2 * a2 + c1 * a1 +
                  N.suffixSeed[j] * rsc;
            F act = (F)(j < n ? 1 : 0);
            w = w * act;
            *Wk.w1(i, j) = w;
            if (needAlign) Wk.revSide[j + 3] = std::fma((F)1.0, w, Wk.revSide[j + 3]);
            rowMax = std::max(rowMax, w);
            Ycur[j] = std::fma(eps, Ynext[j], w) * act;
            rowMax = std::max(rowMax, Ycur[j]);
            F zf = Z[r3] * g;
            Z[r0] = std::fma(b0, zf, w);
            n3 = n2; n2 = n1; n1 = nAt;
            c2 = c1; c1 = w;
        }
        std::swap(Ycur, Ynext);
        if (rowMax > thresh) {
            Wk.bcnt[i]++;
            F *rp = Wk.w1(i, 0);
            for (int j = 0; j < W + 4; ++j) rp[j] *= step;
            for (auto &v : Ynext) v *= step;
            for (auto &v : Ycur) v *= step;
            if (needAlign) for (auto &v : Wk.revSide) v *= step;
        }
    }
    // bestOut, dummest-dp-backward.hh:579-609
    Wk.bwdBest = ES{};
    if (wantBwdBest) {
        ES bm{};
        for (int j = 0; j < W; j++) {
            double nd = (double)N.prefixSeed[j];
            for (int i = M; i >= 0; i--) {
                ES r = promote(*Wk.w1(i, j), Wk.bcnt[i]);
                ES c = batchScale(r, nd);
                if (scoreLess(bm, c)) bm = c;
            }
        }
        Wk.bwdBest = bm;
    }
    // ---- forward, dummest-dp-forward.hh:29-250
    if (Wk.capture) {
        Wk.w0es.assign((size_t)(M + 1) * n, ES{});
        Wk.W1es.assign((size_t)(M + 1) * n, ES{});
    }
    std::fill(Ynext.begin(), Ynext.end(), (F)0);
    std::vector<F> W0c(W + 4, (F)0), W0n(W + 4, (F)0);
    Wk.colBest.assign(W + 4, ES{});
    Wk.colRow.assign(W + 4, 0);
    Wk.fwdSide.assign(W + 4, (F)0);
    if (needAlign) Wk.X.assign(M + 2, W + 4, 0);
    ES gbest{};
    for (int i = 0; i <= M; i++) {
        if (i > 0) Wk.fcnt[i] = Wk.fcnt[i - 1];
        const F *pr = P.par(i);
        const F *em = P.emis(i);
        const F enterC = pr[10] * sC;
        const F a0 = pr[0], a1 = pr[1] * s1, a2 = pr[2] * s2, b0 = pr[3];
        const F d0 = pr[6], d1 = pr[7] * s2, d2 = pr[8] * s1, eps = pr[9];
        F rsc = scale * rowScaleOf(Wk.fcnt[i]);
        F rowMax = 0;
        F Z0[4] = {0, 0, 0, 0}, Z1[4] = {0, 0, 0, 0}, Z2[4] = {0, 0, 0, 0};
        const bool hasNext = (i + 1 <= M);
        const int eoff = (g_mutation == 2) ? 1 : 0;
        F ws[3] = {0, 0, 0};
        for (int j = 0; j < W; j++) {
            F w1v = ws[0], w2v = ws[1], w3v = ws[2];
            int r0 = j & 3, r3 = (j - 3) & 3;
            F e = em[code(j - 2 + eoff)];
            F g = bgAt(j - 2);
            F Xij = enterC * e * w3v;
            if (needAlign) {
                F w1n = hasNext ? *Wk.w1(i + 1, j) : (F)0;
                ES xf = promote(Xij, Wk.fcnt[i]);
                ES wf = promote(w1n, Wk.bcnt[i + 1]);
                ES x = batchMul(xf, wf, 0);
                Wk.X.at(i, j) = x;
                if (x.m != 0.0f && x.e > Wk.rowMaxExp[i]) Wk.rowMaxExp[i] = x.e;
            }
            Z0[r0] = g * sC * (a0 * w3v + b0 * Z0[r3]);
            Z1[r0] = a1 * w1v;
            Z2[r0] = a2 * w2v;
            F w0 = W0c[j];
            w0 = w0 + (Z0[r0] + Z1[r0] + Z2[r0] + N.prefixSeed[j] * rsc);
            if (needAlign) Wk.fwdSide[j] = std::fma((F)1.0, w0, Wk.fwdSide[j]);
            rowMax = std::max(rowMax, w0);
            F w1i = *Wk.w1(i, j);
            F act = (F)(j < n ? 1 : 0);
            ES fm = promote(w0, Wk.fcnt[i]);
            ES bm = promote(w1i, Wk.bcnt[i]);
            if (Wk.capture && j < n) {
                Wk.w0es[(size_t)i * n + j] = fm;
                Wk.W1es[(size_t)i * n + j] = bm;
            }
            if (j >= n) { fm.m = 0.0f; bm.m = 0.0f; }
            ES wMid = batchMul(fm, bm, 0);
            if (!(j < n)) wMid = ES{};
            if (scoreLess(gbest, wMid)) gbest = wMid;
            if (scoreLess(Wk.colBest[j], wMid)) { Wk.colBest[j] = wMid; Wk.colRow[j] = i; }
            w0 = w0 * act;
            W0c[j] = w0;
            Z0[r0] = Z0[r0] * act;
            Z1[r0] = Z1[r0] * act;
            Z2[r0] = Z2[r0] * act;
            ws[2] = ws[1];
            ws[1] = ws[0];
            ws[0] = w0;
            Ycur[j] = d0 * w0 + eps * Ynext[j];
            Ycur[j] = Ycur[j] * act;
            rowMax = std::max(rowMax, Ycur[j]);
            if (hasNext) W0n[j] = W0n[j] + (Xij + Ycur[j] + d1 * w2v + d2 * w1v);
        }
        if (rowMax > thresh) {
            Wk.fcnt[i]++;
            if (hasNext) for (auto &v : W0n) v *= step;
            for (auto &v : Ynext) v *= step;
            for (auto &v : Ycur) v *= step;
            if (needAlign) for (auto &v : Wk.fwdSide) v *= step;
        }
        std::swap(Ycur, Ynext);
        std::swap(W0c, W0n);
        std::fill(W0n.begin(), W0n.end(), (F)0);
    }
    Wk.globalBest = gbest;
    // ---- gate, dummest-dp-forward.hh:252-274
    ES minExp{};
    if (minProbRatio >= 0) minExp = normalizeScore((double)minProbRatio, 0);
    const bool hit = (minProbRatio < 0) || !scoreLess(gbest, minExp);
    if (!hit) return;
    if (needAlign) {
        buildFlankEVs<F>(Wk, P, codes, n, R, rescaleBits);
        int revMax = INT_MIN, fwdMax = INT_MIN;  // sideEVMax, dummest-dp.hh:273-282
        for (int j = 0; j < W; j++) {
            const ES &a = Wk.revEv[j + 3];
            if (a.m != 0.0f && a.e > revMax) revMax = a.e;
            const ES &b = Wk.fwdEv[j];
            if (b.m != 0.0f && b.e > fwdMax) fwdMax = b.e;
        }
        Wk.pre.assign(M + 2, W + 4, 3);
        Wk.suf.assign(M + 2, W + 4, 0);
        // prefixOptimal, dummest-dp-forward.hh:298-371
        {
            int runMax = INT_MIN;
            for (int i = 0; i <= M; i++) {
                if (Wk.rowMaxExp[i] > runMax) runMax = Wk.rowMaxExp[i];
                const int E = (runMax > revMax ? runMax : revMax) + 1;
                float left = 0.0f;
                for (int j = 0; j < W; j++) {
                    const ES &rs = Wk.revEv[j + 3];
                    const ES cx = Wk.X.get(i, j);
                    const float comb = scoreToRowScale(cx.m, cx.e, E);
                    const float rsr = scoreToRowScale(rs.m, rs.e, E);
                    float mx;
                    if (i == 0) {
                        mx = 0.0f;
                        if (left > mx) mx = left;
                        if (comb > mx) mx = comb;
                        if (rsr > mx) mx = rsr;
                    } else {
                        const ES pp = Wk.pre.get(i - 1, j - 3);
                        const float prev = scoreToRowScale(pp.m, pp.e, E);
                        const float cps = comb + prev;
                        const ES up = Wk.pre.get(i - 1, j);
                        const float upr = scoreToRowScale(up.m, up.e, E);
                        mx = upr;
                        if (left > mx) mx = left;
                        if (cps > mx) mx = cps;
                        if (rsr > mx) mx = rsr;
                    }
                    left = mx;
                    Wk.pre.at(i, j) = scoreFromRowScale(mx, E);
                }
            }
        }
        // suffixOptimal, dummest-dp-forward.hh:380-452
        {
            int runMax = INT_MIN;
            for (int i = M; i >= 0; i--) {
                if (Wk.rowMaxExp[i] > runMax) runMax = Wk.rowMaxExp[i];
                const int E = (runMax > fwdMax ? runMax : fwdMax) + 1;
                float right = 0.0f;
                for (int j = W - 1; j >= 0; j--) {
                    const ES &lv = Wk.fwdEv[j];
                    const float fsr = scoreToRowScale(lv.m, lv.e, E);
                    float mx;
                    if (i == M) {
                        mx = 0.0f;
                        if (right > mx) mx = right;
                        if (fsr > mx) mx = fsr;
                    } else {
                        const ES cx = Wk.X.get(i, j + 3);
                        const float comb = scoreToRowScale(cx.m, cx.e, E);
                        const ES sn = Wk.suf.get(i + 1, j + 3);
                        const float snr = scoreToRowScale(sn.m, sn.e, E);
                        const float cps = comb + snr;
                        const ES dn = Wk.suf.get(i + 1, j);
                        const float dnr = scoreToRowScale(dn.m, dn.e, E);
                        mx = dnr;
                        if (right > mx) mx = right;
                        if (fsr > mx) mx = fsr;
                        if (cps > mx) mx = cps;
                    }
                    right = mx;
                    Wk.suf.at(i, j) = scoreFromRowScale(mx, E);
                }
            }
        }
    }
    // anchors, dummest-dp-forward.hh:459-545
    std::vector<Sim> opt((size_t)n, Sim{ES{}, -INFINITY, 0, 0, {}});
    ES bestOverall{};
    for (int j = 0; j < n; j++) {
        const ES &bp = Wk.colBest[j];
        if (bp.m != 0.0) {
            if (scoreLess(bestOverall, bp)) bestOverall = bp;
            opt[j] = Sim{bp, -INFINITY, Wk.colRow[j], j, {}};
        }
    }
    if (minProbRatio >= 0) {
        if (bestOverall.m != 0.0 && !scoreLess(bestOverall, minExp)) {
            std::sort(opt.begin(), opt.end(), std::greater<>());
            std::vector<bool> aligned(n, false);
            for (auto &as : opt) {
                int lj = as.anchor2;
                if (as.pair.m != 0.0 && !scoreLess(as.pair, minExp) && !aligned[lj]) {
                    Sim s;
                    s.pair = as.pair;
                    s.logProbRatio = scoreToLog(as.pair);
                    s.anchor1 = as.anchor1;
                    s.anchor2 = as.anchor2;
                    if (needAlign) forwardAlignment<F>(Wk, M, s.alignment, s.anchor1, s.anchor2);
                    std::reverse(s.alignment.begin(), s.alignment.end());
                    if (needAlign) reverseAlignment<F>(Wk, s.alignment, s.anchor1, s.anchor2);
                    std::reverse(s.alignment.begin(), s.alignment.end());
                    int beg = s.alignment.empty() ? s.anchor2 : s.alignment[0].start2;
                    int en = s.alignment.empty() ? s.anchor2 : s.alignment.back().start2 + s.alignment.back().length;
                    int cb = std::max(beg, 0), ce = std::min(en, n);
                    bool ov = false;
                    for (int p = cb; p < ce; ++p)
                        if (aligned[p]) { ov = true; break; }
                    if (!ov) {
                        std::fill(aligned.begin() + cb, aligned.begin() + ce, true);
                        out.push_back(std::move(s));
                    }
                }
            }
        }
    } else {
        Sim sel = *std::max_element(opt.begin(), opt.end());
        Sim b = sel;
        b.logProbRatio = -INFINITY;
        sel.logProbRatio = scoreToLog(sel.pair);
        out.push_back(b);
        out.push_back(b);
        out.push_back(sel);
    }
}

// ----------------------------------------------------------------------------
// Gumbel fits, dummest-calibration.hh:7-104 (sorts its input in place).
// ----------------------------------------------------------------------------
static double meanOf(const double *x, int n) {
    double s = 0;
    for (int i = 0; i < n; ++i) s += x[i];
    return s / n;
}
static double mmLambda(const double *s, int n, double mu) {
    double pi = 3.1415926535897932, a = 0;
    for (int i = 0; i < n; ++i) a += (s[i] - mu) * (s[i] - mu);
    return pi / std::sqrt(6 * (a / n));
}
static double mmK(double mu, double lam, double len) { return std::exp(lam * mu - 0.57721566490153286) / len; }
static double lmLambda(const double *s, int n, double mu) {
    double a = 0;
    for (int i = 0; i < n; ++i) a += i * s[i];
    double d = 0.5 * n * (n - 1);
    return std::log(2.0) / (a / d - mu);
}
static double shouldBe0(const double *s, int n, double lam) {
    double x = 0, y = 0, z = 0;
    for (int i = 0; i < n; ++i) {
        x += s[i];
        y += std::exp(-lam * s[i]);
        z += s[i] * std::exp(-lam * s[i]);
    }
    return 1 / lam - x / n + z / y;
}
static double mlLambda(const double *s, int n) {
    double lo = 1, hi = 1, x, y;
    int it = 0;
    do {
        if (++it > 1000) return 1.0;
        lo /= 2;
        hi *= 2;
        x = shouldBe0(s, n, lo);
        y = shouldBe0(s, n, hi);
    } while ((x < 0 && y < 0) || (x > 0 && y > 0));
    double gap = hi - lo;
    while (1) {
        gap /= 2;
        double mid = lo + gap;
        if (mid <= lo) return lo;
        double z = shouldBe0(s, n, mid);
        if ((x < 0 && z <= 0) || (x > 0 && z >= 0)) lo = mid;
    }
}
static double mlK(const double *s, int n, double lam, double len) {
    double a = 0;
    for (int i = 0; i < n; ++i) a += std::exp(-lam * s[i]);
    return n / (a * len);
}


// ----------------------------------------------------------------------------
// Calibration sequence for one trial, dummest-calibration.hh:363-376 (aa
// frequencies) and 436-482 (generator). Writes len+border letters.
// Codon lists in the reference's order (dummest-sequence.hh:103-124): the
// synonymous-codon draw indexes into them.
// ----------------------------------------------------------------------------
static const std::vector<std::string> &codonList(char aa) {
    static std::unordered_map<char, std::vector<std::string>> t;
    if (t.empty()) {
        t['A'] = {"GCT", "GCC", "GCA", "GCG"};
        t['R'] = {"CGT", "CGC", "CGA", "CGG", "AGA", "AGG"};
        t['N'] = {"AAT", "AAC"};
        t['D'] = {"GAT", "GAC"};
        t['C'] = {"TGT", "TGC"};
        t['Q'] = {"CAA", "CAG"};
        t['E'] = {"GAA", "GAG"};
        t['G'] = {"GGT", "GGC", "GGA", "GGG"};
        t['H'] = {"CAT", "CAC"};
        t['I'] = {"ATT", "ATC", "ATA"};
        t['L'] = {"TTA", "TTG", "CTT", "CTC", "CTA", "CTG"};
        t['K'] = {"AAA", "AAG"};
        t['M'] = {"ATG"};
        t['F'] = {"TTT", "TTC"};
        t['P'] = {"CCT", "CCC", "CCA", "CCG"};
        t['S'] = {"TCT", "TCC", "TCA", "TCG", "AGT", "AGC"};
        t['T'] = {"ACT", "ACC", "ACA", "ACG"};
        t['W'] = {"TGG"};
        t['Y'] = {"TAT", "TAC"};
        t['V'] = {"GTT", "GTC", "GTA", "GTG"};
        t['*'] = {"TAA", "TAG", "TGA"};
    }
    return t.at(aa);
}

template <class F>
static void calibSequence(const F *letterFreqs20, F bgStop, F bgfs1, F bgfs2, int L, int border, int trial,
                          char *out) {
    std::vector<double> aaFreqs(21);
    for (int k = 0; k < 20; ++k) {
        int cnt = (int)codonList(kAlphabet[k]).size();
        aaFreqs[k] = letterFreqs20[k] * cnt;  // Float * int
    }
    aaFreqs[20] = bgStop;
    std::discrete_distribution<> dist(aaFreqs.begin(), aaFreqs.end());
    std::mt19937_64 gen(5489 + trial);
    const char bases[] = {'A', 'C', 'G', 'T'};
    std::uniform_int_distribution<int> distDNA(0, 3);
    std::uniform_int_distribution<int> distOffset(0, 2);
    std::vector<char> buf(L + border + 16, 'A');
    int offset = distOffset(gen);
    for (int j = 0; j < offset; j++) buf[j] = bases[distDNA(gen)];
    for (int j = offset; j <= L; j += 3) {
        double r = std::generate_canonical<double, 10>(gen);
        if (r < bgfs1) {
            if (j <= L) buf[j] = bases[distDNA(gen)];
            j -= 2;
            continue;
        } else if (r < (F)(bgfs1 + bgfs2)) {
            if (j <= L) buf[j] = bases[distDNA(gen)];
            if (j + 1 <= L) buf[j + 1] = bases[distDNA(gen)];
            j -= 1;
            continue;
        }
        int x = dist(gen);
        char aa = (x < 20) ? kAlphabet[x] : '*';
        const auto &codons = codonList(aa);
        std::uniform_int_distribution<int> dist2(0, (int)codons.size() - 1);
        const auto &xx = codons[dist2(gen)];
        for (int k = 0; k < 3; k++)
            if (j + k <= L) buf[j + k] = xx[k];
    }
    for (int j = 0; j < border; ++j) buf[L + j] = buf[j];
    std::memcpy(out, buf.data(), (size_t)(L + border));
}

}  // namespace orc

using namespace orc;

// ============================================================================
// C ABI for ctypes (tests only).
// ============================================================================
template <class F> static Prof<F> mkProf(int M, const F *values, const F *params, const F *bg) {
    return Prof<F>{M, values, params, bg};
}

template <class F>
static int searchImpl(int M, const F *values, const F *params, const F *bg, const double *opts, const uint8_t *codes,
                      int n, double minProbRatio, int needAlign, int wantBwd, int rescaleBits, double thresh,
                      double *hitScalars /*cap x 6: lnratio, m, e, a1, a2, nseg*/, int *segs /*segCap x 3*/,
                      int hitCap, int segCap, float *gbest /*m,e*/, float *bwdBest, float *colBest /*n x (m,e,row)*/) {
    FtzScope ftz_;
    Rates<F> R;
    R.init(opts);
    Prof<F> P = mkProf<F>(M, values, params, bg);
    Work<F> Wk;
    std::vector<Sim> out;
    search<F>(Wk, out, P, codes, n, (F)minProbRatio, needAlign != 0, wantBwd != 0, R, (F)thresh, rescaleBits);
    if (gbest) { gbest[0] = Wk.globalBest.m; gbest[1] = (float)Wk.globalBest.e; }
    if (bwdBest) { bwdBest[0] = Wk.bwdBest.m; bwdBest[1] = (float)Wk.bwdBest.e; }
    if (colBest)
        for (int j = 0; j < n; j++) {
            colBest[3 * j] = Wk.colBest[j].m;
            colBest[3 * j + 1] = (float)Wk.colBest[j].e;
            colBest[3 * j + 2] = (float)Wk.colRow[j];
        }
    int nh = 0, ns = 0;
    for (auto &s : out) {
        if (nh >= hitCap) return -1;
        double *h = hitScalars + 6 * nh;
        h[0] = s.logProbRatio;
        h[1] = s.pair.m;
        h[2] = s.pair.e;
        h[3] = s.anchor1;
        h[4] = s.anchor2;
        h[5] = (double)s.alignment.size();
        for (auto &g : s.alignment) {
            if (ns >= segCap) return -2;
            segs[3 * ns] = g.start1;
            segs[3 * ns + 1] = g.start2;
            segs[3 * ns + 2] = g.length;
            ns++;
        }
        nh++;
    }
    return nh;
}

// Total path mass computed both ways (ln): forward sum_{i,j} w0(i,j)*suffixSeed(j)
// and backward sum_{i,j} prefixSeed(j)*W1(i,j). Equal for a consistent pair of
// recurrences; used to pin the DP restatement.
template <class F>
static void totalsImpl(int M, const F *values, const F *params, const F *bg, const double *opts, const uint8_t *codes,
                       int n, int rescaleBits, double thresh, double *out) {
    FtzScope ftz_;
    Rates<F> R;
    R.init(opts);
    Prof<F> P = mkProf<F>(M, values, params, bg);
    Work<F> Wk;
    std::vector<Sim> sims;
    search<F>(Wk, sims, P, codes, n, (F)-2.0, true, false, R, (F)thresh, rescaleBits);
    // backward total: W1 stored at row scale bcnt[i]
    double bmax = -INFINITY, fmax = -INFINITY;
    std::vector<double> bl, fl;
    for (int i = 0; i <= M; i++)
        for (int j = 0; j < n; j++) {
            double w = (double)*Wk.w1(i, j) * (double)Wk.N.prefixSeed[j];
            if (w > 0) {
                double l = std::log(w) + (double)rescaleBits * Wk.bcnt[i] * 0.6931471805599453;
                bl.push_back(l);
                bmax = std::max(bmax, l);
            }
        }
    double bs = 0;
    for (double l : bl) bs += std::exp(l - bmax);
    out[0] = bmax + std::log(bs);
    // forward total from the per-row forward values is not kept by search();
    // use the column totals forwardSide (sum over rows of w0 at scale fcnt[M]).
    for (int j = 0; j < n; j++) {
        double w = (double)Wk.fwdSide[j] * (double)Wk.N.suffixSeed[j];
        if (w > 0) {
            double l = std::log(w) + (double)rescaleBits * Wk.fcnt[M] * 0.6931471805599453;
            fl.push_back(l);
            fmax = std::max(fmax, l);
        }
    }
    double fs = 0;
    for (double l : fl) fs += std::exp(l - fmax);
    out[1] = fmax + std::log(fs);
}

extern "C" {

void or_set_mutation(int m) { g_mutation = m; }

int or_finalize_f32(int M, float *values, float *params, float *bg, const double *opts, int bgType,
                    const float *tantan, double thr, double filterSd, int keepNonvarying) {
    Rates<float> R;
    R.init(opts);
    return finalizeProfile<float>(M, values, params, bg, R, bgType, tantan, (float)thr, filterSd, keepNonvarying != 0);
}
int or_finalize_f64(int M, double *values, double *params, double *bg, const double *opts, int bgType,
                    const float *tantan, double thr, double filterSd, int keepNonvarying) {
    Rates<double> R;
    R.init(opts);
    return finalizeProfile<double>(M, values, params, bg, R, bgType, tantan, thr, filterSd, keepNonvarying != 0);
}

void or_tantan(const uint8_t *seq, int n, int isProtein, float *out) { tantanProbs(seq, n, isProtein != 0, out); }
void or_blosum62(signed char *out) { std::memcpy(out, kBlosum62, sizeof(kBlosum62)); }
void or_mask_dna(const char *dna, int n, double thr, char *outDna, float *probs) { maskDna(dna, n, thr, outDna, probs); }

void or_decode(const char *dna, int n, uint8_t *codes) { decodeDna(dna, n, codes); }

// Null model outputs: prefixSeed[n], suffixSeed[n], scales[3] (s1, s2, sC packed),
// prefix/suffix ES as (m, e) pairs for n+4 entries.
#define OR_NULL(F, SUF)                                                                                          \
    void or_null_##SUF(const F *bg29, const uint8_t *codes, int n, const double *opts, F *prefixSeed,          \
                       F *suffixSeed, F *scales, float *prefixM, int *prefixE, float *suffixM, int *suffixE) {  \
        FtzScope ftz_;                                                                                                \
        Rates<F> R;                                                                                              \
        R.init(opts);                                                                                            \
        Null<F> N;                                                                                               \
        buildNull<F>(N, bg29 + 4, codes, n, n, R);                                                               \
        for (int j = 0; j < n; j++) { prefixSeed[j] = N.prefixSeed[j]; suffixSeed[j] = N.suffixSeed[j]; }       \
        scales[0] = N.s1f; scales[1] = N.s2f; scales[2] = N.sCf;                                                 \
        if (prefixM)                                                                                             \
            for (int j = 0; j < n + 4; j++) {                                                                    \
                prefixM[j] = N.prefix[j].m; prefixE[j] = N.prefix[j].e;                                          \
                suffixM[j] = N.suffix[j].m; suffixE[j] = N.suffix[j].e;                                          \
            }                                                                                                    \
    }
OR_NULL(float, f32)
OR_NULL(double, f64)

double or_filter_f32(int M, const float *values, const float *params, const float *bg, const double *opts,
                     const uint8_t *codes, int n) {
    FtzScope ftz_;
    Rates<float> R;
    R.init(opts);
    return (double)filterScore<float>(mkProf<float>(M, values, params, bg), codes, n, R);
}
double or_filter_f64(int M, const double *values, const double *params, const double *bg, const double *opts,
                     const uint8_t *codes, int n) {
    FtzScope ftz_;
    Rates<double> R;
    R.init(opts);
    return filterScore<double>(mkProf<double>(M, values, params, bg), codes, n, R);
}

// Batch of filter scores over many (profile-free) sequences of one profile,
// threaded: the CPU baseline's filter wave.
void or_filter_batch_f32(int M, const float *values, const float *params, const float *bg, const double *opts,
                         const uint8_t *codes, const int64_t *offsets, const int *lens, int nseq, float *out,
                         int nthreads) {
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
    for (int s = 0; s < nseq; s++) {
        FtzScope ftz_;
        Rates<float> R;
        R.init(opts);
        out[s] = filterScore<float>(mkProf<float>(M, values, params, bg), codes + offsets[s], lens[s], R);
    }
}

int or_search_f32(int M, const float *values, const float *params, const float *bg, const double *opts,
                  const uint8_t *codes, int n, double minProbRatio, int needAlign, int wantBwd, int rescaleBits,
                  double thresh, double *hits, int *segs, int hitCap, int segCap, float *gbest, float *bwdBest,
                  float *colBest) {
    return searchImpl<float>(M, values, params, bg, opts, codes, n, minProbRatio, needAlign, wantBwd, rescaleBits,
                             thresh, hits, segs, hitCap, segCap, gbest, bwdBest, colBest);
}
int or_search_f64(int M, const double *values, const double *params, const double *bg, const double *opts,
                  const uint8_t *codes, int n, double minProbRatio, int needAlign, int wantBwd, int rescaleBits,
                  double thresh, double *hits, int *segs, int hitCap, int segCap, float *gbest, float *bwdBest,
                  float *colBest) {
    return searchImpl<double>(M, values, params, bg, opts, codes, n, minProbRatio, needAlign, wantBwd, rescaleBits,
                              thresh, hits, segs, hitCap, segCap, gbest, bwdBest, colBest);
}

// DP matrices of one pair as exact ExpScores (w0 promoted at the forward row
// count, W1 at the backward row count), plus prefixSeed and the scales: the
// inputs of the post-gate processing.
void or_matrices_f32(int M, const float *values, const float *params, const float *bg, const double *opts,
                     const uint8_t *codes, int n, float *w0m, int *w0e, float *W1m, int *W1e, float *prefixSeed,
                     float *scales) {
    FtzScope ftz_;
    Rates<float> R;
    R.init(opts);
    Work<float> Wk;
    Wk.capture = true;
    std::vector<Sim> sims;
    search<float>(Wk, sims, mkProf<float>(M, values, params, bg), codes, n, -2.0f, false, false, R,
                  rescaleThresh<float>(), RESCALE_BITS);
    for (size_t c = 0; c < Wk.w0es.size(); c++) {
        w0m[c] = Wk.w0es[c].m;
        w0e[c] = Wk.w0es[c].e;
        W1m[c] = Wk.W1es[c].m;
        W1e[c] = Wk.W1es[c].e;
    }
    for (int j = 0; j < n; j++) prefixSeed[j] = Wk.N.prefixSeed[j];
    scales[0] = Wk.N.s1f;
    scales[1] = Wk.N.s2f;
    scales[2] = Wk.N.sCf;
}

void or_totals_f32(int M, const float *values, const float *params, const float *bg, const double *opts,
                   const uint8_t *codes, int n, int rescaleBits, double thresh, double *out) {
    totalsImpl<float>(M, values, params, bg, opts, codes, n, rescaleBits, thresh, out);
}
void or_totals_f64(int M, const double *values, const double *params, const double *bg, const double *opts,
                   const uint8_t *codes, int n, int rescaleBits, double thresh, double *out) {
    totalsImpl<double>(M, values, params, bg, opts, codes, n, rescaleBits, thresh, out);
}

// Batched full search (scores only): per pair (global best mid, bwd best), threaded.
// The CPU baseline's fwd+bwd leg.
void or_fwdbwd_batch_f32(int M, const float *values, const float *params, const float *bg, const double *opts,
                         const uint8_t *codes, const int64_t *offsets, const int *lens, int nseq, float *midOut,
                         float *bwdOut, int nthreads) {
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
    for (int s = 0; s < nseq; s++) {
        FtzScope ftz_;
        Rates<float> R;
        R.init(opts);
        Work<float> Wk;
        std::vector<Sim> out;
        search<float>(Wk, out, mkProf<float>(M, values, params, bg), codes + offsets[s], lens[s], (float)-2.0, false,
                      true, R, rescaleThresh<float>(), RESCALE_BITS);
        midOut[2 * s] = Wk.globalBest.m;
        midOut[2 * s + 1] = (float)Wk.globalBest.e;
        bwdOut[2 * s] = Wk.bwdBest.m;
        bwdOut[2 * s + 1] = (float)Wk.bwdBest.e;
    }
}

// Gumbel fits on n scores (sorted in place). out: mmL mmK mmKs mlL mlK mlKs lmL lmK
void or_gumbel(double *scores, int n, double seqLen, double *out) {
    std::sort(scores, scores + n);
    double mu = meanOf(scores, n);
    out[0] = mmLambda(scores, n, mu);
    out[1] = mmK(mu, out[0], seqLen);
    out[2] = mmK(mu, 1, seqLen);
    out[3] = mlLambda(scores, n);
    out[4] = mlK(scores, n, out[3], seqLen);
    out[5] = mlK(scores, n, 1, seqLen);
    out[6] = lmLambda(scores, n, mu);
    out[7] = mmK(mu, out[6], seqLen);
}

void or_calib_sequence_f32(const float *letterFreqs20, const double *opts, int L, int border, int trial, char *out) {
    Rates<float> R;
    R.init(opts);
    calibSequence<float>(letterFreqs20, R.bgstop, R.bgfs1, R.bgfs2, L, border, trial, out);
}

// ExpScore primitives for contract tests (dummest-score.hh, dummest-batch.hh).
// Baum-Welch E-step of dummest-build (dummest-build.cc:561-736), per sequence
// in order: forward (:561-613, the total v summed cell by cell in row-major
// order), backward (:615-666), calculateTransitionCounts (:668-736) with
// weight / (v 2^-480). probs: (L+1) x (7+A); letters: codes 0..A back to back
// plus one pad (the recurrences read one letter past a sequence). counts get
// the sums of every column but the delete starts (left 0); totals[s] = v.
// A letter code A (outside the alphabet) gets a zero likelihood ratio.
void or_bw_estep(int L, int A, const double *probs, int nseq, const unsigned char *letters, const int *lengths,
                 const double *weights, double *counts, double *totals) {
    const int width = 7 + A;
    const double scale = std::ldexp(1.0, -480);  // (1/2^30)^16, dummest-build.cc:36-38
    int maxN = 0;
    for (int s = 0; s < nseq; s++) maxN = std::max(maxN, lengths[s]);
    const size_t area = (size_t)(L + 2) * (maxN + 2);
    std::vector<double> X(area), Y(area), Z(area), Wb(area), Yb(area), Zb(area);
    std::fill(counts, counts + (size_t)(L + 1) * width, 0.0);
    auto rowRates = [&](int r, double &ap, double &bp, double &dp, double &ep) {
        const double alpha = probs[r * width + 1], beta = probs[r * width + 4];
        const double delta = probs[r * width + 2], eps = probs[r * width + 6];
        const double eps1 = (r == L) ? 0.0 : probs[(r + 1) * width + 6];
        ap = alpha * (1.0 - beta);
        bp = beta;
        dp = delta * (1 - eps1);
        ep = eps * (1 - eps1) / (1 - eps);
        if (!std::isfinite(ep)) ep = 0.0;
    };
    auto ratio = [&](int r, int x) {
        if (x >= A) return 0.0;
        double S = (1 - probs[r * width + 1] - probs[r * width + 2]) * probs[r * width + 7 + x] /
                   probs[(size_t)L * width + 7 + x];
        return std::isfinite(S) ? S : 0.0;
    };
    const unsigned char *seq = letters;
    for (int s = 0; s < nseq; s++) {
        const int n = lengths[s], cols = n + 2;
        // forward
        double v = 0.0;
        std::fill_n(&X[0], n + 1, 0.0);
        std::fill_n(&Y[1], n + 1, 0.0);
        for (int i = 1; i <= L + 1; i++) {
            double ap, bp, dp, ep;
            rowRates(i - 1, ap, bp, dp, ep);
            X[(size_t)(i - 1) * cols] = 0.0;
            double z = Z[(size_t)i * cols] = 0.0;
            for (int j = 1; j <= n + 1; j++) {
                const double S = ratio(i - 1, seq[j - 1]);
                const double y = Y[(size_t)(i - 1) * cols + j];
                const double w = X[(size_t)(i - 1) * cols + (j - 1)] + y + z + scale;
                z = ap * w + bp * z;
                v += w;
                X[(size_t)i * cols + j] = S * w;
                Y[(size_t)i * cols + j] = dp * w + ep * y;
                Z[(size_t)i * cols + j] = z;
            }
        }
        totals[s] = v;
        // backward
        std::fill_n(&Wb[(size_t)(L + 1) * cols + 1], n + 1, 0.0);
        std::fill_n(&Yb[(size_t)(L + 1) * cols + 0], n + 1, 0.0);
        for (int i = L; i >= 0; i--) {
            double ap, bp, dp, ep;
            rowRates(i, ap, bp, dp, ep);
            Wb[(size_t)(i + 2) * cols - 1] = 0.0;
            double z = Zb[(size_t)(i + 1) * cols - 1] = 0.0;
            for (int j = n; j >= 0; j--) {
                const double S = ratio(i, seq[j]);
                const double x = S * Wb[(size_t)(i + 1) * cols + (j + 1)];
                const double y = Yb[(size_t)(i + 1) * cols + j];
                const double w = x + dp * y + ap * z + scale;
                z = w + bp * z;
                Wb[(size_t)i * cols + j] = w;
                Yb[(size_t)i * cols + j] = w + ep * y;
                Zb[(size_t)i * cols + j] = z;
            }
        }
        // expected counts
        const double wt = weights[s] / (v * scale);
        double oldGamma = 0;
        for (int i = 1; i <= L + 1; i++) {
            double emis[32] = {0};
            double betap = 0, beta = 0, epsp = 0, eps = 0, etap = 0;
            for (int j = 1; j <= n + 1; j++) {
                const size_t up = (size_t)(i - 1) * cols, here = (size_t)i * cols;
                betap += Z[here + (j - 1)] * Wb[up + (j - 1)];
                beta += Z[here + (j - 1)] * Zb[up + (j - 1)];
                epsp += Y[up + j] * Wb[up + (j - 1)];
                eps += Y[up + j] * Yb[up + (j - 1)];
                etap += Wb[up + (j - 1)];
                emis[seq[j - 1]] += X[here + j] * Wb[here + j];
            }
            beta -= betap;
            eps -= epsp;
            const double gamma = std::accumulate(emis, emis + A + 1, 0.0);
            const double notIns = etap * scale + oldGamma + epsp;
            oldGamma = gamma;
            double *c = counts + (size_t)(i - 1) * width;
            c[0] += gamma * wt;
            c[1] += notIns * wt;
            c[3] += betap * wt;
            c[4] += beta * wt;
            c[5] += epsp * wt;
            c[6] += eps * wt;
            if (i == L + 1) continue;
            for (int k = 0; k < A; k++) c[7 + k] += emis[k] * wt;
        }
        seq += n;
    }
}

void or_es_normalize(double m, int e, float *om, int *oe) { ES r = normalizeScore(m, e); *om = r.m; *oe = r.e; }
void or_es_add(float am, int ae, float bm, int be, float *om, int *oe) { ES r = addScores({am, ae}, {bm, be}); *om = r.m; *oe = r.e; }
void or_es_badd(float am, int ae, float bm, int be, float *om, int *oe) { ES r = batchAdd({am, ae}, {bm, be}); *om = r.m; *oe = r.e; }
void or_es_mul(float am, int ae, float bm, int be, float *om, int *oe) { ES r = mulScores({am, ae}, {bm, be}); *om = r.m; *oe = r.e; }
void or_es_bmul(float am, int ae, float bm, int be, float *om, int *oe) { ES r = batchMul({am, ae}, {bm, be}, 0); *om = r.m; *oe = r.e; }
void or_es_scale(float am, int ae, double f, float *om, int *oe) { ES r = scaleScore({am, ae}, f); *om = r.m; *oe = r.e; }
void or_es_bscale(float am, int ae, double f, float *om, int *oe) { ES r = batchScale({am, ae}, f); *om = r.m; *oe = r.e; }
void or_es_promote(float v, int cum, float *om, int *oe) { FtzScope ftz_; ES r = batchPromote<float>(v, cum); *om = r.m; *oe = r.e; }
double or_es_log(float m, int e) { return scoreToLog({m, e}); }
void or_es_to_rowscale(float m, int e, int E, float *out) { *out = scoreToRowScale(m, e, E); }
void or_es_from_rowscale(float s, int E, float *om, int *oe) { ES r = scoreFromRowScale(s, E); *om = r.m; *oe = r.e; }

}  // extern "C"
