// Data layout shared by the host library and the gfx950 kernels.
//
// HBM layout (all arrays device-resident, uploaded once per call):
//   profiles : per profile p, ProfInfo{M, parOff, emOff}
//              params [parOff + 11*i + f], i = 0..M+1   (alpha'[3] beta'[3] delta'[3] eps' enter,
//                                                        dummest-core.hh:233-239)
//              emis   [emOff + 32*i + code], i = 0..M   (row i holds node i+1's codon emissions,
//                                                        codes 0..24, dummest-profile.hh:206-224)
//              bg     [p*32 + code]                     (background per codon code, bg_probs+4)
//   sequences: decoded codes, uint8, concatenated (dummest-sequence.hh:181-204)
//   pairs    : PairInfo per (profile, sequence) pair, in work-queue order
//   records  : per pair, columns j in [-RPAD, n+RPAD): Rec (12 bytes)
//   units    : work units = up to 64/G pairs of one profile with similar lengths
#pragma once
#include <stdint.h>

namespace dm {

constexpr int kWidth = 29;     // values per profile row (20 letters + 9)
constexpr int kNCode = 25;     // decoded codes 0..24; 24 is the zero-probability pad (zero_idx)
constexpr int kZeroIdx = 24;
constexpr int kEmStride = 32;  // floats per emission row in HBM
constexpr int kRPad = 4;       // record padding columns on each side of a pair's sequence
constexpr int kRecBlock = 16;  // a pair's records start on a multiple of 16 (192 bytes, 64-byte aligned)
inline constexpr int64_t recSpan(int n) { return (n + 2 * kRPad + kRecBlock - 1) / kRecBlock * kRecBlock; }
                               // (every column beyond them reads the zero record at the edge)

// Column record j of a pair (12 bytes): prefix seed, the codon codes at j+1
// (low byte, used by the backward) and j-2 (second byte, used by the forward),
// suffix seed. The background codon probabilities g[j+1] and g[j-2] come from
// the profile's 25-entry table, which the DP kernel keeps in LDS.
struct Rec {
    float ps;
    uint32_t codes;
    float ss;
};

struct ProfInfo {
    int M;
    int pad;
    int64_t parOff;
    int64_t emOff;
};

struct PairInfo {
    int prof;
    int seq;
    int n;
    int status;      // 0 ok; nonzero: null-model failure (sequence shorter than 3)
    int64_t seqOff;  // offset of the sequence codes
    int64_t recOff;  // index of column 0 in the record arrays
    int64_t bufOff;  // float offset of this pair's W1 / w0 buffers (full and dump modes)
    float s1, s2, sC;  // per-pair background scales, packed floats (dummest-dp-backward.hh:244-246)
    int outIdx;        // the pair's result slot (its index in the caller's pair list)
    int64_t nullOff;   // null-kernel scratch: entry j of this pair at nullOff + 64*j
};

struct Unit {
    int prof;
    int first;  // first pair index
    int count;  // pairs in this unit (<= 64/G)
    int maxN;   // longest sequence among them
};

// Output of the full (forward+backward) score pass per pair, as ExpScore pairs.
struct PairScore {
    float midM;  // best anchored score wMid, mantissa in [0.5,1) or 0
    int midE;
    float bwdM;  // best backward start score (bestOut, dummest-dp-backward.hh:579-609)
    int bwdE;
};

// ExpScore: value m * 2^e with m in [0.5,1) or m == 0 (dummest-score.hh:29-33).
struct ES {
    float m;
    int e;
};

enum Mode { kFilter = 0, kScore = 1, kDump = 2 };

struct NullArgs {
    const uint8_t *codes;
    const uint8_t *zeroCode;  // one device byte holding kZeroIdx
    const float *bg;   // [prof*32 + code]
    PairInfo *pairs;
    Rec *rec;
    ES *scratch;       // suffix values, interleaved by the 64 pairs of a null-kernel wave
    int npairs;
    float bgfs1, bgfs2;  // BACKGROUND_FRAMESHIFT_RATE(_2), as Float sums
};

// A strip-edge value of one column: the row value (W1 of the strip's top row
// for the backward, the carried row sum for the forward), its Y, and the
// exponent of the step that produced them.
struct Bnd {
    float v;
    float y;
    int e;
};

struct DPArgs {
    const ProfInfo *profs;
    const float *params;
    const float *emis;
    const PairInfo *pairs;
    const Rec *rec;
    const float *bg;      // [prof*32 + code], background codon probabilities
    const Unit *units;
    int nUnits;
    int *unitCounter;
    float *wbuf;          // W1 per step: per-wave scratch (score mode) or per pair (dump mode)
    float *w0buf;         // dump mode: forward w0 per step, per pair
    int *ebuf;            // per-step backward exponents (same indexing as wbuf / ROWS)
    int *fbuf;            // dump mode: per-step forward exponents
    int64_t waveFloats;   // score mode: scratch floats per wave
    int64_t stripSteps;   // steps per row strip in wbuf/ebuf (>= maxN + G - 1)
    int64_t escGroup;     // score mode: ints of ebuf per group (strips of stripSteps + 8, see kernels.hip)
    struct Bnd *bnd;      // strip kernels: per-group boundary records (3 x bndCols)
    int64_t bndCols;      // >= maxN + 2
    float *filterOut;     // filter mode: best W1*prefixSeed per pair, at the pair's outIdx
    PairScore *scoreOut;  // score mode, at the pair's outIdx
    int wantBwd;
};

#ifdef __HIPCC__
// Implemented in kernels.hip.
int dp_configs(int *gs, int *rs, int cap);
int dp_strip_configs(int *gs, int *rs, int cap);
hipError_t launch_null(const NullArgs &a, hipStream_t s);
hipError_t launch_dp(int G, int R, int mode, bool strips, const DPArgs &a, int grid, hipStream_t s);
hipError_t launch_zero_ints(int *p, int n, hipStream_t s);
hipError_t launch_scatter_floats(const float *src, const int *idx, float *dst, int n, hipStream_t s);

// Waves per workgroup of a DP configuration. Kernels whose emission table is
// large (192 < G*R <= 512 rows: 19-51 KB of LDS) run two (up to 384 rows) or
// four waves per workgroup on as many work units of the same profile, sharing
// one table (in strip kernels, one strip's table: the units of one profile run
// the same strips in the same order), so LDS stops limiting their occupancy;
// the host groups the units (same profile, similar lengths) and pads a
// profile's last group with empty units.
inline constexpr int dp_wg(int G, int R, bool strips, int mode) {
    (void)strips;
    return (mode == kDump || G * R <= 192 || G * R > 512) ? 1 : (G * R > 384 ? 4 : 2);
}

// The column records for the injection lane of each group (score and dump
// modes; the filter, which stores no W1, loads them per step) pass through an
// LDS ring, and so do, in the score-mode strip kernels, the neighbouring
// strip's edge values that lane reads: blocks of kRingBlock steps,
// each loaded with one entry per lane a block ahead, two blocks plus two
// entries that repeat the first ones (so that a four-step iteration reads
// without wrapping). Per entry: the (prefix seed, codes) pair and the suffix
// seed, 12 bytes, and for the edge (value, Y, exponent) another 12.
constexpr int kRingBlock = 8;
constexpr int kRing = 2 * kRingBlock + 2;
inline constexpr int dp_ring_bytes(bool strips, int mode) {
    return mode == kFilter ? 0 : kRing * (12 + ((strips && mode == kScore) ? 12 : 0));
}

// Dynamic LDS of a DP launch: the emission table (kNCode x G R floats), the
// background table (32 floats) and every group's rings.
inline constexpr size_t dp_smem(int G, int R, bool strips, int mode) {
    return sizeof(float) * ((size_t)kNCode * G * R + 32) +
           (size_t)dp_wg(G, R, strips, mode) * (64 / G) * dp_ring_bytes(strips, mode);
}

// tantan on DNA (sequence masking): bases 0..3 per letter, sequences at
// offsets; probs in the same layout; scales: lengths/16 doubles per sequence at
// scaleOff. Sixteen sequences per 64-lane wave, four lanes per sequence.
constexpr int kTantanK = 100;  // repeat-offset limit for DNA (tantan-wrapper.hh:58)
struct TantanArgs {
    const uint8_t *bases;
    const int64_t *offsets;
    const int32_t *lengths;
    const int64_t *scaleOff;
    float *probs;
    double *scales;
    const double *b2f;  // [kTantanK]
    double b2b, f2b, f2f0;
    int nseq;
};
hipError_t launch_tantan_dna(const TantanArgs &a, hipStream_t s);

hipError_t launch_copy_probe(const float *src, float *dst, int64_t n16, hipStream_t s);
#endif

}  // namespace dm
