// gfx950 kernels for DUMMEST's frameshift-aware DNA-vs-profile DP.
//
// Kernel design (one wavefront, 64 lanes):
//   * A wave takes a work unit: up to P = 64/G pairs that share one profile.
//     Each pair owns a group of G lanes (G = 16: four pairs per wave, G = 64:
//     one pair per wave). Lane k of a group holds R consecutive profile rows
//     kR .. kR+R-1, so a group covers G*R rows (M+1 <= G*R).
//   * Lanes are skewed by one DNA column: at step t lane k works on the
//     anti-diagonal D = t (forward) or D = T-1-t (backward), column c = D - k.
//     Every dependency on the neighbouring row block (W1/w0 of the row below or
//     above, the Y chain of the D state) is then a value the neighbour lane
//     produced in the previous steps, read with one DPP row/wave shift. Inside
//     a lane the D chain runs through the R rows in order.
//   * Recurrences: backward dummest-dp-backward.hh:523-548, forward
//     dummest-dp-forward.hh:112-223, score-only filter dummest-dp-backward.hh:389-427.
//   * Scaling: one power-of-two exponent per pair per anti-diagonal step,
//     raised by 64 whenever a value passes 2^100 (the reference's per-row
//     2^-64 halving at 1e30, dummest-score.hh:83-94, moved onto the step the
//     GPU computes at once). Power-of-two scaling is exact in FP32 apart from
//     flushing values more than ~2^226 below the step maximum.
//   * The codon emission table of the unit's profile sits in LDS as
//     [code][row]; per-column inputs (seeds and the codon codes) are 12-byte
//     records that enter a group at one lane and move one lane per step; in
//     score and dump mode the entering records come through a small LDS ring
//     per group, refilled a block of steps ahead, so that the loops issue no
//     per-step memory loads besides the stored W1 (the background codon
//     probability comes from a 25-entry LDS table).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <stdint.h>
#include <float.h>
#include <limits.h>
#include <type_traits>

#include "dm_layout.h"

namespace dm {

// ---------------------------------------------------------------------------
// ExpScore arithmetic (m in [0.5,1) or 0, value m*2^e), dummest-score.hh and
// dummest-batch.hh, evaluated in double exactly as the reference does.
// ---------------------------------------------------------------------------
__device__ __forceinline__ ES es_zero() { return ES{0.0f, 0}; }

__device__ __forceinline__ ES es_normalize(double m, int e) {  // dummest-score.hh:62-67
    if (m == 0.0) return es_zero();
    int ex = 0;
    m = frexp(m, &ex);
    return ES{(float)m, e + ex};
}

__device__ __forceinline__ bool es_less(const ES &a, const ES &b) {  // dummest-score.hh:70-76
    if (a.m == 0.0f) return b.m != 0.0f;
    if (b.m == 0.0f) return false;
    if (a.e != b.e) return a.e < b.e;
    return a.m < b.m;
}

__device__ __forceinline__ ES es_scale(const ES &a, double f) {  // dummest-score.hh:155-160
    if (a.m == 0.0f) return es_zero();
    if (!(f > 0.0) || !isfinite(f)) return es_zero();
    return es_normalize((double)a.m * f, a.e);
}

__device__ __forceinline__ ES es_add(ES a, ES b) {  // dummest-score.hh:169-181
    if (a.m == 0.0f) return b;
    if (b.m == 0.0f) return a;
    if (a.e >= b.e) {
        int d = a.e - b.e;
        if (d > 60) return a;
        return es_normalize((double)a.m + ldexp((double)b.m, -d), a.e);
    }
    int d = b.e - a.e;
    if (d > 60) return b;
    return es_normalize((double)b.m + ldexp((double)a.m, -d), b.e);
}

__device__ __forceinline__ ES es_mul(const ES &a, const ES &b) {  // dummest-score.hh:184-188
    if (a.m == 0.0f || b.m == 0.0f) return es_zero();
    return es_normalize((double)a.m * (double)b.m, a.e + b.e);
}

__device__ __forceinline__ ES es_pow2(int e2) { return ES{0.5f, e2 + 1}; }

// dummest-batch.hh:42-53: zero and subnormal x give (0, 0), infinity and NaN
// give ((float)x, 0), otherwise the mantissa in [0.5, 1) and its exponent.
// Written with selects so that it compiles to straight-line code.
__device__ __forceinline__ void split_double(double x, float &mOut, int &eAdd) {
    const uint64_t u = (uint64_t)__double_as_longlong(x);
    const int ef = (int)((u >> 52) & 0x7FFULL);
    const float mNorm = (float)__longlong_as_double((long long)((u & 0xFFFFFFFFFFFFFULL) | (1022ULL << 52)));
    const bool special = (ef == 0x7FF), zero = (ef == 0);
    mOut = zero ? 0.0f : (special ? (float)x : mNorm);
    eAdd = (zero || special) ? 0 : ef - 1022;
}

__device__ __forceinline__ ES es_bmul(const ES &A, const ES &B) {  // dummest-batch.hh:58-72
    const bool nz = (A.m != 0.0f) & (B.m != 0.0f);
    const double p = (double)A.m * (double)B.m;
    const int e = A.e + B.e;
    const bool lo = (p < 0.5);
    const double ps = lo ? p * 2.0 : p;
    const int es = lo ? e - 1 : e;
    return nz ? ES{(float)ps, es} : es_zero();
}

__device__ __forceinline__ ES es_badd(const ES &A, const ES &B) {  // dummest-batch.hh:78-108
    const bool a0 = (A.m == 0.0f), b0 = (B.m == 0.0f);
    const bool agebe = (A.e >= B.e);
    const float mh = agebe ? A.m : B.m;
    const float ml = agebe ? B.m : A.m;
    const int eh = agebe ? A.e : B.e;
    const int el = agebe ? B.e : A.e;
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

__device__ __forceinline__ ES es_bscale(const ES &A, double fv) {  // dummest-batch.hh:114-130
    const bool a0 = (A.m == 0.0f);
    const bool fok = (fv > 0.0) && isfinite(fv);
    float m;
    int ex;
    split_double(a0 ? 0.0 : (double)A.m * (fok ? fv : 1.0), m, ex);
    return (a0 || !fok) ? es_zero() : ES{m, A.e + ex};
}

__device__ __forceinline__ float es_to_float_clamped(const ES &a) {  // dummest-score.hh:224-230 (FP32: 120)
    if (a.m == 0.0f) return 0.0f;
    int e = a.e;
    if (e > 120) e = 120;
    return ldexpf(a.m, e);
}

// ---------------------------------------------------------------------------
// Null model, one thread per pair: dummest-dp-backward.hh:6-188 for one lane
// (the reference's lane arithmetic is independent of the batch). Writes the
// pair's record arrays and scales.
//
// One 64-lane wave takes 64 pairs. Three passes over the columns, each with its
// recurrence window in registers: the suffix (descending; its values go to the
// interleaved scratch, entry j at nullOff + 64*j, so a wave's accesses are
// coalesced), the prefix (ascending; only the last three values are kept, for
// the totals), and a final ascending pass that recomputes the prefix, forms
// both seeds and writes the records. The operations and their order are those
// of the reference, so the results are bit-identical to a single stored pass.
//
// Memory: the column loops are wave-uniform (each lane predicated to its own
// length), so the codes are staged through an LDS window that the wave fills
// with coalesced loads (64 pairs x 40 columns at a time), and the records
// are collected in LDS and written 8 columns (96 bytes; every pair's record
// block starts on a multiple of 16 records, 64-byte aligned) per pair at a
// time. Per-lane scattered byte loads and 12-byte stores would each touch a
// different cache line per lane and thrash L2.
// ---------------------------------------------------------------------------
constexpr int kNullWin = 32;                          // columns per code window step
constexpr int kNullMargin = 4;                        // window margin on either side
constexpr int kNullWinWords = (kNullWin + 2 * kNullMargin) / 4;  // 10 dwords of codes per pair
constexpr int kNullWinStride = kNullWinWords + 1;     // odd stride: conflict-free per-lane reads
constexpr int kNullFlush = 8;                         // records per pair per flush (96 bytes)
constexpr int kNullRecWords = kNullFlush * 3;         // 24 dwords
constexpr int kNullRecStride = kNullRecWords + 1;
static_assert(sizeof(Rec) == 12, "record flush assumes 12-byte records");
static_assert(kNullWin % kNullFlush == 0 && kRecBlock % kNullFlush == 0, "window and record blocks must nest");
static_assert(kNullMargin >= 2 && kNullMargin % 4 == 0, "records read codes at j-2 and j+1");

// About 10 KB per wave, so four waves share a SIMD.
struct NullShared {
    int64_t seqOff[64];
    int n[64];  // -kNullNone for lanes without a pair
    uint32_t win[64 * kNullWinStride];
    uint32_t out[64 * kNullRecStride];
};
constexpr int kNullNone = 1 << 20;

// Codes of columns [c0 - 4, c0 + 36) of every pair of the wave, kZeroIdx
// outside [0, n). Each lane assembles one dword of four codes per iteration.
// A column outside the sequence loads the kZeroIdx byte at zeroCode, so every
// load is unconditional and all of them are in flight together.
__device__ __forceinline__ void null_fill(NullShared &S, const uint8_t *codes, const uint8_t *zeroCode, int lane,
                                          int c0) {
    __syncthreads();
    uint32_t v[kNullWinWords];
#pragma unroll
    for (int k = 0; k < kNullWinWords; k++) {
        const int f = k * 64 + lane;
        const int p = f / kNullWinWords, w = f - p * kNullWinWords;
        const int np = S.n[p];
        const uint8_t *cp = codes + S.seqOff[p];
        const int col = c0 - kNullMargin + 4 * w;
        v[k] = 0;
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int c = col + q;
            const uint8_t *src = (c >= 0 && c < np) ? cp + c : zeroCode;
            v[k] |= (uint32_t)*src << (8 * q);
        }
    }
#pragma unroll
    for (int k = 0; k < kNullWinWords; k++) {
        const int f = k * 64 + lane;
        const int p = f / kNullWinWords, w = f - p * kNullWinWords;
        S.win[p * kNullWinStride + w] = v[k];
    }
    __syncthreads();
}

__global__ void __launch_bounds__(64) null_kernel(NullArgs a) {
    __shared__ NullShared S;
    const int lane = threadIdx.x;
    const int pi = blockIdx.x * 64 + lane;
    const bool valid = pi < a.npairs;
    PairInfo P{};
    if (valid) P = a.pairs[pi];
    const int n = valid ? P.n : -kNullNone;
    // The reference asserts n >= 3 (dummest-dp.hh:84). Such pairs get records
    // with zero seeds and zero-probability codes, so the DP produces nothing;
    // status reports the rejection.
    const bool ok = valid && n >= 3;
    S.seqOff[lane] = valid ? P.seqOff : 0;
    S.n[lane] = n;
    const int64_t myRecOff = valid ? P.recOff : 0;
    int N = ok ? n : 0;  // wave-uniform loop bound: the longest sequence of the wave
    for (int off = 32; off > 0; off >>= 1) N = max(N, __shfl_xor(N, off));
    int NR = valid ? n : -kRPad;  // record columns run to n + kRPad for short pairs too
    for (int off = 32; off > 0; off >>= 1) NR = max(NR, __shfl_xor(NR, off));

    const uint8_t *winb = reinterpret_cast<const uint8_t *>(S.win) + lane * 4 * kNullWinStride;
    int c0 = 0;  // base column of the current window
    auto code = [&](int j) -> int { return (int)winb[j - c0 + kNullMargin]; };

    const float *bgp = a.bg + (int64_t)(valid ? P.prof : 0) * 32;
    ES *suf = a.scratch + (valid ? P.nullOff : 0);  // suf(j) = suf[64 * j]
    const double single = (double)a.bgfs1, dbl = (double)a.bgfs2;
    const double codonKeep = 1 - single - dbl;
    const double singleEmit = single * 0.25;
    const double doubleEmit = dbl * 0.0625;
    const double frame = 1.0 / 3.0;

    // buildNullSuffix, dummest-dp-backward.hh:6-55 (window s1,s2,s3 = suf(i+1..i+3))
    {
        float tm;
        int te;
        split_double(0.5 * codonKeep, tm, te);
        ES s1 = es_pow2(0), s2 = es_zero(), s3 = es_zero();
        if (ok) suf[(int64_t)64 * n] = s1;
        for (int b = (N - 1) / kNullWin; N > 0 && b >= 0; b--) {
            c0 = b * kNullWin;
            null_fill(S, a.codes, a.zeroCode, lane, c0);
            // the next column's background value is loaded one column ahead
            float bgNext = bgp[code(c0 + kNullWin - 1)];
            for (int i = c0 + kNullWin - 1; i >= c0; i--) {
                const float bgc = bgNext;
                bgNext = bgp[code(i - 1)];
                if (!ok || i >= n) continue;
                const double codonScale = (double)bgc * codonKeep;
                const ES doubleIn = (i + 1 < n) ? s2 : es_pow2(2);
                const ES codonRaw = es_bscale(s3, codonScale);
                const ES sb = es_bscale(s1, singleEmit);
                const ES db = es_bscale(doubleIn, doubleEmit);
                const ES cb = (i + 2 < n) ? codonRaw : ES{tm, te - 2 * (n - i) + 1};
                const ES v = es_badd(cb, es_badd(sb, db));
                suf[(int64_t)64 * i] = v;
                s3 = s2;
                s2 = s1;
                s1 = v;
            }
        }
    }
    // buildNullPrefixFirstCodon, dummest-dp-backward.hh:58-100, as a function of
    // the window (p1, p2, p3) = prefix(i-1..i-3)
    // (bgc = background value of code(i - 2)).
    auto prefixStep = [&](int i, float bgc, const ES &p1, const ES &p2, const ES &p3) -> ES {
        const double codonScale = (double)bgc * codonKeep;
        const ES ps = (i > 0) ? p1 : es_zero();
        const ES pd = (i >= 2) ? p2 : es_zero();
        ES cb;
        if (i >= 3) cb = es_bscale(p3, codonScale);
        else if (i == 2) cb = es_scale(es_normalize((double)bgc, 0), frame);
        else if (i == 1) cb = es_scale(es_pow2(-4), frame);
        else cb = es_scale(es_pow2(-2), frame);
        return es_badd(cb, es_badd(es_bscale(ps, singleEmit), es_bscale(pd, doubleEmit)));
    };
    ES p1 = es_zero(), p2 = es_zero(), p3 = es_zero();
    for (int b = 0; b * kNullWin < N; b++) {
        c0 = b * kNullWin;
        null_fill(S, a.codes, a.zeroCode, lane, c0);
        float bgNext = bgp[code(c0 - 2)];
        for (int i = c0; i < c0 + kNullWin; i++) {
            const float bgc = bgNext;
            bgNext = bgp[code(i - 1)];
            if (!ok || i >= n) continue;
            const ES v = prefixStep(i, bgc, p1, p2, p3);
            p3 = p2;
            p2 = p1;
            p1 = v;
        }
    }
    // buildNullTotalAndScales, dummest-dp-backward.hh:103-129 (pre[n-1..n-3] = p1..p3)
    ES perBase = es_zero(), perBaseInv = es_zero(), s1 = es_zero(), s2 = es_zero(), sC = es_zero();
    if (ok) {
        const double keepSingle = 1.0 - single;
        const double keepBoth = keepSingle - dbl;
        const ES ts = es_scale(p2, keepSingle * 0.25);
        const ES td = es_scale(p3, keepBoth * 0.0625);
        const ES total = es_add(es_add(p1, ts), td);
        const float nullLog = (total.m == 0.0f) ? -INFINITY : (float)(log2((double)total.m) + (double)total.e);
        double perBaseLog = -(double)nullLog / (double)(float)n;
        if (!(perBaseLog < 120.0)) perBaseLog = 120.0;
        const double raw = exp2(perBaseLog);
        perBase = es_normalize(raw, 0);
        perBaseInv = es_normalize(1.0 / raw, 0);
        s1 = es_mul(perBase, es_normalize(0.25, 0));
        s2 = es_mul(es_mul(perBase, perBase), es_normalize(0.0625, 0));
        sC = es_mul(es_mul(perBase, perBase), perBase);
    }
    // buildNullSeedModels, dummest-dp-backward.hh:133-165, and the records:
    // column j holds the seeds and the codon codes at j+1 and j-2.
    {
        ES runS = es_pow2(0);
        for (int t = 0; ok && t < n - 1; t++) runS = es_bmul(runS, perBase);
        ES runP = perBase;
        ES q1 = es_zero(), q2 = es_zero(), q3 = es_zero();
        uint32_t *rec = reinterpret_cast<uint32_t *>(a.rec);
        // suf(j + 1) and the background value of code(j - 2) are loaded one
        // column ahead; the suffix index stays inside [0, n] of the lane's pair
        auto sufIdx = [&](int j) -> int64_t { return ok ? (int64_t)64 * min(max(j, 0), n) : 0; };
        ES sufNext = suf[sufIdx(-kRPad + 1)];
        float bgNext = 0.f;
        for (int jb = -kRPad; jb < NR + kRPad; jb += kNullFlush) {
            if (((jb + kRPad) % kNullWin) == 0) {
                c0 = jb;
                null_fill(S, a.codes, a.zeroCode, lane, c0);
                bgNext = bgp[code(c0 - 2)];
            }
            for (int u = 0; u < kNullFlush; u++) {
                const int j = jb + u;
                if (j >= n + kRPad) break;
                const float bgc = bgNext;
                bgNext = bgp[code(j - 1)];
                const ES sufj = sufNext;
                sufNext = suf[sufIdx(j + 2)];
                const uint32_t cds = ok ? (uint32_t)(code(j + 1) | (code(j - 2) << 8)) : (uint32_t)(kZeroIdx | (kZeroIdx << 8));
                float psd = 0.f, ssd = 0.f;
                if (ok && j >= 0 && j < n) {
                    const ES pre = prefixStep(j, bgc, q1, q2, q3);
                    q3 = q2;
                    q2 = q1;
                    q1 = pre;
                    ssd = es_to_float_clamped(es_bmul(sufj, runS));
                    runS = es_bmul(runS, perBaseInv);
                    psd = es_to_float_clamped(es_bmul(pre, runP));
                    runP = es_bmul(runP, perBase);
                }
                uint32_t *o = S.out + lane * kNullRecStride + 3 * u;
                o[0] = __float_as_uint(psd);
                o[1] = cds;
                o[2] = __float_as_uint(ssd);
            }
            __syncthreads();
            // 64 pairs x 24 dwords, 64 lanes per store: each store covers
            // contiguous 96-byte record blocks of two or three pairs
#pragma unroll 4
            for (int k = 0; k < kNullRecWords; k++) {
                const int f = k * 64 + lane;
                const int p = f / kNullRecWords, d = f - p * kNullRecWords;
                const int col = jb + d / 3;
                const int64_t ro = __shfl(myRecOff, p);
                if (col < S.n[p] + kRPad) rec[(ro + col) * 3 + d % 3] = S.out[p * kNullRecStride + d];
            }
            __syncthreads();
        }
    }
    if (valid) {
        a.pairs[pi].s1 = es_to_float_clamped(s1);
        a.pairs[pi].s2 = es_to_float_clamped(s2);
        a.pairs[pi].sC = es_to_float_clamped(sC);
        a.pairs[pi].status = ok ? 0 : 1;
    }
}

// ---------------------------------------------------------------------------
// DPP neighbour reads. Backward reads the lane below (k+1), forward the lane
// above (k-1); the group's edge lane reads 0 (bound_ctrl).
// row_shl/row_shr act inside 16-lane rows (G = 16), wave_shl/wave_shr across
// the whole wave (G = 64).
// ---------------------------------------------------------------------------
//
// G = 32 uses the whole-wave shifts, so the edge lanes of the two groups read
// the other group's values. Those reads are harmless: the group's edge rows are
// row M or past it, and row M has zero delete/enter coefficients
// (dummest-profile.hh:166-191 leaves them unset for i == length), so a finite
// neighbour value is multiplied by zero. The score-only filter never rescales
// and can hold +inf, so there the edge lane is zeroed explicitly.
template <int G> __device__ __forceinline__ float from_below(float v) {
    constexpr int ctrl = (G == 16) ? 0x101 : 0x130;
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), ctrl, 0xf, 0xf, true));
}
template <int G> __device__ __forceinline__ float from_above(float v) {
    constexpr int ctrl = (G == 16) ? 0x111 : 0x138;
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), ctrl, 0xf, 0xf, true));
}

// A column record moved one lane along the group (same DPP as the DP values).
template <int G> __device__ __forceinline__ Rec rec_from_below(const Rec &r) {
    constexpr int ctrl = (G == 16) ? 0x101 : 0x130;
    Rec o;
    o.ps = __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(r.ps), ctrl, 0xf, 0xf, true));
    o.codes = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)r.codes, ctrl, 0xf, 0xf, true);
    o.ss = __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(r.ss), ctrl, 0xf, 0xf, true));
    return o;
}
template <int G> __device__ __forceinline__ Rec rec_from_above(const Rec &r) {
    constexpr int ctrl = (G == 16) ? 0x111 : 0x138;
    Rec o;
    o.ps = __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(r.ps), ctrl, 0xf, 0xf, true));
    o.codes = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)r.codes, ctrl, 0xf, 0xf, true);
    o.ss = __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(r.ss), ctrl, 0xf, 0xf, true));
    return o;
}

// Single-rounding products and sums that the compiler may not fuse into
// multiply-adds: the matrices handed to the post-gate steps (dump mode)
// evaluate every cell's expression in the reference's order of operations
// (dummest-dp-backward.hh:523-548, dummest-dp-forward.hh:123,158-165,218-222),
// fusing only where it calls fma itself, so that each cell equals the
// reference's FP32 value (the per-pair power-of-two scaling, different from
// the reference's per-row rescaling, is exact) and the hits traced through
// near-equal anchors and max-plus ties come out the same.
__device__ __forceinline__ float mul_nc(float a, float b) {
#pragma clang fp contract(off)
    return a * b;
}
__device__ __forceinline__ float add_nc(float a, float b) {
#pragma clang fp contract(off)
    return a + b;
}

// Maximum of non-negative floats through their bit patterns, which order like
// unsigned integers: one integer max, where fmaxf would first quiet both
// operands. Only for values that cannot be NaN (the rescaled passes).
__device__ __forceinline__ float nnmax(float a, float b) {
    return __uint_as_float(max(__float_as_uint(a), __float_as_uint(b)));
}

// Exact running maximum of float values whose power-of-two offset changes
// rarely: cur holds the max at offset off; (sm, se) the exact max of all
// values retired at earlier offsets.
struct ExpMax {
    float cur;
    int off;
    float sm;
    int se;
    __device__ __forceinline__ void init(int o) { cur = 0.f; off = o; sm = 0.f; se = 0; }
    __device__ __forceinline__ void retire() {
        if (cur > 0.f) {
            int ex;
            float m = frexpf(cur, &ex);
            ES c{m, ex + off};
            if (es_less(ES{sm, se}, c)) { sm = c.m; se = c.e; }
        }
    }
    // The kept maximum seeds the running max at the new offset when it is a
    // float there; otherwise the running max starts at 0 (retire() compares
    // exactly with (sm, se), so nothing is lost). Clamping a larger maximum to
    // FLT_MAX would be wrong: FLT_MAX * 2^off exceeds sm * 2^se when se - off
    // is 128, and would then be retired as the maximum.
    __device__ __forceinline__ void rebase(int newOff) {
        retire();
        off = newOff;
        const int d = se - newOff;
        cur = (sm == 0.f || d > 127 || d < -126) ? 0.f : ldexpf(sm, d);
    }
};

// The same for products of two floats formed in double (exact, and without
// the float range limits): cur holds the max at offset off.
struct ExpMaxD {
    double cur;
    int off;
    float sm;
    int se;
    __device__ __forceinline__ void init(int o) { cur = 0.0; off = o; sm = 0.f; se = 0; }
    __device__ __forceinline__ void retire() {
        if (cur > 0.0) {
            int ex;
            const double m = frexp(cur, &ex);
            ES c{(float)m, ex + off};
            if (es_less(ES{sm, se}, c)) { sm = c.m; se = c.e; }
        }
    }
    // As ExpMax::rebase: a kept maximum outside the double range at the new
    // offset leaves the running max at 0 (never DBL_MAX, which could exceed it).
    __device__ __forceinline__ void rebase(int newOff) {
        retire();
        off = newOff;
        const int d = se - newOff;
        cur = (sm == 0.f || d > 1000 || d < -1000) ? 0.0 : ldexp((double)sm, d);
    }
};

// Exact ExpScore max over the G lanes of a group; result valid in every lane.
template <int G> __device__ __forceinline__ ES group_max(ES v) {
#pragma unroll
    for (int s = 1; s < G; s <<= 1) {
        ES o;
        o.m = __shfl_xor(v.m, s, 64);
        o.e = __shfl_xor(v.e, s, 64);
        if (es_less(v, o)) v = o;
    }
    return v;
}

constexpr float kBig = 1.2676506e30f;  // 2^100: rescale trigger
constexpr float kSmall = 6.8719477e10f; // 2^36: the live state below this is raised by 2^64
constexpr int kGhostTop = 100;          // strip ghost values are kept below 2^100 after conversion
constexpr float kGhostTopV = 0x1p100f;  // 2^kGhostTop

// True (in every lane of the group) when the whole group's live maximum is
// below kSmall and not all zero.
template <int G> __device__ __forceinline__ bool group_small(float live, int grp) {
    const unsigned long long bs = __ballot(live < kSmall);
    const unsigned long long bn = __ballot(live > 0.f);
    const unsigned long long gm = (G == 64) ? ~0ull : (((1ull << (G % 64)) - 1ull) << (grp * G));
    return ((bs & gm) == gm) && ((bn & gm) != 0ull);
}
// A buffer resource over bytes [0, size) from p, which must be the same for
// every lane of the wave. Buffer accesses at an offset past size are dropped
// (stores) or return zeros (loads) without touching memory; callers pass
// kOutOfRange for the lanes that must not access.
constexpr uint32_t kOutOfRange = 0x80000000u;
__device__ __forceinline__ __amdgpu_buffer_rsrc_t wave_rsrc(const void *p, uint32_t size) {
    const uint64_t v = (uint64_t)p;
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
    return __builtin_amdgcn_make_buffer_rsrc((void *)(((uint64_t)hi << 32) | lo), 0, (int)size, 0x00020000);
}

// The value of lane `at` of this lane's group, in every lane of the group:
// scalar reads of each group's lane (no LDS round trip).
template <int G> __device__ __forceinline__ int group_lane(int v, int grp, int at) {
    int r = __builtin_amdgcn_readlane(v, at);
#pragma unroll
    for (int g = 1; g < 64 / G; g++) {
        const int x = __builtin_amdgcn_readlane(v, g * G + at);
        r = (grp == g) ? x : r;
    }
    return r;
}
constexpr float kStep = 5.421011e-20f;  // 2^-64
constexpr int kEscGuard = 4;            // score mode: guard ints around a strip's exponents

// Four exponents moved with one 16-byte access at a 4-byte-aligned address.
typedef unsigned v2u __attribute__((ext_vector_type(2)));
typedef unsigned v3u __attribute__((ext_vector_type(3)));
typedef unsigned v4u __attribute__((ext_vector_type(4)));

struct __attribute__((aligned(4))) Int4u {
    int v0, v1, v2, v3;
    template <int I> __device__ __forceinline__ int at() const {
        if constexpr (I == 0) return v0;
        else if constexpr (I == 1) return v1;
        else if constexpr (I == 2) return v2;
        else return v3;
    }
};
constexpr int kPre = 100;               // bulk-mode prescale of stored W1 (2^-100)

// Ghost reads across row strips (STR kernels, G = 64, one pair per wave):
// boundary buffers hold one value pair per column plus the exponent of the
// step that produced it; index c+1 for columns c in [-1, n], zero outside [0, n).
__device__ __forceinline__ int bidx(int c, int n) { return c < -1 ? 0 : (c > n ? n + 1 : c + 1); }

// Two waves per SIMD for the R = 7 and R = 8 single-strip kernels and the
// strip kernels up to R = 6: they fit 256 registers without spilling, and the
// second wave hides latency (8-23% faster per step than the unconstrained
// allocation, which takes one wave).
template <int G, int R, int MODE, bool STR>
__global__ void __launch_bounds__(64 * dp_wg(G, R, STR, MODE))
__attribute__((amdgpu_waves_per_eu(((!STR && (R == 7 || R == 8)) || (STR && R <= 6)) ? 2 : 1))) dp_kernel(DPArgs a) {
    constexpr int ROWS = G * R;
    constexpr int WG = dp_wg(G, R, STR, MODE);  // waves per workgroup (sharing the emission table)
    // prefetch distance in steps (divides 4); the strip kernels keep 2, which
    // lets all but G = 32, R = 8 fit 256 registers (two waves per SIMD)
    constexpr int kPF = ((R < 8 && !STR) || (STR && G == 32 && R == 8)) ? 4 : 2;
    extern __shared__ float etab[];  // [kNCode][ROWS], the profile's background table (32), the rings
    float *bgl = etab + kNCode * ROWS;
    const int lane = threadIdx.x & 63;
    const int wv = (WG > 1) ? (int)(threadIdx.x >> 6) : 0;  // wave of the workgroup
    const int wslot = (int)blockIdx.x * WG + wv;             // this wave's scratch slot
    const int grp = lane / G;
    const int k = lane - grp * G;
    // the backward start score is wanted (a scalar, so that its tests stay scalar)
    const int wantB = __builtin_amdgcn_readfirstlane(a.wantBwd);
    // this group's rings: kRing (prefix seed, codes) pairs, then (after every
    // group's pairs) kRing suffix seeds; the score-mode strip kernels' edge
    // values, Y values and exponents follow in the same way
    constexpr int kRings = WG * (64 / G);
    // The score-only filter stores no W1: its record loads never wait behind
    // stores, and it keeps the per-step loads (the ring's LDS traffic costs it
    // more than it saves).
    constexpr bool kRecRing = MODE != kFilter;
    constexpr bool kGhostRing = STR && MODE == kScore;
    const int ringId = wv * (64 / G) + grp;
    uint2 *ringPC = reinterpret_cast<uint2 *>(bgl + 32) + ringId * kRing;
    float *ringSS = reinterpret_cast<float *>(reinterpret_cast<uint2 *>(bgl + 32) + kRings * kRing) + ringId * kRing;
    float *ringGV = ringSS + kRings * kRing;
    float *ringGY = ringGV + kRings * kRing;
    int *ringGE = reinterpret_cast<int *>(ringGY + kRings * kRing);
    // The injection lane's record of step tt of a pass (the last lane's in the
    // backward, the first lane's in the forward) sits in ring entry tt mod
    // 2 kRingBlock (entries 2 kRingBlock and 2 kRingBlock + 1 repeat entries 0
    // and 1). Lane k < kRingBlock of the group writes entry k of a block; every
    // lane reads the entry of its group. A block is written at the start of the
    // iteration before it is first read and was loaded a block earlier, so the
    // loops' memory waits no longer cover the stores and loads of the steps
    // just before them. The passes run one after the other and share the ring.
    auto ringPut = [&](int base, const Rec &r) {
        if (k < kRingBlock) {
            ringPC[base + k] = uint2{__float_as_uint(r.ps), r.codes};
            ringSS[base + k] = r.ss;
            if (base == 0 && k < 2) {
                ringPC[2 * kRingBlock + k] = uint2{__float_as_uint(r.ps), r.codes};
                ringSS[2 * kRingBlock + k] = r.ss;
            }
        }
    };
    auto ringGet = [&](int e) {
        const uint2 pc = ringPC[e];
        return Rec{__uint_as_float(pc.x), pc.y, ringSS[e]};
    };
    // the strip edge values of step tt (score-mode strip kernels), alike
    auto gringPut = [&](int base, const Bnd &g) {
        if (k < kRingBlock) {
            ringGV[base + k] = g.v;
            ringGY[base + k] = g.y;
            ringGE[base + k] = g.e;
            if (base == 0 && k < 2) {
                ringGV[2 * kRingBlock + k] = g.v;
                ringGY[2 * kRingBlock + k] = g.y;
                ringGE[2 * kRingBlock + k] = g.e;
            }
        }
    };
    auto gringGet = [&](int e) { return Bnd{ringGV[e], ringGY[e], ringGE[e]}; };
    __shared__ int sItem;
    for (;;) {
        int u;
        if (WG == 1) {
            int got = 0;
            if (lane == 0) got = atomicAdd(a.unitCounter, 1);
            u = __builtin_amdgcn_readfirstlane(got);
            if (u >= a.nUnits) break;
        } else {
            // one item = WG consecutive units of one profile (the host pairs them)
            __syncthreads();  // the previous item is done with sItem and the table
            if (threadIdx.x == 0) sItem = atomicAdd(a.unitCounter, 1);
            __syncthreads();
            const int item = sItem;
            if (item * WG >= a.nUnits) break;
            u = item * WG + wv;
        }
        const Unit U = a.units[u];
        const ProfInfo PI = a.profs[U.prof];
        const int M = PI.M;
        const int S = STR ? (M + ROWS) / ROWS : 1;  // row strips of ROWS rows
        const bool active = grp < U.count;
        const int pi = U.first + (active ? grp : 0);
        const PairInfo PR = a.pairs[pi];
        const int n = active ? PR.n : 0;
        const Rec *recs = a.rec + PR.recOff;
        const float s1 = PR.s1, s2 = PR.s2, sC = PR.sC;
        // Lanes of a strip whose rows all lie past row M never feed a live row
        // (rows past M and row M itself have zero delete/enter coefficients), so
        // score and filter modes start a strip's anti-diagonals at n + live - 1
        // rather than n + G - 1, live = the lanes that hold a row <= M (dump mode
        // keeps G: the host reads its matrices in that layout).
        auto liveLanes = [&](int strip) {
            if (MODE == kDump) return G;
            const int l = (M + 1 - strip * ROWS + R - 1) / R;
            return l < G ? l : G;
        };
        const int64_t strideS = (int64_t)a.stripSteps * ROWS;  // floats per strip of W1 steps
        float *wsc0;
        int *esc0;
        if (MODE == kDump) {
            wsc0 = a.wbuf + PR.bufOff;
            esc0 = a.ebuf + PR.bufOff / ROWS;
        } else {
            // strip-major: strip s of group g at (s (64/G) + g) stripSteps ROWS
            wsc0 = a.wbuf + (int64_t)wslot * a.waveFloats + (int64_t)grp * strideS;
            // per strip stripSteps + 8 ints: 4 guard ints on either side of
            // anti-diagonals 0..stripSteps-1 (exponents move four at a time)
            esc0 = a.ebuf + ((int64_t)wslot * (64 / G) + grp) * a.escGroup + kEscGuard;
        }
        const int64_t escStrip = (MODE == kDump) ? a.stripSteps : a.stripSteps + 2 * kEscGuard;
        const int64_t wStrip = (MODE == kDump) ? strideS : strideS * (64 / G);  // W1 floats between strips
        // strip boundary buffers (per group), one (value, Y, exponent) record per
        // column, n + 2 entries each: two parities for the backward, one for the forward
        Bnd *bB = nullptr, *bF = nullptr;
        if (STR) {
            const int64_t nb = (int64_t)a.bndCols;
            const int64_t gslot = (int64_t)wslot * (64 / G) + grp;
            bB = a.bnd + gslot * 3 * nb;  // [2][nb]
            bF = bB + 2 * nb;
            for (int c = k; c < 3 * nb; c += G) bB[c] = Bnd{0.f, 0.f, 0};
        }

        // Score mode's W1 block of a step (G R floats per group): lane k's rows
        // in chunks of four, chunk j holding rows 4j..4j+3 of every lane of the
        // group side by side (G x 4 floats, the last chunk narrower), so that
        // each wide store or load instruction covers one contiguous run.
        auto w1Off = [&](int kk, int q) {
            const int j = q >> 2, wj = (R - 4 * j < 4) ? R - 4 * j : 4;
            return G * 4 * j + kk * wj + (q & 3);
        };
        // Lanes of a run of chunk j: L_j lanes of chunk j fill about one 128-byte line.
        auto runLanes = [](int j) {
            const int wj = (R - 4 * j < 4) ? R - 4 * j : 4;
            const int l = (wj >= 3) ? 8 : (wj == 2 ? 16 : 32);
            return l < G ? l : G;
        };
        float enterC[R], d0[R], d1s2[R], d2s1[R], a1s1[R], a2s2[R], eps[R];
        float fbest = 0.f;  // filter mode: max W1 * prefixSeed
        ExpMaxD bwdMax;  // max of W1 * prefixSeed, at offset Eb
        bwdMax.init(0);
        const float *etk = etab + k * R;

        // Per-strip setup: LDS emission rows and per-row coefficients
        // (dummest-dp-backward.hh:481-491, dummest-dp-forward.hh:72-80).
        // fill: load the strip's emission table (the forward's first strip is the
        // backward's last, whose table is still in LDS).
        auto loadStrip = [&](int strip, float *a0v, float *b0v, bool fill) {
            const int rb = strip * ROWS;
            if (fill) {
                __syncthreads();  // the previous table is no longer read
                // Thread x fills rows x, x + 64 WG, ...: a row's 25 values come with
                // seven 16-byte loads (the HBM rows hold 32 floats), and every load
                // of a batch of rows is issued before the first LDS write, so the
                // fill waits on memory once per batch rather than once per value.
                constexpr int kPer = (ROWS + 64 * WG - 1) / (64 * WG);
                constexpr int kBatch = (R >= 12) ? 1 : 2;
#pragma unroll 1
                for (int i0 = 0; i0 < kPer; i0 += kBatch) {
                    float4 v[kBatch][7];
#pragma unroll
                    for (int i = 0; i < kBatch; i++) {
                        const int rr = (int)threadIdx.x + (i0 + i) * 64 * WG;
                        const int r = min(rb + rr, M);  // rows past M read row M, stored as 0
                        const float4 *src = (const float4 *)(a.emis + PI.emOff + (int64_t)r * kEmStride);
#pragma unroll
                        for (int w = 0; w < 7; w++) v[i][w] = src[w];
                    }
#pragma unroll
                    for (int i = 0; i < kBatch; i++) {
                        const int rr = (int)threadIdx.x + (i0 + i) * 64 * WG;
                        if (i0 + i < kPer && rr < ROWS) {
                            const bool live = rb + rr <= M;
#pragma unroll
                            for (int c = 0; c < kNCode; c++) {
                                const float4 q4 = v[i][c >> 2];
                                const float x = (c & 3) == 0 ? q4.x : (c & 3) == 1 ? q4.y : (c & 3) == 2 ? q4.z : q4.w;
                                etab[c * ROWS + rr] = live ? x : 0.f;
                            }
                        }
                    }
                }
                if (threadIdx.x < 32) bgl[threadIdx.x] = a.bg[(int64_t)U.prof * 32 + threadIdx.x];
                __syncthreads();
            }
#pragma unroll
            for (int q = 0; q < R; q++) {
                const int r = rb + k * R + q;
                const bool ok = active && r <= M;
                const float *pp = a.params + PI.parOff + (int64_t)(ok ? r : 0) * 11;
                enterC[q] = ok ? pp[10] * sC : 0.f;
                d0[q] = ok ? pp[6] : 0.f;
                d1s2[q] = ok ? pp[7] * s2 : 0.f;
                d2s1[q] = ok ? pp[8] * s1 : 0.f;
                a0v[q] = ok ? pp[0] : 0.f;
                a1s1[q] = ok ? pp[1] * s1 : 0.f;
                a2s2[q] = ok ? pp[2] * s2 : 0.f;
                b0v[q] = ok ? pp[3] : 0.f;
                eps[q] = ok ? pp[9] : 0.f;
            }
        };

        // ------------------------------------------------------------- backward
        for (int strip = S - 1; strip >= 0; strip--) {
            const int T = U.maxN + liveLanes(strip) - 1;  // steps of this strip's pass
            const int Tmine = n + liveLanes(strip) - 1;   // this pair's
            const int Tb = Tmine;  // anti-diagonal of the backward's first step
            float a0sC[R], b0sC[R];
            loadStrip(strip, a0sC, b0sC, true);
            // rows of this lane that exist (r <= M): a prefix of its R rows; the
            // seed enters only there
            const int nValid = active ? max(0, min(R, M - (strip * ROWS + k * R) + 1)) : 0;
#pragma unroll
            for (int q = 0; q < R; q++) {  // backward alpha0Coeff/beta0Coeff carry sC
                a0sC[q] *= sC;
                b0sC[q] *= sC;
            }
            float *wsc = wsc0 + (int64_t)strip * wStrip;
            int *esc = esc0 + (int64_t)strip * escStrip;
            const bool below = STR && strip < S - 1;  // a strip below feeds lane G-1
            const bool above = STR && strip > 0;      // lane 0 feeds the strip above
            const int par = strip & 1;
            Bnd *bOut = STR ? bB + par * a.bndCols : nullptr;
            const Bnd *bIn = STR ? bB + (par ^ 1) * a.bndCols : nullptr;

            float H[4][R], Zr[4][R], Yc[R], GW[4], LV[4];
#pragma unroll
            for (int s4 = 0; s4 < 4; s4++) {
                GW[s4] = 0.f;
                LV[s4] = 0.f;
#pragma unroll
                for (int q = 0; q < R; q++) { H[s4][q] = 0.f; Zr[s4][q] = 0.f; }
            }
#pragma unroll
            for (int q = 0; q < R; q++) Yc[q] = 0.f;
            int Eb = 0;      // backward exponent of the current step
            float rs = 1.f;  // 2^-Eb, the seed's row scale

            // Multiply the pair's whole backward state by 2^s (exact, or a flush
            // to zero for large negative s) and move the exponent to match.
            auto scaleB = [&](int s) {
                const float f = ldexpf(1.f, s);
#pragma unroll
                for (int s4 = 0; s4 < 4; s4++) {
                    GW[s4] *= f;
                    LV[s4] *= f;
#pragma unroll
                    for (int q = 0; q < R; q++) { H[s4][q] *= f; Zr[s4][q] *= f; }
                }
#pragma unroll
                for (int q = 0; q < R; q++) Yc[q] *= f;
                Eb -= s;
                rs = ldexpf(1.f, -Eb);
                if (wantB) bwdMax.rebase(Eb);
            };

            // Column records: lane k uses at step t the record lane k+1 used at
            // step t-1, so a record enters the group at its last lane (loaded kPF
            // steps ahead there) and moves up one lane per step by DPP.
            auto brecAt = [&](int tt, int kk) {
                const int c = Tb - 1 - tt - kk;
                return recs[c < -kRPad ? -kRPad : (c > n + kRPad - 1 ? n + kRPad - 1 : c)];
            };
            const bool injB = (k == G - 1);
            Rec pend{}, rnext{};
            Rec rq[kPF];  // filter mode: the last lane's records of the next steps (slot = step mod kPF)
            if constexpr (kRecRing) {
                ringPut(0, brecAt(k & (kRingBlock - 1), G - 1));
                pend = brecAt(kRingBlock + (k & (kRingBlock - 1)), G - 1);  // the next block
                rnext = ringGet(1);  // the last lane's record of the next step
            } else {
#pragma unroll
                for (int f = 1; f <= kPF; f++) rq[f % kPF] = brecAt(f, G - 1);
            }
            Rec rcur = brecAt(0, k);
            // strip kernels: the strip below's top row (W1, Y, exponent), also ahead
            Bnd gq4[kPF];
            // (loaded in every strip so that the load is not conditional, which
            // would make the compiler wait for all memory operations each step;
            // used only when there is a strip below)
            // Only the last lane uses the value; score and filter modes load its
            // column in every lane of the group (one address per group, so the
            // load moves one record rather than G), dump mode each lane's own
            // column (whose values also enter the group's scale-up test).
            auto ghostAt = [&](int tt) -> Bnd {
                if (STR) return bIn[bidx(Tb - 1 - tt - (MODE == kDump ? k : G - 1), n)];
                return Bnd{0.f, 0.f, 0};
            };
            // In score mode the group's edge value of step tt comes through
            // the ghost ring, entry tt mod 2 kRingBlock, like the records.
            Bnd gpend{0.f, 0.f, 0}, gnext{0.f, 0.f, 0};
            if constexpr (kGhostRing) {
                gringPut(0, ghostAt(k & (kRingBlock - 1)));
                gpend = ghostAt(kRingBlock + (k & (kRingBlock - 1)));
                gnext = gringGet(0);  // step 0
            } else {
#pragma unroll
                for (int f = 0; f < kPF; f++) gq4[f] = ghostAt(f);
            }
            // emission values of the next step, read from LDS one step ahead
            float eq[R];
#pragma unroll
            for (int q = 0; q < R; q++) eq[q] = etk[(rcur.codes & 255) * ROWS + q];
            float gq = bgl[rcur.codes & 255];  // g[c+1] of the next step
            // lanes whose rows all lie past row M store and load no W1
            const bool rowsLive = strip * ROWS + k * R <= M;
            // Score mode: chunk j of a step's block is stored by runs of L_j
            // lanes (L_j lanes of chunk j fill about one 128-byte line); a run
            // is skipped when none of its lanes holds a cell of the matrix: all
            // its columns lie outside [0, n) (the skew), or all its rows past M.
            // Skipping whole runs keeps every write a whole line.
            int runLo[(R + 3) / 4], runSpan[(R + 3) / 4];  // run of chunk j: columns D - runLo - runSpan + 1 .. D - runLo
#pragma unroll
            for (int j = 0; j * 4 < R; j++) {
                const int k0 = k & ~(runLanes(j) - 1);
                runLo[j] = k0;
                runSpan[j] = (active && strip * ROWS + k0 * R <= M) ? n + runLanes(j) - 1 : 0;
            }

            // W1 of step tt (ring slot sl) and its exponent into the scratch.
            // Score mode stores the runs of lanes described above for every
            // step of the pair (anti-diagonals 0..Tb-1, inside the group's own
            // scratch slot). The forward reads the same runs (the same test on
            // the same anti-diagonal) and gets zeros for the others; within a
            // run it gives no weight to the lanes whose column is at or past n
            // or whose rows all lie past M (at columns before 0, w0 is zero).
            // Stored cells outside the matrix
            // cannot raise the anchor maximum either: columns outside the
            // sequence meet a zero on one side, rows past M in a live lane hold
            // seed-only values that never exceed row M's at the same column.
            // Dump mode stores the live cells of active pairs only (its buffers
            // are the pairs' own). Score mode writes the exponents of a four-step
            // iteration with one store (storeE4) when its W1 blocks are stored
            // late, else one per step.
            // Score mode stores through a buffer resource over the wave's blocks
            // of this strip: a skipped run's offset lies past its end, so the
            // store writes nothing and needs no branch.
            const __amdgpu_buffer_rsrc_t w1s = wave_rsrc(a.wbuf + (int64_t)wslot * a.waveFloats + (int64_t)strip * wStrip,
                                                         MODE == kScore ? (uint32_t)(wStrip * sizeof(float)) : 0u);
            uint32_t boff[(R + 3) / 4];  // byte offset of this lane's part of chunk j at anti-diagonal 0
#pragma unroll
            for (int j = 0; j * 4 < R; j++) boff[j] = (uint32_t)(((int64_t)grp * strideS + w1Off(k, 4 * j)) * sizeof(float));
            auto storeW1 = [&](int tt, int sl, bool withExp) {
                const int D = Tb - 1 - tt;
                if (MODE == kScore) {
                    const uint32_t sb = (uint32_t)D * (uint32_t)(ROWS * sizeof(float));
#pragma unroll
                    for (int j = 0; j * 4 < R; j++) {
                        // some column of the run lies in [0, n) (rows checked in
                        // runSpan; D < 0 fails the test too)
                        const uint32_t off = ((unsigned)(D - runLo[j]) < (unsigned)runSpan[j]) ? boff[j] + sb : kOutOfRange;
                        const int wj = (R - 4 * j < 4) ? R - 4 * j : 4;
                        const float *h = &H[sl][4 * j];
                        if (wj == 4) {
                            const v4u v{__float_as_uint(h[0]), __float_as_uint(h[1]), __float_as_uint(h[2]), __float_as_uint(h[3])};
                            __builtin_amdgcn_raw_buffer_store_b128(v, w1s, off, 0, 0);
                        } else if (wj == 3) {
                            const v3u v{__float_as_uint(h[0]), __float_as_uint(h[1]), __float_as_uint(h[2])};
                            __builtin_amdgcn_raw_buffer_store_b96(v, w1s, off, 0, 0);
                        } else if (wj == 2) {
                            const v2u v{__float_as_uint(h[0]), __float_as_uint(h[1])};
                            __builtin_amdgcn_raw_buffer_store_b64(v, w1s, off, 0, 0);
                        } else {
                            __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(h[0]), w1s, off, 0, 0);
                        }
                    }
                    if (withExp && D >= 0) esc[D] = Eb;  // every lane of the group writes the same value
                } else {
                    const int c = D - k;
                    if (D >= 0 && active) {
                        if (rowsLive && c >= 0 && c < n) {
                            float *dst = wsc + (int64_t)D * ROWS + k * R;
#pragma unroll
                            for (int q = 0; q < R; q++) dst[q] = H[sl][q];
                        }
                        if (k == 0) esc[D] = Eb;
                    }
                }
            };

            // The exponent of the four steps tt0..tt0+3 (anti-diagonals D0-3..D0,
            // D0 = Tb-1-tt0), all the current one: their W1 blocks are stored in
            // its units. The guard ints before anti-diagonal 0 take the entries
            // of a last iteration that runs past it.
            auto storeE4 = [&](int tt0) {
                const int D0 = Tb - 1 - tt0;
                if (D0 >= 0) *(Int4u *)(esc + D0 - 3) = Int4u{Eb, Eb, Eb, Eb};
            };

            // DEFER (score mode, main loop): the step's W1 is stored at the start
            // of the next four-step iteration rather than at once (see below)
            auto bstep = [&](auto PHc, auto DEFc, int t) {
                constexpr int PH = decltype(PHc)::value;
                constexpr bool DEFER = decltype(DEFc)::value;
                constexpr int S1 = (PH + 3) & 3, S2 = (PH + 2) & 3, S3 = (PH + 1) & 3, S4 = PH;
                const int c = Tb - 1 - t - k;  // column: D = Tb-1-t
                const Rec rc = rcur;
                const float rcg = gq;
                // the next step's record (the lane below's, or a new one at the last
                // lane) and its emission values, read now so that they arrive in time
                Rec nx = rec_from_below<G>(rcur);
                if constexpr (kRecRing) {
                    if (injB) nx = rnext;
                    rnext = ringGet(((t - PH) & (2 * kRingBlock - 1)) + PH + 2);  // step t + 2
                } else {
                    constexpr int SN = (PH + 1) % kPF;
                    if (injB) nx = rq[SN];
                    rq[SN] = brecAt(t + 1 + kPF, G - 1);  // same address in every lane of the group
                }
                float eqN[R];
#pragma unroll
                for (int q = 0; q < R; q++) eqN[q] = etk[(nx.codes & 255) * ROWS + q];
                const float gqN = bgl[nx.codes & 255];
                float gw = 0.f, gy = 0.f;
                Bnd gcur{0.f, 0.f, 0};
                if constexpr (kGhostRing) {
                    gcur = gnext;
                    gnext = gringGet(((t - PH) & (2 * kRingBlock - 1)) + PH + 1);  // step t + 1
                } else if (STR) {
                    gcur = gq4[PH % kPF];
                    gq4[PH % kPF] = ghostAt(t + kPF);
                }
                if (STR && below) {
                    // last lane: the strip below's top row from the boundary buffer.
                    // When that row sits far above this strip's scale, raise the
                    // exponent first (before any neighbour value is read).
                    // The exponent is raised when max(v, y) 2^(e - Eb) >= 2^kGhostTop
                    // on the last lane (frexp exponent + e - Eb > kGhostTop): tested
                    // on the converted values, the exact amount worked out only then.
                    const float rw = gcur.v, ry = gcur.y;
                    const int ge = gcur.e;
                    gw = ldexpf(rw, ge - Eb);
                    gy = ldexpf(ry, ge - Eb);
                    if (MODE != kFilter && __builtin_expect(__any(k == G - 1 && fmaxf(gw, gy) >= kGhostTopV), 0)) {
                        int need = 0;
                        if (k == G - 1 && (rw > 0.f || ry > 0.f)) {
                            int ex;
                            (void)frexpf(fmaxf(rw, ry), &ex);
                            need = ex + ge - Eb - kGhostTop;
                        }
                        need = group_lane<G>(need, grp, G - 1);  // the group's last lane
                        if (need > 0) scaleB(-((need + 63) & ~63));
                        gw = ldexpf(rw, ge - Eb);
                        gy = ldexpf(ry, ge - Eb);
                    }
                }
                // lane below (k+1): its row 0 at columns c+1, c+2, c+3 and its Y at c
                float bw1 = from_below<G>(H[S2][0]);
                float bw2 = from_below<G>(H[S3][0]);
                float bw3 = from_below<G>(H[S4][0]);
                float bY = from_below<G>(Yc[0]);
                if (G == 32 && MODE == kFilter && k == G - 1) { bw1 = 0.f; bw2 = 0.f; bw3 = 0.f; bY = 0.f; }
                if (STR) {
                    GW[S4] = gw;
                    if (k == G - 1) {
                        bw1 = GW[S1];
                        bw2 = GW[S2];
                        bw3 = GW[S3];
                        bY = gy;
                    }
                }
                const float ssr = rc.ss * rs;
                const float psp = rc.ps;
                const double pspD = (double)rc.ps;
                float smax = 0.f;
#pragma unroll
                for (int q = R - 1; q >= 0; q--) {
                    const float e = eq[q];
                    const float zg = Zr[S3][q] * rcg;
                    float w;
                    if constexpr (MODE == kDump) {
                        // n3 e enter + Y d0 + n2 d1 + n1 d2 + Z g a0 + c2 a2 + c1 a1 + seed, left to right
                        w = mul_nc(mul_nc(bw3, e), enterC[q]);
                        w = add_nc(w, mul_nc(bY, d0[q]));
                        w = add_nc(w, mul_nc(bw2, d1s2[q]));
                        w = add_nc(w, mul_nc(bw1, d2s1[q]));
                        w = add_nc(w, mul_nc(zg, a0sC[q]));
                        w = add_nc(w, mul_nc(H[S2][q], a2s2[q]));
                        w = add_nc(w, mul_nc(H[S1][q], a1s1[q]));
                        w = add_nc(w, q < nValid ? ssr : 0.f);
                    } else {
                        const float eE = e * enterC[q];
                        // score mode adds the seed to the rows past M too: their
                        // coefficients are zero, so they feed nothing, and their
                        // seed-only values never exceed a real row's (every real
                        // row holds the seed plus non-negative terms)
                        w = fmaf(bw3, eE, (MODE == kScore || q < nValid) ? ssr : 0.f);
                        w = fmaf(bw2, d1s2[q], w);
                        w = fmaf(bw1, d2s1[q], w);
                        w = fmaf(zg, a0sC[q], w);
                        w = fmaf(H[S2][q], a2s2[q], w);
                        w = fmaf(H[S1][q], a1s1[q], w);
                        w = fmaf(bY, d0[q], w);
                    }
                    const float y = fmaf(eps[q], bY, w);
                    const float z = fmaf(b0sC[q], zg, w);
                    if (MODE == kFilter) fbest = fmaxf(fbest, w * psp);  // +inf * 0 may give NaN: fmaxf skips it
                    else smax = nnmax(smax, nnmax(y, z));  // y, z >= w
                    // row q at columns c+1, c+2, c+3 for row q-1 (same lane, same column)
                    bw1 = H[S1][q];
                    bw2 = H[S2][q];
                    bw3 = H[S3][q];
                    bY = y;
                    H[S4][q] = w;
                    Zr[S4][q] = z;
                    Yc[q] = y;
                }
                rcur = nx;
#pragma unroll
                for (int q = 0; q < R; q++) eq[q] = eqN[q];
                gq = gqN;
                if (MODE != kFilter && wantB) {
                    // backward start score: max W1 * prefixSeed. The seed is
                    // shared by the lane's rows and non-negative, and a product
                    // of two floats is exact in double, so the max of the
                    // products is the product of the float max.
                    float hm = H[S4][0];
#pragma unroll
                    for (int q = 1; q < R; q++) hm = nnmax(hm, H[S4][q]);
                    bwdMax.cur = fmax(bwdMax.cur, (double)hm * pspD);
                }
                if (STR && above && k == 0 && c >= 0 && c < n) {
                    bOut[c + 1] = Bnd{H[S4][0], Yc[0], Eb};
                }
                if (MODE == kDump || (MODE == kScore && !DEFER)) storeW1(t, S4, true);
                if (MODE != kFilter) {
                    // One exponent per pair per step: lower the scale by 2^-64 when
                    // a value passes 2^100; every 4 steps, raise it by 2^64 when the
                    // whole live state (the last 4 steps) of the pair is below 2^36
                    // and the exponent is at least 64, so regions that decay after
                    // a strong hit keep their precision.
                    LV[S4] = nnmax(smax, nnmax(gw, gy));
                    const bool big = smax > kBig;
                    if (__builtin_expect(__any(big), 0)) {
                        bool gbig = big;
#pragma unroll
                        for (int s = 1; s < G; s <<= 1) gbig |= __shfl_xor((int)gbig, s, 64) != 0;
                        if (gbig) scaleB(-64);
                    }
                    if (PH == 3) {  // never below exponent 0: the seeds enter at 2^-Eb <= 1
                        const float live = nnmax(nnmax(LV[0], LV[1]), nnmax(LV[2], LV[3]));
                        if (__builtin_expect(group_small<G>(live, grp), 0) && Eb >= 64) scaleB(64);
                    }
                }
            };
            int t = 0;
            for (; t + 4 <= T; t += 4) {
                // the ring block that starts at step t + 4, and the load of the one after it
                if (kRecRing && ((t + 4) & (kRingBlock - 1)) == 0) {
                    ringPut((t + 4) & (2 * kRingBlock - 1), pend);
                    pend = brecAt(t + 4 + kRingBlock + (k & (kRingBlock - 1)), G - 1);
                    if constexpr (kGhostRing) {
                        gringPut((t + 4) & (2 * kRingBlock - 1), gpend);
                        gpend = ghostAt(t + 4 + kRingBlock + (k & (kRingBlock - 1)));
                    }
                }
                // Score mode stores an iteration's four W1 blocks (still in the
                // ring, in the units of the current exponent) when the next one
                // starts: the loop's waits for prefetched loads also wait for the
                // stores issued before those loads, and stores issued a whole
                // iteration earlier have mostly completed by then.
                if (MODE == kScore && t > 0) {
#pragma unroll
                    for (int j = 0; j < 4; j++) storeW1(t - 4 + j, j, false);
                    storeE4(t - 4);
                }
                bstep(std::integral_constant<int, 0>{}, std::bool_constant<MODE == kScore>{}, t);
                bstep(std::integral_constant<int, 1>{}, std::bool_constant<MODE == kScore>{}, t + 1);
                bstep(std::integral_constant<int, 2>{}, std::bool_constant<MODE == kScore>{}, t + 2);
                bstep(std::integral_constant<int, 3>{}, std::bool_constant<MODE == kScore>{}, t + 3);
            }
            if (MODE == kScore && t > 0) {
#pragma unroll
                for (int j = 0; j < 4; j++) storeW1(t - 4 + j, j, false);
                storeE4(t - 4);
            }
            if (t < T) bstep(std::integral_constant<int, 0>{}, std::false_type{}, t++);
            if (t < T) bstep(std::integral_constant<int, 1>{}, std::false_type{}, t++);
            if (t < T) bstep(std::integral_constant<int, 2>{}, std::false_type{}, t++);
            if (STR) __syncthreads();  // boundary written above is read by the next strip
            if (wantB && MODE != kFilter) bwdMax.rebase(0);  // next strip starts at exponent 0
        }
        if (MODE == kFilter) {
            // group max of the plain float best (no rescale in the filter)
            float v = fbest;
#pragma unroll
            for (int s = 1; s < G; s <<= 1) v = fmaxf(v, __shfl_xor(v, s, 64));
            if (active && k == 0) a.filterOut[PR.outIdx] = v;
            continue;
        }
        __syncthreads();  // W1 steps and exponents stored above are read below
        ES bwdBest = es_zero();
        if (wantB) {
            bwdMax.retire();
            bwdBest = group_max<G>(ES{bwdMax.sm, bwdMax.se});
        }

        // -------------------------------------------------------------- forward
        ExpMax midMax;
        midMax.init(kPre);
        for (int strip = 0; strip < S; strip++) {
            const int T = U.maxN + liveLanes(strip) - 1;  // steps of this strip's pass
            const int Tmine = n + liveLanes(strip) - 1;   // this pair's
            float a0[R], b0[R];
            loadStrip(strip, a0, b0, strip > 0);
            const float *wsc = wsc0 + (int64_t)strip * wStrip;
            const int *esc = esc0 + (int64_t)strip * escStrip;
            const bool above = STR && strip > 0;      // lane 0 reads the strip above
            const bool below = STR && strip < S - 1;  // last lane feeds the strip below
            float Wf[4][R], Zf[4][R], LVf[4];
#pragma unroll
            for (int s4 = 0; s4 < 4; s4++) {
                LVf[s4] = 0.f;
#pragma unroll
                for (int q = 0; q < R; q++) { Wf[s4][q] = 0.f; Zf[s4][q] = 0.f; }
            }
            float lastN = 0.f, lastY = 0.f;  // row R-1 outputs of the previous step
            int Ef = 0;
            float rf = 1.f;
            int EbPrev = 0;
            bool eChg = false;  // score mode: the current block has an exponent change
            if (MODE == kScore) midMax.rebase(Ef + EbPrev + kPre);
            float *w0d = (MODE == kDump) ? a.w0buf + PR.bufOff + (int64_t)strip * strideS : nullptr;
            int *efd = (MODE == kDump) ? a.fbuf + PR.bufOff / ROWS + (int64_t)strip * a.stripSteps : nullptr;

            // Multiply the pair's forward state by 2^s and move the exponent.
            auto scaleF = [&](int s) {
                const float f = ldexpf(1.f, s);
#pragma unroll
                for (int s4 = 0; s4 < 4; s4++) {
                    LVf[s4] *= f;
#pragma unroll
                    for (int q = 0; q < R; q++) { Wf[s4][q] *= f; Zf[s4][q] *= f; }
                }
                lastN *= f;
                lastY *= f;
                Ef -= s;
                rf = ldexpf(1.f, -Ef);
                if (MODE == kScore) midMax.rebase(Ef + EbPrev + kPre);
            };

            // Column records and the stored W1 block of a step (with its
            // exponent) are loaded kPF steps ahead of their use.
            constexpr int kNoStep = INT_MIN;
            // Column records move down one lane per step (lane k uses at step t
            // the record lane k-1 used at t-1); they enter at the group's first lane.
            auto frecAt = [&](int tt, int kk) {
                const int c = tt - kk;
                return recs[c < -kRPad ? -kRPad : (c > n + kRPad - 1 ? n + kRPad - 1 : c)];
            };
            const bool injF = (k == 0);
            float Wq[kPF][R];
            int Eq[kPF];
            // Loads are unconditional (from a clamped step; runs the backward did
            // not store read zeros, see below) so that no select waits for them;
            // which of them count is decided at use.
            const int Tw = Tmine;
            const int nW1 = (active && strip * ROWS + k * R <= M) ? n : 0;  // columns of this lane with W1
            // Score mode reads the wave's blocks of this strip through a buffer
            // resource. A run of lanes whose store the backward skipped (no
            // column in [0, n), or all rows past M) reads with an offset past
            // the resource's end: the load returns zeros and fetches nothing.
            // the host keeps a strip of the wave's blocks below 2^31 bytes (score mode)
            const __amdgpu_buffer_rsrc_t w1r = wave_rsrc(a.wbuf + (int64_t)wslot * a.waveFloats + (int64_t)strip * wStrip,
                                                         MODE == kScore ? (uint32_t)(wStrip * sizeof(float)) : 0u);
            uint32_t foff[(R + 3) / 4];  // byte offset of this lane's part of chunk j at anti-diagonal 0
            int flo[(R + 3) / 4], fspan[(R + 3) / 4];  // the backward's runLo / runSpan
            if constexpr (MODE == kScore) {
#pragma unroll
                for (int j = 0; j * 4 < R; j++) {
                    foff[j] = (uint32_t)(((int64_t)grp * strideS + w1Off(k, 4 * j)) * sizeof(float));
                    flo[j] = k & ~(runLanes(j) - 1);
                    fspan[j] = (active && strip * ROWS + flo[j] * R <= M) ? n + runLanes(j) - 1 : 0;
                }
            }
            auto loadW1 = [&](int tt, float *dst, int &e) {
                const int ts = min(tt, Tw - 1);  // Tw = n + live - 1 >= 1
                if constexpr (MODE == kScore) {
                    const uint32_t sb = (uint32_t)ts * (uint32_t)(ROWS * sizeof(float));
#pragma unroll
                    for (int j = 0; j * 4 < R; j++) {
                        const uint32_t off = ((unsigned)(ts - flo[j]) < (unsigned)fspan[j]) ? foff[j] + sb : kOutOfRange;
                        const int wj = (R - 4 * j < 4) ? R - 4 * j : 4;
                        if (wj == 4) {
                            const auto v = __builtin_amdgcn_raw_buffer_load_b128(w1r, off, 0, 0);
#pragma unroll
                            for (int q = 0; q < 4; q++) dst[4 * j + q] = __uint_as_float(v[q]);
                        } else if (wj == 3) {
                            const auto v = __builtin_amdgcn_raw_buffer_load_b96(w1r, off, 0, 0);
#pragma unroll
                            for (int q = 0; q < 3; q++) dst[4 * j + q] = __uint_as_float(v[q]);
                        } else if (wj == 2) {
                            const auto v = __builtin_amdgcn_raw_buffer_load_b64(w1r, off, 0, 0);
                            dst[4 * j] = __uint_as_float(v[0]);
                            dst[4 * j + 1] = __uint_as_float(v[1]);
                        } else {
                            dst[4 * j] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(w1r, off, 0, 0));
                        }
                    }
                    e = 0;
                } else {
                    e = esc[ts];
                }
            };
            // score mode: the exponents of the four-step block starting at tb
            // (entries past the pair's last step are unused; the guard ints after
            // the strip keep the load in bounds). An eight-step iteration keeps
            // steps 0..3 in e4A and 4..7 in e4B, each loaded four steps before its
            // first use and after its previous block's last, so that no register
            // copy at the loop's end has to wait for a load.
            auto loadE4 = [&](int tb) { return *(const Int4u *)(esc + (tb < Tw ? tb : 0)); };
            Int4u e4A = loadE4(0), e4B{0, 0, 0, 0};
#pragma unroll
            for (int f = 0; f < kPF; f++) loadW1(f, Wq[f], Eq[f]);
            ringPut(0, frecAt(k & (kRingBlock - 1), 0));  // steps 0..7; the loop writes the later blocks
            Rec fnext = ringGet(1);  // the first lane's record of the next step
            Rec fcur = frecAt(0, k);

            // strip kernels: the strip above's last row (carried sum, Y, exponent), also ahead
            Bnd fq4[kPF];
            // (loaded in every strip, used only when there is a strip above)
            // (used by the first lane; one address per group outside dump mode)
            auto fghostAt = [&](int tt) -> Bnd {
                if (STR) return bF[bidx(tt - (MODE == kDump ? k : 0), n)];
                return Bnd{0.f, 0.f, 0};
            };
            Bnd fgnext{0.f, 0.f, 0};
            if constexpr (kGhostRing) {  // score mode: through the ghost ring
                gringPut(0, fghostAt(k & (kRingBlock - 1)));
                fgnext = gringGet(0);  // step 0
            } else {
#pragma unroll
                for (int f = 0; f < kPF; f++) fq4[f] = fghostAt(f);
            }
            float eqf[R];  // emission values of the next step, one step ahead
#pragma unroll
            for (int q = 0; q < R; q++) eqf[q] = etk[((fcur.codes >> 8) & 255) * ROWS + q];
            float gqf = bgl[(fcur.codes >> 8) & 255];  // g[c-2] of the next step

            auto fstep = [&](auto PHc, int t) {
                constexpr int PH = decltype(PHc)::value;
                constexpr int P4 = PH & 3;  // step within a four-step exponent block
                constexpr int S1 = (P4 + 3) & 3, S2 = (P4 + 2) & 3, S3 = (P4 + 1) & 3, S4 = P4;
                constexpr int SL = PH % kPF;
                const int D = t;
                const int c = D - k;
                const Rec rc = fcur;
                const float rcg = gqf;
                // the next step's record (the lane above's, or a new one at the first
                // lane) and its emission values, read now so that they arrive in time
                Rec nx = rec_from_above<G>(fcur);
                if (injF) nx = fnext;
                fnext = ringGet(((t - PH) & (2 * kRingBlock - 1)) + PH + 2);  // step t + 2
                float eqN[R];
#pragma unroll
                for (int q = 0; q < R; q++) eqN[q] = etk[((nx.codes >> 8) & 255) * ROWS + q];
                const float gqN = bgl[(nx.codes >> 8) & 255];
                // the stored block of this step scaled by 2^-100 (so that w0 W1
                // cannot overflow), or zero where the lane's column lies at or
                // past n or all its rows past M (no W1 of the matrix there)
                float W1v[R];
                const float w1f = (MODE == kScore && t - k < nW1) ? 7.888609052e-31f : 0.f;
#pragma unroll
                for (int q = 0; q < R; q++) W1v[q] = Wq[SL][q] * w1f;
                const Int4u &e4 = (PH < 4) ? e4A : e4B;  // this step's block
                const int EbNext = (t < Tw) ? ((MODE == kScore) ? e4.at<P4>() : Eq[SL]) : kNoStep;
                if (MODE == kScore && P4 == 0) {
                    // one test per four-step block: does any exponent of the block
                    // (of this pair's steps) differ from the current one
                    const bool d = ((t < Tw) & (e4.v0 != EbPrev)) | ((t + 1 < Tw) & (e4.v1 != EbPrev)) |
                                   ((t + 2 < Tw) & (e4.v2 != EbPrev)) | ((t + 3 < Tw) & (e4.v3 != EbPrev));
                    eChg = __any(d);
                    // the other register: the block four steps on
                    if (PH < 4) e4B = loadE4(t + 4);
                    else e4A = loadE4(t + 4);
                }
                loadW1(t + kPF, Wq[SL], Eq[SL]);
                float gn = 0.f, gy = 0.f;
                Bnd fcurB{0.f, 0.f, 0};
                if constexpr (kGhostRing) {
                    fcurB = fgnext;
                    fgnext = gringGet(((t - PH) & (2 * kRingBlock - 1)) + PH + 1);  // step t + 1
                } else if (STR) {
                    fcurB = fq4[SL];
                    fq4[SL] = fghostAt(t + kPF);
                }
                if (STR && above) {
                    // lane 0: the strip above's last row from the boundary buffer;
                    // raise this strip's exponent first when that row is far above it
                    // (tested as in the backward, on the first lane)
                    const float rn = fcurB.v, ry = fcurB.y;
                    const int ge = fcurB.e;
                    gn = ldexpf(rn, ge - Ef);
                    gy = ldexpf(ry, ge - Ef);
                    if (__builtin_expect(__any(k == 0 && fmaxf(gn, gy) >= kGhostTopV), 0)) {
                        int need = 0;
                        if (k == 0 && (rn > 0.f || ry > 0.f)) {
                            int ex;
                            (void)frexpf(fmaxf(rn, ry), &ex);
                            need = ex + ge - Ef - kGhostTop;
                        }
                        need = group_lane<G>(need, grp, 0);  // the group's first lane
                        if (need > 0) scaleF(-((need + 63) & ~63));
                        gn = ldexpf(rn, ge - Ef);
                        gy = ldexpf(ry, ge - Ef);
                    }
                }
                float Nin = from_above<G>(lastN);
                float Ya = from_above<G>(lastY);
                if (STR && k == 0) {  // the strip above's row, or 0 in the top strip
                    Nin = gn;
                    Ya = gy;
                }
                const bool inD = D < Tmine;
                if (MODE != kScore || __builtin_expect(eChg, 0)) {
                    const int EbD = (EbNext != kNoStep) ? EbNext : EbPrev;
                    if (MODE == kScore && EbD != EbPrev) midMax.rebase(Ef + EbD + kPre);
                    EbPrev = EbD;
                }
                const float gsC = rcg * sC;
                const float psr = rc.ps * rf;
                float smax = 0.f;
#pragma unroll
                for (int q = 0; q < R; q++) {
                    const float w1 = Wf[S1][q], w2 = Wf[S2][q], w3 = Wf[S3][q];
                    const float e = eqf[q];
                    float X, Z0, w0, y, nn;
                    if constexpr (MODE == kDump) {
                        // X = enter e w3; Z0 = g sC (a0 w3 + b0 Z0'); w0 = W0in + (Z0 + a1 w1 + a2 w2 + seed);
                        // Y = d0 w0 + eps Y'; next row's W0in = X + Y + d1 w2 + d2 w1, left to right
                        X = mul_nc(mul_nc(enterC[q], e), w3);
                        Z0 = mul_nc(gsC, add_nc(mul_nc(a0[q], w3), mul_nc(b0[q], Zf[S3][q])));
                        w0 = add_nc(add_nc(add_nc(Z0, mul_nc(a1s1[q], w1)), mul_nc(a2s2[q], w2)), psr);
                        w0 = add_nc(Nin, w0);
                        y = add_nc(mul_nc(d0[q], w0), mul_nc(eps[q], Ya));
                        nn = add_nc(add_nc(add_nc(X, y), mul_nc(d1s2[q], w2)), mul_nc(d2s1[q], w1));
                    } else {
                        X = enterC[q] * e * w3;
                        Z0 = gsC * fmaf(b0[q], Zf[S3][q], a0[q] * w3);
                        w0 = fmaf(a1s1[q], w1, Z0);
                        w0 = fmaf(a2s2[q], w2, w0);
                        w0 = Nin + (w0 + psr);
                        y = fmaf(d0[q], w0, eps[q] * Ya);
                        nn = X + y;
                        nn = fmaf(d1s2[q], w2, nn);
                        nn = fmaf(d2s1[q], w1, nn);
                    }
                    if (MODE == kScore) midMax.cur = nnmax(midMax.cur, w0 * W1v[q]);
                    if (MODE == kDump && inD && active) w0d[(int64_t)D * ROWS + k * R + q] = w0;
                    smax = nnmax(smax, nnmax(w0, y));
                    Wf[S4][q] = w0;
                    Zf[S4][q] = Z0;
                    Nin = nn;
                    Ya = y;
                }
                lastN = Nin;
                lastY = Ya;
                fcur = nx;
#pragma unroll
                for (int q = 0; q < R; q++) eqf[q] = eqN[q];
                gqf = gqN;
                if (STR && below && k == G - 1 && c >= 0 && c < n) {
                    bF[c + 1] = Bnd{lastN, lastY, Ef};
                }
                if (MODE == kDump && inD && active && k == 0) efd[D] = Ef;
                // exponent moves as in the backward (up at 2^100, down below 2^36)
                LVf[S4] = nnmax(nnmax(smax, nnmax(lastN, lastY)), nnmax(gn, gy));
                const bool big = smax > kBig;
                if (__builtin_expect(__any(big), 0)) {
                    bool gbig = big;
#pragma unroll
                    for (int s = 1; s < G; s <<= 1) gbig |= __shfl_xor((int)gbig, s, 64) != 0;
                    if (gbig) scaleF(-64);
                }
                if (P4 == 3) {  // never below exponent 0: the seeds enter at 2^-Ef <= 1
                    const float live = nnmax(nnmax(LVf[0], LVf[1]), nnmax(LVf[2], LVf[3]));
                    if (__builtin_expect(group_small<G>(live, grp), 0) && Ef >= 64) scaleF(64);
                }
            };
            int t = 0;
            for (; t + 8 <= T; t += 8) {
                // the ring block that starts at step t + 8: loaded now, written
                // after step t + 5 (step t + 6 reads its first entry)
                const Rec fnew = frecAt(t + 8 + (k & (kRingBlock - 1)), 0);
                Bnd fgnew{0.f, 0.f, 0};
                if constexpr (kGhostRing) fgnew = fghostAt(t + 8 + (k & (kRingBlock - 1)));
                fstep(std::integral_constant<int, 0>{}, t);
                fstep(std::integral_constant<int, 1>{}, t + 1);
                fstep(std::integral_constant<int, 2>{}, t + 2);
                fstep(std::integral_constant<int, 3>{}, t + 3);
                fstep(std::integral_constant<int, 4>{}, t + 4);
                fstep(std::integral_constant<int, 5>{}, t + 5);
                ringPut((t + 8) & (2 * kRingBlock - 1), fnew);
                if constexpr (kGhostRing) gringPut((t + 8) & (2 * kRingBlock - 1), fgnew);
                fstep(std::integral_constant<int, 6>{}, t + 6);
                fstep(std::integral_constant<int, 7>{}, t + 7);
            }
            if (t < T) fstep(std::integral_constant<int, 0>{}, t++);
            if (t < T) fstep(std::integral_constant<int, 1>{}, t++);
            if (t < T) fstep(std::integral_constant<int, 2>{}, t++);
            if (t < T) fstep(std::integral_constant<int, 3>{}, t++);
            if (t < T) fstep(std::integral_constant<int, 4>{}, t++);
            if (t < T) fstep(std::integral_constant<int, 5>{}, t++);
            if (t < T) fstep(std::integral_constant<int, 6>{}, t++);
            if (STR) __syncthreads();  // boundary written above is read by the next strip
        }
        if (MODE == kScore) {
            midMax.retire();
            const ES mid = group_max<G>(ES{midMax.sm, midMax.se});
            if (active && k == 0) a.scoreOut[PR.outIdx] = PairScore{mid.m, mid.e, bwdBest.m, bwdBest.e};
        }
    }
}

}  // namespace dm

// ---------------------------------------------------------------------------
// Launch table: (G, R) configurations compiled for gfx950.
// ---------------------------------------------------------------------------
namespace dm {

#define DM_CFGS(X)                                                                                 \
    X(16, 2) X(16, 3) X(16, 4) X(16, 5) X(16, 6) X(16, 7) X(16, 8)                               \
    X(32, 2) X(32, 3) X(32, 4) X(32, 5) X(32, 6) X(32, 7) X(32, 8)                               \
    X(64, 1) X(64, 2) X(64, 3) X(64, 4) X(64, 5) X(64, 6) X(64, 7) X(64, 8) X(64, 12) X(64, 16)

// Row-strip kernels: any profile length, rows in strips of G*R.
#define DM_STRIP_CFGS(X) X(32, 4) X(32, 6) X(32, 8) X(64, 4) X(64, 6) X(64, 8) X(64, 16)

#define DM_LIST(g, r)        \
    if (n < cap) {           \
        gs[n] = g;           \
        rs[n] = r;           \
    }                        \
    n++;
int dp_configs(int *gs, int *rs, int cap) {
    int n = 0;
    DM_CFGS(DM_LIST)
    return n;
}
int dp_strip_configs(int *gs, int *rs, int cap) {
    int n = 0;
    DM_STRIP_CFGS(DM_LIST)
    return n;
}
#undef DM_LIST

hipError_t launch_null(const NullArgs &a, hipStream_t s) {
    const int blocks = (a.npairs + 63) / 64;
    if (blocks == 0) return hipSuccess;
    hipLaunchKernelGGL(null_kernel, dim3(blocks), dim3(64), 0, s, a);
    return hipGetLastError();
}

// Tables above 64 KiB of LDS need the per-kernel dynamic-LDS limit raised.
// grid = workgroups; wg = waves per workgroup (dp_wg).
template <class K>
static hipError_t launch_one(K kernel, size_t smem, int grid, int wg, hipStream_t s, const DPArgs &a) {
    if (smem > 65536) {
        hipError_t r = hipFuncSetAttribute((const void *)kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem);
        if (r != hipSuccess) return r;
    }
    hipLaunchKernelGGL(kernel, dim3(grid), dim3(64 * wg), smem, s, a);
    return hipGetLastError();
}

hipError_t launch_dp(int G, int R, int mode, bool strips, const DPArgs &a, int grid, hipStream_t s) {
    const size_t smem = dp_smem(G, R, strips, mode);
    const int wg = dp_wg(G, R, strips, mode);
#define DM_LAUNCH(g, r)                                                                                 \
    if (G == g && R == r && !strips) {                                                                  \
        if (mode == kFilter) return launch_one(dp_kernel<g, r, kFilter, false>, smem, grid, wg, s, a);  \
        if (mode == kScore) return launch_one(dp_kernel<g, r, kScore, false>, smem, grid, wg, s, a);    \
        return launch_one(dp_kernel<g, r, kDump, false>, smem, grid, wg, s, a);                         \
    }
#define DM_LAUNCH_STRIPS(g, r)                                                                          \
    if (G == g && R == r && strips) {                                                                   \
        if (mode == kFilter) return launch_one(dp_kernel<g, r, kFilter, true>, smem, grid, wg, s, a);   \
        if (mode == kScore) return launch_one(dp_kernel<g, r, kScore, true>, smem, grid, wg, s, a);     \
        return launch_one(dp_kernel<g, r, kDump, true>, smem, grid, wg, s, a);                          \
    }
    DM_CFGS(DM_LAUNCH)
    DM_STRIP_CFGS(DM_LAUNCH_STRIPS)
#undef DM_LAUNCH
#undef DM_LAUNCH_STRIPS
    return hipErrorInvalidValue;
}

}  // namespace dm

namespace dm {

// Clears the work-unit counters (a kernel, not a memset node, so that plans
// captured into a hipGraph replay correctly).
__global__ void zero_ints_kernel(int *p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

// dst[idx[k]] = src[k]: results staged in work-queue order back to the
// caller's pair order.
__global__ void scatter_floats_kernel(const float *src, const int *idx, float *dst, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[idx[i]] = src[i];
}

// A plain copy of n16 16-byte words: the reference point for the HBM traffic
// counters (its algorithmic traffic is exactly 16 bytes read + 16 written per word).
__global__ void __launch_bounds__(256) copy_probe_kernel(const float4 *src, float4 *dst, int64_t n16) {
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n16; i += (int64_t)gridDim.x * 256)
        dst[i] = src[i];
}

// ---------------------------------------------------------------------------
// tantan on DNA (sequence masking, -m 2): makeMaskedSequence
// (dummest-sequence.hh:206-236) -> calcTantanProbabilities(isProtein = false)
// (tantan-wrapper.hh:13-64) -> the gap-free forward-backward of tantan.cc
// (40-151, 222-397): per letter, 100 repeat-offset states and one background
// state in FP64, rescaled every 16 letters.
//
// Four lanes per sequence: lane l of a group holds the offsets i = 4k + l
// (k < 25), which is exactly the partial sum the reference's AVX2 lane l
// accumulates (mcf_simd.hh:74). The partial sums meet as (s0+s2)+(s1+s3)
// through two lane exchanges (mcf_simd.hh:175-179) and the offsets past the
// last full quad (the first 100 letters) are added afterwards in order, as the
// reference's scalar tail does; multiply-adds are fused where the reference's
// x86-64-v3 build fuses them. The last 100 letters sit in 2-bit fields of
// seven registers (field i = the letter i+1 back), so the likelihood ratio of
// offset i is one field extract. Loops are wave-uniform (sequences of a wave
// are length-sorted by the host); a group's state is left unchanged at
// positions past its own length.
// ---------------------------------------------------------------------------
constexpr int kTQ = kTantanK / 4;  // offsets per lane

__device__ __forceinline__ int tantan_base(const uint8_t *bs, int n, int p) {
    return (n > 0) ? (int)bs[p < 0 ? 0 : (p >= n ? n - 1 : p)] : 0;
}

__global__ void __launch_bounds__(64) tantan_dna_kernel(TantanArgs a) {
#pragma clang fp contract(off)
    const int lane = threadIdx.x;
    const int l = lane & 3;
    const int g0 = lane & ~3;  // first lane of the group
    const int sq = blockIdx.x * 16 + (lane >> 2);
    const bool valid = sq < a.nseq;
    const int n = valid ? a.lengths[sq] : 0;
    const int64_t off = valid ? a.offsets[sq] : 0;
    const uint8_t *bs = a.bases + off;
    float *out = a.probs + off;
    double *sc = a.scales + (valid ? a.scaleOff[sq] : 0);
    int N = n;
#pragma unroll
    for (int s = 1; s < 64; s <<= 1) N = max(N, __shfl_xor(N, s));

    // the offset weights are the same for every sequence: LDS, read at use
    __shared__ double sb2f[kTantanK];
    for (int i = lane; i < kTantanK; i += 64) sb2f[i] = a.b2f[i];
    __syncthreads();
    auto b2f = [&](int k) { return sb2f[4 * k + l]; };
    double fp[kTQ];
#pragma unroll
    for (int k = 0; k < kTQ; ++k) fp[k] = 0.0;
    uint32_t h[7] = {0, 0, 0, 0, 0, 0, 0};  // field i (bits 2i, 2i+1): letter pos-1-i
    auto field = [&](int k) -> int { return (int)((h[k >> 2] >> (8 * (k & 3) + 2 * l)) & 3u); };
    const double f2f0 = a.f2f0, f2b = a.f2b, b2b = a.b2b;

    // forward (tantan.cc:283-330 with rescaleForward 364-371)
    double bg = 1.0;
    int curNext = tantan_base(bs, n, 0);
    for (int pos = 0; pos < N; ++pos) {
        const bool act = pos < n;
        const int cur = curNext;
        curNext = tantan_base(bs, n, pos + 1);
        const int m = pos < kTantanK ? pos : kTantanK;
        const int q4 = m >> 2, t = m & 3;
        const double b = bg;
        double s = 0.0, tailf = 0.0;
#pragma unroll
        for (int k = 0; k < kTQ; ++k) {
            const double r = (field(k) == cur) ? 3.0 : (1.0 / 3.0);
            const double f = fp[k];
            if (k < q4) s += f;
            if (k == q4 && l < t) tailf = f;
            const double nf = __fma_rn(b, b2f(k), f * f2f0) * r;
            fp[k] = (act && 4 * k + l < m) ? nf : f;
        }
        double x = s + __shfl_xor(s, 2);
        x = x + __shfl_xor(x, 1);
        for (int u = 0; u < t; ++u) x += __shfl(tailf, g0 + u);
        double nb = __fma_rn(b, b2b, x * f2b);
        if ((pos & 15) == 15) {
            const double scl = 1.0 / nb;
            if (act && l == 0) sc[pos >> 4] = scl;
            nb *= scl;
#pragma unroll
            for (int k = 0; k < kTQ; ++k) fp[k] = act ? fp[k] * scl : fp[k];
        }
        bg = act ? nb : bg;
        if (act && l == 0) out[pos] = (float)bg;
#pragma unroll
        for (int w = 6; w > 0; --w) h[w] = (h[w] << 2) | (h[w - 1] >> 30);
        h[0] = (h[0] << 2) | (uint32_t)cur;
    }
    // forwardTotal (tantan.cc:135-141): the offsets summed in order
    double all = 0.0;
#pragma unroll
    for (int i = 0; i < kTantanK; ++i) all += __shfl(fp[i >> 2], g0 + (i & 3));
    const double z = __fma_rn(all, f2b, bg * b2b);

    // backward (tantan.cc:147-151, 332-361, 384-397)
    bg = b2b;
#pragma unroll
    for (int k = 0; k < kTQ; ++k) fp[k] = f2b;
    auto shiftBack = [&](int p) {  // window of position p-1 from that of p
#pragma unroll
        for (int w = 0; w < 6; ++w) h[w] = (h[w] >> 2) | (h[w + 1] << 30);
        h[6] = ((h[6] >> 2) & ~(3u << 6)) | ((uint32_t)tantan_base(bs, n, p - 101) << 6);
    };
    shiftBack(N);
    for (int pos = N - 1; pos >= 0; --pos) {
        const bool act = pos < n;
        const int cur = tantan_base(bs, n, pos);
        if (act && l == 0) {
            const double nonRepeat = (double)out[pos] * bg / z;
            out[pos] = 1.0f - (float)nonRepeat;
        }
        if ((pos & 15) == 15) {
            double scl = (act && l == 0) ? sc[pos >> 4] : 1.0;
            scl = __shfl(scl, g0);
            if (act) {
                bg *= scl;
#pragma unroll
                for (int k = 0; k < kTQ; ++k) fp[k] *= scl;
            }
        }
        const double toBg = f2b * bg;
        const int m = pos < kTantanK ? pos : kTantanK;
        const int q4 = m >> 2, t = m & 3;
        double s = 0.0, tb = 0.0, tf = 0.0;
#pragma unroll
        for (int k = 0; k < kTQ; ++k) {
            const double r = (field(k) == cur) ? 3.0 : (1.0 / 3.0);
            const double f = fp[k] * r;
            if (k < q4) s = __fma_rn(b2f(k), f, s);
            if (k == q4 && l < t) {
                tb = b2f(k);
                tf = f;
            }
            const double nf = __fma_rn(f, f2f0, toBg);
            fp[k] = (act && 4 * k + l < m) ? nf : fp[k];
        }
        double x = s + __shfl_xor(s, 2);
        x = x + __shfl_xor(x, 1);
        for (int u = 0; u < t; ++u) x = __fma_rn(__shfl(tb, g0 + u), __shfl(tf, g0 + u), x);
        if (act) bg = __fma_rn(b2b, bg, x);
        shiftBack(pos);
    }
}

hipError_t launch_tantan_dna(const TantanArgs &a, hipStream_t s) {
    const int blocks = (a.nseq + 15) / 16;
    if (blocks == 0) return hipSuccess;
    hipLaunchKernelGGL(tantan_dna_kernel, dim3(blocks), dim3(64), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_copy_probe(const float *src, float *dst, int64_t n16, hipStream_t s) {
    if (n16 <= 0) return hipSuccess;
    const int blocks = (int)std::min<int64_t>((n16 + 255) / 256, 8192);
    hipLaunchKernelGGL(copy_probe_kernel, dim3(blocks), dim3(256), 0, s, (const float4 *)src, (float4 *)dst, n16);
    return hipGetLastError();
}

hipError_t launch_zero_ints(int *p, int n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(zero_ints_kernel, dim3((n + 63) / 64), dim3(64), 0, s, p, n);
    return hipGetLastError();
}

hipError_t launch_scatter_floats(const float *src, const int *idx, float *dst, int n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(scatter_floats_kernel, dim3((n + 255) / 256), dim3(256), 0, s, src, idx, dst, n);
    return hipGetLastError();
}

}  // namespace dm
