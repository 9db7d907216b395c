// CPU BASELINE -- measurement infrastructure, never part of the product.
//
// The reference's lane-batched CPU forward+backward score path, restated with
// one DNA sequence per SIMD lane: 16 lanes of AVX-512 or 8 of AVX2 (the
// compiler's -march picks), the pairs of each profile sorted by sequence
// length and cut into lane batches (dummest-search.hh:76-285), the batches
// spread over a pool of threads that run with FTZ + DAZ
// (dummest-core.hh:143-146, 156-202). Per batch:
//   the null model of every lane (buildNull of oracle.cpp,
//     dummest-dp-backward.hh:6-188);
//   the backward over rows M..0 with per-lane row rescaling by 2^-64
//     (dummest-dp-backward.hh:444-615), the full W1 matrix kept;
//   the best backward start max_{i,j} prefixSeed(j) W1(i,j)
//     (dummest-dp-backward.hh:579-609);
//   the forward over rows 0..M with the mid score max_{i,j} w0(i,j) W1(i,j)
//     (dummest-dp-forward.hh:29-250).
// Per-cell ExpScore products are formed exactly: a product of two floats is
// exact in double, and per row every lane has one rescale exponent, so the
// row's largest product, normalized once, is the ExpScore the reference's
// per-cell promote-multiply-compare keeps (rounding to float is monotone).
//
// Checked against the oracle (or_fwdbwd_batch_f32, its scalar restatement of
// the same path) by tests/test_simd_baseline.py: the same scores to float
// rounding. Built with the reference's release flags (-O3, FMA contraction
// left to the compiler as in the reference's own build).
#include "oracle.cpp"

#include <immintrin.h>
#include <omp.h>

#include <chrono>
#include <memory>

namespace sb {

#if defined(__AVX512F__)
constexpr int L = 16;
typedef __m512 vf;
struct vd {
    __m512d lo, hi;
};
static inline vf set1(float x) { return _mm512_set1_ps(x); }
static inline vf ld(const float *p) { return _mm512_load_ps(p); }
static inline void st(float *p, vf x) { _mm512_store_ps(p, x); }
static inline vf vmax(vf a, vf b) { return _mm512_max_ps(a, b); }
// a > t ? y : n, lane by lane
static inline vf gtsel(vf a, vf t, vf y, vf n) { return _mm512_mask_blend_ps(_mm512_cmp_ps_mask(a, t, _CMP_GT_OQ), n, y); }
struct Table {  // 32 floats in two registers
    __m512 lo, hi;
    explicit Table(const float *t) : lo(_mm512_loadu_ps(t)), hi(_mm512_loadu_ps(t + 16)) {}
    vf operator()(const int32_t *idx) const { return _mm512_permutex2var_ps(lo, _mm512_load_si512(idx), hi); }
};
static inline vd dzero() { return vd{_mm512_setzero_pd(), _mm512_setzero_pd()}; }
static inline vd dmaxprod(vd m, vf a, vf b) {  // max(m, a b) with the products in double (exact)
    const __m512d alo = _mm512_cvtps_pd(_mm512_castps512_ps256(a)), blo = _mm512_cvtps_pd(_mm512_castps512_ps256(b));
    const __m512d ahi = _mm512_cvtps_pd(_mm512_extractf32x8_ps(a, 1)), bhi = _mm512_cvtps_pd(_mm512_extractf32x8_ps(b, 1));
    return vd{_mm512_max_pd(m.lo, _mm512_mul_pd(alo, blo)), _mm512_max_pd(m.hi, _mm512_mul_pd(ahi, bhi))};
}
static inline void dst(double *p, vd x) {
    _mm512_storeu_pd(p, x.lo);
    _mm512_storeu_pd(p + 8, x.hi);
}
#else
constexpr int L = 8;
typedef __m256 vf;
struct vd {
    __m256d lo, hi;
};
static inline vf set1(float x) { return _mm256_set1_ps(x); }
static inline vf ld(const float *p) { return _mm256_load_ps(p); }
static inline void st(float *p, vf x) { _mm256_store_ps(p, x); }
static inline vf vmax(vf a, vf b) { return _mm256_max_ps(a, b); }
static inline vf gtsel(vf a, vf t, vf y, vf n) { return _mm256_blendv_ps(n, y, _mm256_cmp_ps(a, t, _CMP_GT_OQ)); }
struct Table {  // gathered from memory
    const float *t;
    explicit Table(const float *p) : t(p) {}
    vf operator()(const int32_t *idx) const {
        return _mm256_i32gather_ps(t, _mm256_load_si256((const __m256i *)idx), 4);
    }
};
static inline vd dzero() { return vd{_mm256_setzero_pd(), _mm256_setzero_pd()}; }
static inline vd dmaxprod(vd m, vf a, vf b) {
    const __m256d alo = _mm256_cvtps_pd(_mm256_castps256_ps128(a)), blo = _mm256_cvtps_pd(_mm256_castps256_ps128(b));
    const __m256d ahi = _mm256_cvtps_pd(_mm256_extractf128_ps(a, 1)), bhi = _mm256_cvtps_pd(_mm256_extractf128_ps(b, 1));
    return vd{_mm256_max_pd(m.lo, _mm256_mul_pd(alo, blo)), _mm256_max_pd(m.hi, _mm256_mul_pd(ahi, bhi))};
}
static inline void dst(double *p, vd x) {
    _mm256_storeu_pd(p, x.lo);
    _mm256_storeu_pd(p + 4, x.hi);
}
#endif

template <class T>
struct Buf {  // 64-byte aligned, grown on demand
    std::unique_ptr<T[], void (*)(T *)> p{nullptr, [](T *q) { std::free(q); }};
    size_t n = 0;
    T *get(size_t need) {
        if (need > n) {
            const size_t bytes = ((need * sizeof(T) + 63) / 64) * 64;
            p.reset((T *)std::aligned_alloc(64, bytes));
            n = need;
        }
        return p.get();
    }
};

struct Work {
    Buf<float> W1, Yn, Yc, W0c, W0n, pre, suf, bgv, act;
    Buf<int32_t> idx;
    Buf<int> bcnt;
};

// One profile's data: emission rows padded to 32 floats (codes 0..24 used).
struct ProfData {
    int M;
    const float *params, *bg;
    std::vector<float> emis;  // (M+1) x 32
};

static inline void fold(orc::ES &best, double v, int e2) {
    if (!(v > 0.0)) return;
    int ex = 0;
    const double m = std::frexp(v, &ex);
    const orc::ES c{(float)m, ex + e2};
    if (orc::scoreLess(best, c)) best = c;
}

// One lane batch: sequences seqs[0..cnt) (cnt <= L), the longest first; the
// best backward start only when kBwd.
template <bool kBwd>
static void runBatch(Work &Wk, const ProfData &P, const orc::Rates<float> &R, const uint8_t *codes,
                     const int64_t *offsets, const int *lens, const int *seqs, int cnt, orc::ES *midOut,
                     orc::ES *bwdOut) {
    const int M = P.M;
    int Wd = 0;
    for (int k = 0; k < cnt; k++) Wd = std::max(Wd, lens[seqs[k]]);
    const int WC = Wd + 4;  // columns kept per row (W..W+3 stay zero)
    float *pre = Wk.pre.get((size_t)WC * L), *suf = Wk.suf.get((size_t)WC * L);
    float *bgv = Wk.bgv.get((size_t)(Wd + 8) * L), *act = Wk.act.get((size_t)WC * L);
    int32_t *idx = Wk.idx.get((size_t)(Wd + 8) * L);  // column j at (j + 4) L
    alignas(64) float s1[L], s2[L], sC[L];
    for (int k = 0; k < L; k++) {
        const bool live = k < cnt;
        const int n = live ? lens[seqs[k]] : 0;
        const uint8_t *c = live ? codes + offsets[seqs[k]] : nullptr;
        for (int j = -4; j < Wd + 4; j++) {
            const int code = (j >= 0 && j < n) ? c[j] : orc::kZeroIdx;
            idx[(size_t)(j + 4) * L + k] = code;
            bgv[(size_t)(j + 4) * L + k] = P.bg[4 + code];
        }
        for (int j = 0; j < WC; j++) act[(size_t)j * L + k] = j < n ? 1.0f : 0.0f;
        if (live) {
            orc::Null<float> N;
            orc::buildNull<float>(N, P.bg + 4, c, n, n, R);
            for (int j = 0; j < WC; j++) {
                pre[(size_t)j * L + k] = j < n ? N.prefixSeed[j] : 0.0f;
                suf[(size_t)j * L + k] = j < n ? N.suffixSeed[j] : 0.0f;
            }
            s1[k] = N.s1f;
            s2[k] = N.s2f;
            sC[k] = N.sCf;
        } else {
            for (int j = 0; j < WC; j++) pre[(size_t)j * L + k] = suf[(size_t)j * L + k] = 0.0f;
            s1[k] = s2[k] = sC[k] = 0.0f;
        }
    }
    const vf vs1 = ld(s1), vs2 = ld(s2), vsC = ld(sC);
    const float thresh = orc::rescaleThresh<float>();
    const float step = std::ldexp(1.0f, -orc::RESCALE_BITS);
    const vf vthr = set1(thresh), vstep = set1(step), one = set1(1.0f), zero = set1(0.0f);
    const size_t RS = (size_t)WC * L;  // floats per W1 row
    float *W1 = Wk.W1.get((size_t)(M + 2) * RS);
    float *Yn = Wk.Yn.get(RS), *Yc = Wk.Yc.get(RS);
    int *bcnt = Wk.bcnt.get((size_t)(M + 2) * 2 * L), *fcnt = bcnt + (size_t)(M + 2) * L;
    std::memset(W1 + (size_t)(M + 1) * RS, 0, RS * sizeof(float));
    std::memset(Yn, 0, RS * sizeof(float));
    std::memset(Yc, 0, RS * sizeof(float));
    orc::ES bbest[L], mbest[L];
    alignas(64) float rm[L];
    alignas(64) double dv[L];
    auto scaleOf = [&](const int *cnt) {
        alignas(64) float s[L];
        for (int k = 0; k < L; k++) s[k] = std::ldexp(1.0f, -orc::RESCALE_BITS * orc::clampRescaleCount(cnt[k]));
        return ld(s);
    };
    // ---- backward, dummest-dp-backward.hh:473-575
    for (int i = M; i >= 0; i--) {
        int *bc = bcnt + (size_t)i * L;
        for (int k = 0; k < L; k++) bc[k] = i < M ? bc[k + L] : 0;
        const float *pr = P.params + 11 * (size_t)i;
        const Table em(P.emis.data() + 32 * (size_t)i);
        const vf enterC = set1(pr[10]) * vsC, d0 = set1(pr[6]), d1 = set1(pr[7]) * vs2, d2 = set1(pr[8]) * vs1;
        const vf a0 = set1(pr[0]) * vsC, a1 = set1(pr[1]) * vs1, a2 = set1(pr[2]) * vs2, b0 = set1(pr[3]) * vsC;
        const vf eps = set1(pr[9]), rsc = scaleOf(bc);
        float *row = W1 + (size_t)i * RS;
        const float *below = row + RS;
        for (int j = Wd; j < WC; j++) st(row + (size_t)j * L, zero);
        vf n3 = ld(below + (size_t)(Wd + 2) * L), n2 = ld(below + (size_t)(Wd + 1) * L), n1 = ld(below + (size_t)Wd * L);
        vf c2 = zero, c1 = zero, z1 = zero, z2 = zero, z3 = zero, rowMax = zero;
        vd bb = dzero();
        for (int j = Wd - 1; j >= 0; j--) {
            const size_t o = (size_t)j * L;
            const vf e = em(idx + o + 5 * L), g = ld(bgv + o + 5 * L), a = ld(act + o);
            const vf nAt = ld(below + o), yn = ld(Yn + o);
            vf w = n3 * e * enterC + yn * d0 + n2 * d1 + n1 * d2 + z3 * g * a0 + c2 * a2 + c1 * a1 + ld(suf + o) * rsc;
            w = w * a;
            st(row + o, w);
            const vf y = (eps * yn + w) * a;
            st(Yc + o, y);
            rowMax = vmax(rowMax, vmax(w, y));
            const vf z0 = b0 * (z3 * g) + w;
            if constexpr (kBwd) bb = dmaxprod(bb, w, ld(pre + o));
            z3 = z2;
            z2 = z1;
            z1 = z0;
            n3 = n2;
            n2 = n1;
            n1 = nAt;
            c2 = c1;
            c1 = w;
        }
        st(rm, rowMax);
        const vf m = gtsel(rowMax, vthr, vstep, one);
        bool any = false;
        for (int k = 0; k < L; k++) {
            if (rm[k] > thresh) {
                bc[k]++;
                any = true;
            }
        }
        if (any) {
            for (int j = 0; j < WC; j++) {
                const size_t o = (size_t)j * L;
                st(row + o, ld(row + o) * m);
                st(Yn + o, ld(Yn + o) * m);
                st(Yc + o, ld(Yc + o) * m);
            }
        }
        // bestOut, dummest-dp-backward.hh:579-609: the row's largest
        // prefixSeed(j) W1(i,j) at the row's final scale
        if constexpr (kBwd) {
            dst(dv, bb);
            for (int k = 0; k < cnt; k++)
                fold(bbest[k], rm[k] > thresh ? dv[k] * (double)step : dv[k],
                     orc::RESCALE_BITS * orc::clampRescaleCount(bc[k]));
        }
        std::swap(Yc, Yn);
    }
    // ---- forward, dummest-dp-forward.hh:29-250
    float *W0c = Wk.W0c.get(RS), *W0n = Wk.W0n.get(RS);
    std::memset(W0c, 0, RS * sizeof(float));
    std::memset(Yn, 0, RS * sizeof(float));
    for (int i = 0; i <= M; i++) {
        int *fc = fcnt + (size_t)i * L;
        for (int k = 0; k < L; k++) fc[k] = i > 0 ? fc[k - L] : 0;
        const int *bc = bcnt + (size_t)i * L;
        const float *pr = P.params + 11 * (size_t)i;
        const Table em(P.emis.data() + 32 * (size_t)i);
        const vf enterC = set1(pr[10]) * vsC, a0 = set1(pr[0]), a1 = set1(pr[1]) * vs1, a2 = set1(pr[2]) * vs2;
        const vf b0 = set1(pr[3]), d0 = set1(pr[6]), d1 = set1(pr[7]) * vs2, d2 = set1(pr[8]) * vs1;
        const vf eps = set1(pr[9]), rsc = scaleOf(fc);
        const bool hasNext = i + 1 <= M;
        const float *row = W1 + (size_t)i * RS;
        vf w1v = zero, w2v = zero, w3v = zero, z1 = zero, z2 = zero, z3 = zero, rowMax = zero;
        vd mm = dzero();
        for (int j = 0; j < Wd; j++) {
            const size_t o = (size_t)j * L;
            const vf e = em(idx + o + 2 * L), g = ld(bgv + o + 2 * L), a = ld(act + o);
            const vf X = enterC * e * w3v;
            const vf z0 = g * vsC * (a0 * w3v + b0 * z3);
            vf w0 = ld(W0c + o) + (z0 + a1 * w1v + a2 * w2v + ld(pre + o) * rsc);
            w0 = w0 * a;
            rowMax = vmax(rowMax, w0);
            mm = dmaxprod(mm, w0, ld(row + o));
            const vf y = (d0 * w0 + eps * ld(Yn + o)) * a;
            st(Yc + o, y);
            rowMax = vmax(rowMax, y);
            if (hasNext) st(W0n + o, X + y + d1 * w2v + d2 * w1v);
            z3 = z2;
            z2 = z1;
            z1 = z0 * a;
            w3v = w2v;
            w2v = w1v;
            w1v = w0;
        }
        // the mid scores at this row's scales (forward before its rescale)
        dst(dv, mm);
        for (int k = 0; k < cnt; k++)
            fold(mbest[k], dv[k],
                 orc::RESCALE_BITS * (orc::clampRescaleCount(fc[k]) + orc::clampRescaleCount(bc[k])));
        st(rm, rowMax);
        const vf m = gtsel(rowMax, vthr, vstep, one);
        bool any = false;
        for (int k = 0; k < L; k++) {
            if (rm[k] > thresh) {
                fc[k]++;
                any = true;
            }
        }
        if (any) {
            for (int j = 0; j < Wd; j++) {
                const size_t o = (size_t)j * L;
                if (hasNext) st(W0n + o, ld(W0n + o) * m);
                st(Yc + o, ld(Yc + o) * m);
            }
        }
        std::swap(Yc, Yn);
        std::swap(W0c, W0n);
    }
    for (int k = 0; k < cnt; k++) {
        midOut[k] = mbest[k];
        bwdOut[k] = bbest[k];
    }
}

}  // namespace sb

extern "C" {

int sb_lanes() { return sb::L; }

// Forward+backward scores of npairs (profile, sequence) pairs. Profiles back
// to back: profile p has Ms[p] positions, values ((M+1) x 29) at valOff[p],
// params ((M+2) x 11) at parOff[p], bg (29) at 29 p. Sequences: codes at
// offsets[s], lens[s] letters. Per pair (caller's order) the best mid score and
// the best backward start (want_bwd != 0; else zero) as (mantissa, exponent)
// in mid / bwd. Returns the
// wall time of the batched work in seconds (lane batching, null models and
// DP; the threads' pool started beforehand).
double sb_fwdbwd_pairs(int nprof, const int *Ms, const int64_t *valOff, const int64_t *parOff, const float *values,
                       const float *params, const float *bgs, const double *opts, const uint8_t *codes,
                       const int64_t *offsets, const int *lens, int npairs, const int *pairProf, const int *pairSeq,
                       float *mid, float *bwd, int want_bwd, int nthreads) {
    orc::Rates<float> R;
    R.init(opts);
#pragma omp parallel num_threads(nthreads)
    { orc::FtzScope ftz_; }  // start the pool before the clock
    const auto t0 = std::chrono::steady_clock::now();
    std::vector<sb::ProfData> P(nprof);
    for (int p = 0; p < nprof; p++) {
        P[p].M = Ms[p];
        P[p].params = params + parOff[p];
        P[p].bg = bgs + 29 * (size_t)p;
        P[p].emis.assign((size_t)(Ms[p] + 1) * 32, 0.0f);
        for (int i = 0; i <= Ms[p]; i++)
            for (int c = 0; c < 25; c++) P[p].emis[32 * (size_t)i + c] = values[valOff[p] + 29 * (size_t)i + 4 + c];
    }
    // pairs by profile, longest sequence first, cut into lane batches
    std::vector<int> order(npairs);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
        if (pairProf[a] != pairProf[b]) return pairProf[a] < pairProf[b];
        return lens[pairSeq[a]] > lens[pairSeq[b]];
    });
    std::vector<std::pair<int, int>> batches;  // [first, count) in order
    for (int a = 0; a < npairs;) {
        int b = a;
        while (b < npairs && b - a < sb::L && pairProf[order[b]] == pairProf[order[a]]) b++;
        batches.push_back({a, b - a});
        a = b;
    }
    // the largest batches (in cells) first
    std::vector<int64_t> cost(batches.size());
    for (size_t q = 0; q < batches.size(); q++)
        cost[q] = (int64_t)(Ms[pairProf[order[batches[q].first]]] + 1) * lens[pairSeq[order[batches[q].first]]];
    std::vector<int> bq(batches.size());
    std::iota(bq.begin(), bq.end(), 0);
    std::stable_sort(bq.begin(), bq.end(), [&](int a, int b) { return cost[a] > cost[b]; });
#pragma omp parallel num_threads(nthreads)
    {
        orc::FtzScope ftz_;
        sb::Work Wk;
        std::vector<int> seqs(sb::L);
        orc::ES mo[sb::L], bo[sb::L];
#pragma omp for schedule(dynamic, 1)
        for (size_t q = 0; q < bq.size(); q++) {
            const auto [first, cnt] = batches[bq[q]];
            const int p = pairProf[order[first]];
            for (int k = 0; k < cnt; k++) seqs[k] = pairSeq[order[first + k]];
            if (want_bwd)
                sb::runBatch<true>(Wk, P[p], R, codes, offsets, lens, seqs.data(), cnt, mo, bo);
            else
                sb::runBatch<false>(Wk, P[p], R, codes, offsets, lens, seqs.data(), cnt, mo, bo);
            for (int k = 0; k < cnt; k++) {
                const int pr = order[first + k];
                mid[2 * (size_t)pr] = mo[k].m;
                mid[2 * (size_t)pr + 1] = (float)mo[k].e;
                bwd[2 * (size_t)pr] = bo[k].m;
                bwd[2 * (size_t)pr + 1] = (float)bo[k].e;
            }
        }
    }
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

}  // extern "C"
