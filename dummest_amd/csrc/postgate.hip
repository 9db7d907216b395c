// Batched post-gate path on gfx950: for every gated pair of a batch, from the
// dump-mode DP matrices (FP32 values with a power-of-two exponent per
// anti-diagonal step) to its hits and their aligned segments.
//
// The steps are the reference's forwardPass epilogue, in its order and with
// its ExpScore arithmetic (the host restatement post.cpp computes the same
// numbers one pair at a time): anchor products w0 W1 and the per-column best
// (dummest-dp-forward.hh:174-206), the posterior match matrix
// X(i,j) = X_ij W1(i+1,j) (:126-156), the gate (:252-274), the flank totals
// (buildFlankEVs, dummest-dp.hh:137-268), the row-scaled MEA tables
// prefixOptimal / suffixOptimal (dummest-dp-forward.hh:294-452,
// dummest-scalerow.hh) and the tracebacks of the ranked anchors
// (dummest-dp-forward.hh:456-545, dummest-traceback.hh:35-95). The anchors
// are ranked on the host, where std::sort gives the reference's order of tied
// anchors.
//
// Work split: one thread per column for the per-column pass (it walks the
// rows in order, so the column totals are summed in the reference's order);
// one thread per pair for the flank recurrences; one workgroup per pair and
// table for the MEA tables, whose rows are an elementwise step and a running
// maximum along the row (a parallel prefix maximum: exact in any order); one
// thread per pair for the tracebacks, which read only the cells along their
// paths.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>

#include "postgate.h"

namespace dm {
namespace {

// ExpScore arithmetic (dummest-score.hh:62-188, dummest-batch.hh:216-235)
__device__ __forceinline__ ES es_norm(double m, int e) {
    if (m == 0.0) return ES{0.f, 0};
    int ex = 0;
    m = frexp(m, &ex);
    return ES{(float)m, e + ex};
}
__device__ __forceinline__ bool es_less(const ES &a, const ES &b) {
    if (a.m == 0.0f) return b.m != 0.0f;
    if (b.m == 0.0f) return false;
    if (a.e != b.e) return a.e < b.e;
    return a.m < b.m;
}
__device__ __forceinline__ ES es_scale(const ES &a, double f) {
    if (a.m == 0.0f || !(f > 0.0) || !isfinite(f)) return ES{0.f, 0};
    return es_norm((double)a.m * f, a.e);
}
__device__ __forceinline__ ES es_add(ES a, ES b) {
    if (a.m == 0.0f) return b;
    if (b.m == 0.0f) return a;
    if (a.e >= b.e) {
        const int d = a.e - b.e;
        if (d > 60) return a;
        return es_norm((double)a.m + ldexp((double)b.m, -d), a.e);
    }
    const int d = b.e - a.e;
    if (d > 60) return b;
    return es_norm((double)b.m + ldexp((double)a.m, -d), b.e);
}
__device__ __forceinline__ ES es_mul(const ES &a, const ES &b) {
    if (a.m == 0.0f || b.m == 0.0f) return ES{0.f, 0};
    return es_norm((double)a.m * (double)b.m, a.e + b.e);
}
__device__ __forceinline__ ES es_bmul(const ES &A, const ES &B) {
    if (A.m == 0.0f || B.m == 0.0f) return ES{0.f, 0};
    double p = (double)A.m * (double)B.m;
    int e = A.e + B.e;
    if (p < 0.5) {
        p *= 2.0;
        e -= 1;
    }
    return ES{(float)p, e};
}
// a stored FP32 value at a power-of-two offset (promoteRescaledScore)
__device__ __forceinline__ ES es_stored(float v, int off) {
    if (!(v > 0.f)) return ES{0.f, 0};
    if (!(v < INFINITY)) return ES{0x1.fffffep-1f, 1023 + off};
    int ex = 0;
    const float m = frexpf(v, &ex);
    return ES{m, ex + off};
}
// scoreToRowScale / scoreFromRowScale, dummest-scalerow.hh:26-55
__device__ __forceinline__ float es_to_row(const ES &s, int E) {
    if (s.m == 0.0f) return 0.0f;
    int d = s.e - E;
    d = d > 0 ? 0 : (d < -1000 ? -1000 : d);
    const int ne = 126 + d;
    if (ne <= 0) return 0.0f;
    return __uint_as_float((__float_as_uint(s.m) & 0x807FFFFFu) | ((unsigned)ne << 23));
}
__device__ __forceinline__ ES es_from_row(float s, int E) {
    const unsigned u = __float_as_uint(s);
    const unsigned ef = (u >> 23) & 0xFFu;
    if (!(s > 0.0f) || ef == 0) return ES{0.f, 0};
    return ES{__uint_as_float((u & 0x807FFFFFu) | (126u << 23)), (int)ef - 126 + E};
}
// order-preserving key of an ExpScore (exponent, then mantissa; zero lowest)
__device__ __forceinline__ unsigned long long es_key(const ES &a) {
    if (a.m == 0.0f) return 0ull;
    return ((unsigned long long)(unsigned)(a.e + 0x40000000) << 32) | __float_as_uint(a.m);
}
__device__ __forceinline__ ES es_unkey(unsigned long long k) {
    if (k == 0ull) return ES{0.f, 0};
    return ES{__uint_as_float((unsigned)(k & 0xffffffffu)), (int)(k >> 32) - 0x40000000};
}

struct PairView {
    const PostPair &P;
    const PostArgs &a;
    int ROWS;
    __device__ PairView(const PostPair &P_, const PostArgs &a_) : P(P_), a(a_), ROWS(P_.G * P_.R) {}
    // dump cell (i, j): strip i / ROWS, lane k = (i % ROWS) / R, step j + k
    __device__ __forceinline__ ES dump(const float *vals, const int *exps, int i, int j) const {
        const int s = i / ROWS, ii = i - s * ROWS;
        const int k = ii / P.R, q = ii - k * P.R;
        const int64_t st = (int64_t)s * P.T + j + k;
        return es_stored(vals[P.wOff + st * ROWS + k * P.R + q], exps[P.eOff + st]);
    }
    __device__ __forceinline__ ES w0(int i, int j) const { return dump(a.w0buf, a.fbuf, i, j); }
    __device__ __forceinline__ ES W1(int i, int j) const { return dump(a.wbuf, a.ebuf, i, j); }
    __device__ __forceinline__ int stride() const { return P.n + 7; }
    __device__ __forceinline__ ES *X() const { return a.mat + P.matOff; }
    __device__ __forceinline__ ES *pre() const { return a.mat + P.matOff + (int64_t)(P.M + 2) * stride(); }
    __device__ __forceinline__ ES *suf() const { return a.mat + P.matOff + 2 * (int64_t)(P.M + 2) * stride(); }
    __device__ __forceinline__ int64_t at(int i, int j) const { return (int64_t)i * stride() + j + 3; }
    __device__ __forceinline__ ES *col(int c) const { return a.col + P.colOff + (int64_t)c * (P.n + 8); }
    __device__ __forceinline__ int *colRow() const { return a.ints + P.intOff; }
    __device__ __forceinline__ int *rowMaxExp() const { return a.ints + P.intOff + P.n + 8; }
    __device__ __forceinline__ int *aligned() const { return a.ints + P.intOff + P.n + 8 + P.M + 2; }
    __device__ __forceinline__ int code(int j) const {
        return (j >= 0 && j < P.n) ? (int)a.codes[a.pairs[P.pairIdx].seqOff + j] : kZeroIdx;
    }
};

// Zero the tables and column arrays of every pair; rowMaxExp starts at INT_MIN.
__global__ void post_init_kernel(PostArgs a) {
    const int p = blockIdx.y;
    const PostPair P = a.pp[p];
    const int64_t matCells = 3 * (int64_t)(P.M + 2) * (P.n + 7), colCells = (int64_t)kPostCols * (P.n + 8);
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < matCells; c += stride)
        a.mat[P.matOff + c] = ES{0.f, 0};
    for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < colCells; c += stride)
        a.col[P.colOff + c] = ES{0.f, 0};
    const int nInt = 2 * (P.n + 8) + P.M + 2;
    for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < nInt; c += stride) {
        const bool rowMax = c >= P.n + 8 && c < P.n + 8 + P.M + 2;
        a.ints[P.intOff + c] = rowMax ? INT_MIN : 0;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) a.res[p] = PostRes{0ull, INT_MIN, INT_MIN, 0, 0, 0, 0};
}

__device__ __forceinline__ int wave_max_int(int v) {
    for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o));
    return v;
}
__device__ __forceinline__ unsigned long long wave_max_key(unsigned long long v) {
    for (int o = 32; o > 0; o >>= 1) {
        const unsigned long long w = __shfl_xor(v, o);
        v = w > v ? w : v;
    }
    return v;
}

// Per column j (one thread): over the rows in order, the anchor product
// w0 W1 and the column's best (first row on ties), the column totals of w0
// and W1 (fwdSide, revSide), and X(i,j) = (enter e(code(j-2)) w0(i,j-3)) W1(i+1,j)
// (post.cpp; dummest-dp-forward.hh:126-206). Row i of the dump is (strip s,
// lane k, slot q) with i = s G R + k R + q: the R rows of one (s, k) at
// column j are R consecutive floats of step s T + j + k, loaded together (the
// dump's steps lie ROWS floats apart, so a thread reads a whole chunk per R
// rows instead of one value per row). One launch per rows-per-lane R; pairs
// of another R return at once.
template <int R>
__global__ __launch_bounds__(256) void post_columns_kernel(PostArgs a) {
#pragma clang fp contract(off)
    const int p = blockIdx.y;
    const PostPair &P = a.pp[p];
    if (P.R != R || (int)(blockIdx.x * blockDim.x) >= P.n) return;
    const PairView V(P, a);
    const int M = P.M, n = P.n, G = P.G, T = P.T, ROWS = G * R;
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const float sC = a.pairs[P.pairIdx].sC;
    const ProfInfo PI = a.profs[P.prof];
    const float *w0b = a.w0buf + P.wOff, *w1b = a.wbuf + P.wOff;
    const int *efb = a.fbuf + P.eOff, *ebb = a.ebuf + P.eOff;
    ES best{0.f, 0}, fs{0.f, 0}, rs{0.f, 0}, gb{0.f, 0};
    int bestRow = 0;
    if (j < n) {
        const int cj = V.code(j - 2);
        const bool x3 = j >= 3;
        ES *X = V.X();
        // the chunk of group (s, k): w0 and W1 at column j, w0 at column j-3
        float A[R], B[R], C[R];
        int ea = 0, eb = 0, ec = 0;
        auto loadGroup = [&](int s, int k, float *Aa, float *Bb, float *Cc, int &a_, int &b_, int &c_) {
            const int64_t st = (int64_t)s * T + j + k;
            const float *pa = w0b + st * ROWS + k * R, *pb = w1b + st * ROWS + k * R;
#pragma unroll
            for (int q = 0; q < R; q++) {
                Aa[q] = pa[q];
                Bb[q] = pb[q];
            }
            a_ = efb[st];
            b_ = ebb[st];
            if (x3) {
                const float *pc = w0b + (st - 3) * ROWS + k * R;
#pragma unroll
                for (int q = 0; q < R; q++) Cc[q] = pc[q];
                c_ = efb[st - 3];
            }
        };
        loadGroup(0, 0, A, B, C, ea, eb, ec);
        int i = 0;
        for (int s = 0; i <= M; s++) {
            for (int k = 0; k < G && i <= M; k++) {
                // the next group's chunk, loaded ahead (its first W1 is row i+1 of this group's last row)
                float A2[R], B2[R], C2[R];
                int ea2 = 0, eb2 = 0, ec2 = 0;
                const bool more = i + R <= M;
                if (more) {
                    int s2 = s, k2 = k + 1;
                    if (k2 == G) {
                        k2 = 0;
                        ++s2;
                    }
                    loadGroup(s2, k2, A2, B2, C2, ea2, eb2, ec2);
                }
#pragma unroll
                for (int q = 0; q < R; q++) {
                    if (i + q > M) break;
                    const int row = i + q;
                    const ES wa = es_stored(A[q], ea), wb = es_stored(B[q], eb);
                    const ES wm = es_bmul(wa, wb);
                    if (es_less(gb, wm)) gb = wm;
                    if (es_less(best, wm)) {
                        best = wm;
                        bestRow = row;
                    }
                    fs = es_add(fs, wa);
                    rs = es_add(rs, wb);
                    if (row < M && x3) {
                        const ES w3 = es_stored(C[q], ec);
                        if (w3.m != 0.f) {
                            const ES wd = q + 1 < R ? es_stored(B[q + 1], eb) : es_stored(B2[0], eb2);
                            const float enterC = a.params[PI.parOff + (int64_t)row * 11 + 10] * sC;
                            const float xf = (enterC * a.emis[PI.emOff + (int64_t)row * kEmStride + cj]) * w3.m;
                            X[V.at(row, j)] = es_bmul(es_stored(xf, w3.e), wd);
                        }
                    }
                }
                i += R;
                if (more) {
#pragma unroll
                    for (int q = 0; q < R; q++) {
                        A[q] = A2[q];
                        B[q] = B2[q];
                        C[q] = C2[q];
                    }
                    ea = ea2;
                    eb = eb2;
                    ec = ec2;
                }
            }
        }
        V.col(kColBest)[j] = best;
        V.colRow()[j] = bestRow;
        V.col(kFwdSide)[j] = fs;
        V.col(kRevSide)[j] = rs;
        a.cands[P.candOff + j] = PostCand{best.m, best.e, bestRow, j};
    }
    const unsigned long long key = wave_max_key(es_key(gb));
    if ((threadIdx.x & 63) == 0 && key != 0ull) atomicMax(&a.res[p].gkey, key);
}

// The largest exponent of each row of X (rows 0..M-1): one workgroup per
// pair, its four wavefronts taking every fourth row.
__global__ __launch_bounds__(256) void post_rowmax_kernel(PostArgs a) {
    const int p = blockIdx.x, lane = threadIdx.x & 63;
    const PostPair &P = a.pp[p];
    const PairView V(P, a);
    const ES *X = V.X();
    int *rowMax = V.rowMaxExp();
    for (int i = threadIdx.x >> 6; i < P.M; i += 4) {
        int m = INT_MIN;
        for (int j = lane; j < P.n; j += 64) {
            const ES x = X[V.at(i, j)];
            if (x.m != 0.0f) m = max(m, x.e);
        }
        m = wave_max_int(m);
        if (lane == 0) rowMax[i] = m;
    }
}

// Per pair (three wavefronts): the gate, then the flank totals of
// buildFlankEVs (dummest-dp.hh:137-268) as post.cpp forms them, their running
// sums revEv / fwdEv and the largest exponents of those (sideEVMax,
// dummest-dp.hh:273-282). The per-column terms (entry weights, totals) are
// computed across all threads; the linear recurrences by the first lane of a
// wavefront each (mid, sr and fa at once, then the two running sums).
__global__ __launch_bounds__(192) void post_flanks_kernel(PostArgs a) {
#pragma clang fp contract(off)
    const int p = blockIdx.x, t = threadIdx.x, lead = (t & 63) == 0 ? t >> 6 : -1;
    const PostPair &P = a.pp[p];
    const PairView V(P, a);
    PostRes &R = a.res[p];
    const ES gbest = es_unkey(R.gkey);
    if (P.minRatio < 0.f) return;  // the best anchor only: no tables
    const ES minExp = es_norm((double)P.minRatio, 0);
    if (es_less(gbest, minExp)) return;
    const int W = P.n, L = P.n;
    const PairInfo PR = a.pairs[P.pairIdx];
    const Rec *rec = a.rec + PR.recOff;
    const float *bg = a.bg + 32 * (int64_t)P.prof;
    const double s1d = a.bgfs1, s2d = a.bgfs2;
    const double keep = 1.0 - s1d - s2d;
    const ES zero{0.f, 0};
    const ES aS = es_norm((double)PR.s1, 0), aD = es_norm((double)PR.s2, 0), aC = es_norm((double)PR.sC, 0);
    const ES bS = es_scale(aS, 4.0), bD = es_scale(aD, 16.0);
    auto bgc = [&](int pos) { return es_norm((double)bg[V.code(pos)], 0); };
    auto pm = [&](int q) { return es_norm((double)rec[q].ps, 0); };
    ES *mid = V.col(kMid), *sr = V.col(kSr), *uni = V.col(kUni), *revEv = V.col(kRevEv), *fwdEv = V.col(kFwdEv);
    ES *fa = V.col(kFa), *we = V.col(kWe);
    const ES *fwdSide = V.col(kFwdSide), *revSide = V.col(kRevSide);
    // entry weights eS + 2 eD + 3 eC of every column
    for (int q = t; q < W; q += 192) {
        const ES eS = q == 0 ? pm(0) : es_mul(es_scale(pm(q - 1), s1d), aS);
        const ES eD = q == 1 ? es_scale(aD, 1.0 / 3.0) : (q >= 2 ? es_mul(es_scale(pm(q - 2), s2d), aD) : zero);
        const ES eC = q == 2 ? es_mul(es_scale(aC, 1.0 / 3.0), bgc(0))
                             : (q >= 3 ? es_mul(es_scale(es_mul(pm(q - 3), aC), keep), bgc(q - 2)) : zero);
        we[q] = es_add(eS, es_add(es_scale(eD, 2.0), es_scale(eC, 3.0)));
    }
    __syncthreads();
    if (lead == 0) {
        // mid: the flank into each column from the right, through revSide
        ES m1 = zero, m2 = zero, m3 = zero;  // mid[q+1..q+3]
        for (int q = L - 1; q >= 0; q--) {
            ES c = es_mul(es_scale(m1, s1d), aS);
            c = es_add(c, es_mul(es_scale(m2, s2d), aD));
            c = es_add(c, es_mul(es_scale(es_mul(m3, aC), keep), bgc(q + 1)));
            const ES v = es_add(revSide[q], c);
            mid[q] = v;
            m3 = m2;
            m2 = m1;
            m1 = v;
        }
    } else if (lead == 1) {
        // sr: the null suffix after each column
        ES r1 = ES{0.5f, 1}, r2 = zero, r3 = zero;  // sr[q+1..q+3], sr[L] = 1
        for (int q = L - 1; q >= 0; q--) {
            const ES st = es_mul(es_scale(r1, s1d), aS);
            const ES dt = (q + 1 <= L - 1) ? es_mul(es_scale(r2, s2d), aD) : es_scale(bS, s2d * 0.25);
            ES ct;
            if (q + 2 <= L - 1) ct = es_mul(es_scale(es_mul(r3, aC), keep), bgc(q));
            else if (q + 1 <= L - 1) ct = es_scale(bD, keep * 0.0625);
            else ct = es_scale(bS, keep * 0.25);
            const ES v = es_add(st, es_add(dt, ct));
            sr[q] = v;
            r3 = r2;
            r2 = r1;
            r1 = v;
        }
        sr[L] = ES{0.5f, 1};
    } else if (lead == 2) {
        // fa: the flank from the left, through fwdSide
        ES f1 = zero, f2 = zero, f3 = zero;  // fa[q-1..q-3]
        for (int q = 0; q < W; q++) {
            ES av = fwdSide[q];
            if (q >= 1) av = es_add(av, es_mul(es_scale(f1, s1d), aS));
            if (q >= 2) av = es_add(av, es_mul(es_scale(f2, s2d), aD));
            if (q >= 3) av = es_add(av, es_mul(es_scale(es_mul(f3, aC), keep), bgc(q - 2)));
            fa[q] = av;
            f3 = f2;
            f2 = f1;
            f1 = av;
        }
    }
    __syncthreads();
    // the per-column totals
    for (int q = t; q < W; q += 192) {
        ES tot = es_mul(mid[q], we[q]);
        if (q >= 1 && q < L) tot = es_add(tot, es_mul(fa[q - 1], es_mul(es_scale(sr[q + 1], s1d), aS)));
        if (q >= 2 && q < L) tot = es_add(tot, es_scale(es_mul(fa[q - 2], es_mul(es_scale(sr[q + 1], s2d), aD)), 2.0));
        if (q >= 3 && q < L)
            tot = es_add(tot,
                         es_scale(es_mul(fa[q - 3], es_mul(es_scale(es_mul(sr[q + 1], aC), keep), bgc(q - 2))), 3.0));
        if (q == L - 1) {
            tot = es_add(tot, es_scale(es_mul(fa[q - 1], es_scale(bS, s2d * 0.25)), 2.0));
            tot = es_add(tot, es_scale(es_mul(fa[q - 1], es_scale(bS, keep * 0.25)), 3.0));
            if (q >= 2) tot = es_add(tot, es_scale(es_mul(fa[q - 2], es_scale(bD, keep * 0.0625)), 3.0));
        }
        uni[q] = es_scale(tot, 1.0 / 3.0);
    }
    __syncthreads();
    if (lead == 0) {
        int revMax = INT_MIN;
        ES cs = zero;
        for (int q = 0; q < W; q++) {
            cs = es_add(cs, uni[q]);
            revEv[q + 3] = cs;
            if (cs.m != 0.f && cs.e > revMax) revMax = cs.e;
        }
        R.revMax = revMax;
    } else if (lead == 1) {
        int fwdMax = INT_MIN;
        ES cs = zero;
        for (int q = W - 1; q >= 0; q--) {
            cs = es_add(cs, uni[q]);
            fwdEv[q] = cs;
            if (cs.m != 0.f && cs.e > fwdMax) fwdMax = cs.e;
        }
        R.fwdMax = fwdMax;
        R.live = 1;
    }
}

// The MEA tables of one pair (blockIdx.y = 0: prefixOptimal over rows 0..M,
// 1: suffixOptimal over rows M..0; dummest-dp-forward.hh:298-452). Each row
// at the row's scale E: an elementwise candidate per column, then the running
// maximum along the row (left to right for pre, right to left for suf) in
// chunks of 64 WV columns. WV = 1 (one wavefront) takes the pairs of at most
// kMeaShort letters and keeps the row before in LDS; WV = 4 takes the longer
// ones and reads it back from memory.
constexpr int kMeaShort = 1024;
template <int WV>
__global__ __launch_bounds__(64 * WV) void post_mea_kernel(PostArgs a) {
#pragma clang fp contract(off)
    constexpr bool kLds = WV == 1;
    __shared__ float sw[WV];
    __shared__ ES rows[kLds ? 2 : 1][kLds ? kMeaShort + 8 : 1];  // the row before and this row, column j at j+3
    const int p = blockIdx.x;
    const PostPair &P = a.pp[p];
    const PostRes &R = a.res[p];
    if (!R.live || (P.n > kMeaShort) != (WV > 1)) return;
    const PairView V(P, a);
    constexpr int CH = 64 * WV;
    const int M = P.M, W = P.n, t = threadIdx.x, lane = t & 63, wv = t >> 6;
    const bool fwd = blockIdx.y == 0;
    const ES *X = V.X();
    ES *T = fwd ? V.pre() : V.suf();
    const ES *ev = fwd ? V.col(kRevEv) : V.col(kFwdEv);
    const int *rowMax = V.rowMaxExp();
    const int side = fwd ? R.revMax : R.fwdMax;
    if constexpr (kLds) {
        for (int c = t; c < kMeaShort + 8; c += 64) rows[0][c] = rows[1][c] = ES{0.f, 0};
        __syncthreads();
    }
    int runMax = INT_MIN;
    for (int r = 0; r <= M; r++) {
        const int i = fwd ? r : M - r;
        ES *prevL = kLds ? rows[r & 1] : nullptr, *curL = kLds ? rows[(r & 1) ^ 1] : nullptr;
        // the row before (i-1 for pre, i+1 for suf) at column c
        auto prev = [&](int c) -> ES {
            if constexpr (kLds) return prevL[c + 3];
            else return T[V.at(fwd ? i - 1 : i + 1, c)];
        };
        runMax = max(runMax, rowMax[i]);
        const int E = max(runMax, side) + 1;
        float carry = 0.f;
        for (int c0 = 0; c0 < W; c0 += CH) {
            const int j = fwd ? c0 + t : W - 1 - c0 - t;
            const bool valid = fwd ? j < W : j >= 0;
            float b = 0.f;
            if (valid) {
                if (fwd) {
                    const float comb = es_to_row(X[V.at(i, j)], E);
                    const float rsr = es_to_row(ev[j + 3], E);
                    if (i == 0) {
                        b = fmaxf(comb, rsr);
                    } else {
                        const float cps = comb + es_to_row(prev(j - 3), E);
                        b = fmaxf(fmaxf(es_to_row(prev(j), E), cps), rsr);
                    }
                } else {
                    const float fsr = es_to_row(ev[j], E);
                    if (i == M) {
                        b = fsr;
                    } else {
                        const float cps = es_to_row(X[V.at(i, j + 3)], E) + es_to_row(prev(j + 3), E);
                        b = fmaxf(fmaxf(es_to_row(prev(j), E), fsr), cps);
                    }
                }
            }
            // running maximum within the wave, then across the waves
            for (int o = 1; o < 64; o <<= 1) {
                const float u = __shfl_up(b, o);
                if (lane >= o) b = fmaxf(b, u);
            }
            float mx;
            if constexpr (WV == 1) {
                mx = fmaxf(b, carry);
                carry = fmaxf(carry, __shfl(b, 63));
            } else {
                if (lane == 63) sw[wv] = b;
                __syncthreads();
                float in = carry;
                for (int w = 0; w < wv; w++) in = fmaxf(in, sw[w]);
                mx = fmaxf(b, in);
                for (int w = 0; w < WV; w++) carry = fmaxf(carry, sw[w]);
                __syncthreads();  // sw is rewritten by the next chunk
            }
            if (valid) {
                const ES v = es_from_row(mx, E);
                T[V.at(i, j)] = v;
                if constexpr (kLds) curL[j + 3] = v;
            }
        }
        __syncthreads();  // the row is complete before the next one reads it
    }
}

// The alignment of one anchor as a stack of segments (al of
// dummest-traceback.hh:6-118), kept in full in memory.
struct SegStack {
    PostSeg *s;
    int n = 0, cap;
    bool full = false;
    __device__ SegStack(PostSeg *p, int c) : s(p), cap(c) {}
    __device__ bool empty() const { return n == 0; }
    __device__ PostSeg back() const { return s[n - 1]; }
    __device__ void setBack(const PostSeg &x) { s[n - 1] = x; }
    __device__ void push(const PostSeg &x) {
        if (n < cap) s[n++] = x;
        else full = true;
    }
    __device__ void reverse() {
        for (int x = 0, y = n - 1; x < y; x++, y--) {
            const PostSeg t = s[x];
            s[x] = s[y];
            s[y] = t;
        }
    }
    __device__ int size() const { return n; }
    __device__ bool stop() { return false; }
};

// The same stack reduced to what decides whether an alignment is kept: its
// first and last segments and its length. The forward walk pushes f1..fa (a
// merge extends the newest); after the first reversal the back is f1; the
// backward walk pushes r1..rb (a merge changes the back: f1 before r1 exists);
// after the second reversal the stack reads rb..r1 f1..fa.
//
// The columns an alignment spans, [front start2, last start2 + length), only
// grow as its walks go on; stop() reports (and then ends the walks) when they
// come to reach a column an alignment kept earlier holds, which rejects this
// one whatever follows.
struct SegEnds {
    PostSeg f1{0, 0, 0}, fa{0, 0, 0}, r1{0, 0, 0}, rb{0, 0, 0};
    int a = 0, b = 0;
    bool reversed = false;  // the forward walk is done (the first reversal)
    bool full = false;
    volatile const int *aligned;
    int n;
    int clo = 0, chi = 0;  // the columns checked so far
    bool blocked = false;
    __device__ SegEnds(volatile const int *al, int n_) : aligned(al), n(n_) {}
    __device__ bool stop() {
        if (blocked || a + b == 0) return blocked;
        const PostSeg f = front(), l = last();
        const int lo = max(f.start2, 0), hi = min(l.start2 + l.length, n);
        if (clo == chi) {  // first segment
            for (int q = lo; q < hi && !blocked; q++) blocked = aligned[q] != 0;
            clo = lo;
            chi = max(lo, hi);
        } else {
            for (int q = lo; q < clo && !blocked; q++) blocked = aligned[q] != 0;
            for (int q = max(chi, lo); q < hi && !blocked; q++) blocked = aligned[q] != 0;
            clo = min(clo, lo);
            chi = max(chi, hi);
        }
        return blocked;
    }
    __device__ bool empty() const { return a + b == 0; }
    __device__ PostSeg back() const { return !reversed ? fa : (b > 0 ? rb : f1); }
    __device__ void setBack(const PostSeg &x) {
        if (!reversed) {
            fa = x;
            if (a == 1) f1 = x;
        } else if (b > 0) {
            rb = x;
            if (b == 1) r1 = x;
        } else {
            f1 = x;
            if (a == 1) fa = x;
        }
    }
    __device__ void push(const PostSeg &x) {
        if (!reversed) {
            if (a == 0) f1 = x;
            fa = x;
            a++;
        } else {
            if (b == 0) r1 = x;
            rb = x;
            b++;
        }
    }
    __device__ void reverse() { reversed = true; }  // only the first reversal changes what is tracked
    __device__ int size() const { return a + b; }
    __device__ PostSeg front() const { return b > 0 ? rb : f1; }
    __device__ PostSeg last() const { return a > 0 ? fa : r1; }
};

// The forward walk over suffixOptimal and the backward walk over
// prefixOptimal from anchor (a1, a2) (dummest-traceback.hh:35-95), with the
// reversals between and after them.
template <class Stack>
__device__ void trace(const PairView &V, const ES *X, const ES *pre, const ES *suf, const ES *fwdEv, const ES *revEv,
                      int a1, int a2, Stack &al) {
    const int M = V.P.M, n = V.P.n, cols = n + 4, bound = n;
    {
        int i = a1, q = a2;
        while (i <= M && q < bound) {
            ES best{0.f, 0};
            int bi = INT_MAX, bj = INT_MAX;
            bool em = false;
            auto consider = [&](const ES &v, int ni, int nj, bool e) {
                if (es_less(best, v)) {
                    best = v;
                    bi = ni;
                    bj = nj;
                    em = e;
                }
            };
            if (q + 3 < cols) consider(es_add(X[V.at(i, q + 3)], suf[V.at(i + 1, q + 3)]), i + 1, q + 3, true);
            consider(suf[V.at(i + 1, q)], i + 1, q, false);
            consider(suf[V.at(i, q + 1)], i, q + 1, false);
            consider(fwdEv[q], INT_MAX, INT_MAX, false);
            if (em) {
                const int p1 = i, p2 = q + 1;
                bool push = true;
                if (!al.empty()) {
                    PostSeg x = al.back();
                    if (x.start1 + x.length == p1 && x.start2 + x.length == p2) {
                        ++x.length;
                        al.setBack(x);
                    }
                    if (x.start1 + x.length > p1 || x.start2 + x.length > p2) push = false;
                }
                if (push) al.push(PostSeg{p1, p2, 1});
                if (al.stop()) return;
            }
            i = bi;
            q = bj;
        }
    }
    al.reverse();
    {
        int i = a1 - 1, q = a2;
        while (i >= 0 && q >= 0) {
            ES best{0.f, 0};
            int bi = -1, bj = -1;
            bool em = false;
            auto consider = [&](const ES &v, int ni, int nj, bool e) {
                if (es_less(best, v)) {
                    best = v;
                    bi = ni;
                    bj = nj;
                    em = e;
                }
            };
            if (q >= 3) {
                const ES os = (i >= 1) ? pre[V.at(i - 1, q - 3)] : ES{0.f, 0};
                consider(es_add(X[V.at(i, q)], os), i - 1, q - 3, true);
            }
            if (i >= 1) consider(pre[V.at(i - 1, q)], i - 1, q, false);
            if (q > 0) consider(pre[V.at(i, q - 1)], i, q - 1, false);
            consider(revEv[q + 3], -1, -1, false);
            if (em) {
                const int p1 = i, p2 = q - 2;
                bool push = true;
                if (!al.empty()) {
                    PostSeg x = al.back();
                    if (x.start1 - 1 == p1 && x.start2 - 1 == p2) {
                        --x.start1;
                        --x.start2;
                        ++x.length;
                        al.setBack(x);
                    }
                    if (x.start1 <= p1 || x.start2 <= p2) push = false;
                }
                if (push) al.push(PostSeg{p1, p2, 1});
                if (al.stop()) return;
            }
            i = bi;
            q = bj;
        }
    }
    al.reverse();
}

// Per pair (one wavefront): the ranked anchors in order; an anchor whose
// column an earlier alignment holds is skipped, the others are traced and
// kept when their columns overlap no earlier alignment
// (dummest-dp-forward.hh:500-545). The lanes trace 64 anchors at a time
// (their ends only), then take the keep-or-skip decisions one anchor after
// another in rank order, so the result is the one-at-a-time result; the
// kept anchors are traced again, writing their segments.
__global__ __launch_bounds__(64) void post_walk_kernel(PostArgs a) {
    const int p = blockIdx.x, lane = threadIdx.x;
    const PostPair &P = a.pp[p];
    PostRes &R = a.res[p];
    if (P.ncand == 0) return;
    const PairView V(P, a);
    const int n = P.n;
    const ES *X = V.X(), *pre = V.pre(), *suf = V.suf();
    const ES *fwdEv = V.col(kFwdEv), *revEv = V.col(kRevEv);
    // read and written by the lanes in turn: volatile, so every access goes to memory
    volatile int *aligned = V.aligned();
    PostSeg *out = a.segs + P.segOff;
    PostHit *hits = a.hits + P.hitOff;
    const PostCand *cand = a.cands + P.candOff;
    int nh = 0, ns = 0;  // uniform
    bool overflow = false;
    for (int c0 = 0; c0 < P.ncand; c0 += 64) {
        const int ci = c0 + lane;
        const bool has = ci < P.ncand;
        const PostCand c = has ? cand[ci] : PostCand{0.f, 0, 0, 0};
        // lanes whose anchor an earlier alignment already holds stay idle
        bool walk = has && !aligned[c.a2];
        int cb = 0, ce = 0, nal = 0;
        if (walk) {
            SegEnds ends(aligned, n);
            trace(V, X, pre, suf, fwdEv, revEv, c.a1, c.a2, ends);
            walk = !ends.blocked;  // reaches an alignment of an earlier round: rejected
            nal = ends.size();
            const int beg = nal == 0 ? c.a2 : ends.front().start2;
            const PostSeg l = ends.last();
            const int en = nal == 0 ? c.a2 : l.start2 + l.length;
            cb = max(beg, 0);
            ce = min(en, n);
        }
        // keep-or-skip in rank order (every lane takes part in each decision)
        int seg0 = -1, slot = -1;
        for (int l = 0; l < 64; l++) {
            if (!__shfl(walk ? 1 : 0, l)) continue;
            const int a2l = __shfl(c.a2, l), cbl = __shfl(cb, l), cel = __shfl(ce, l), nall = __shfl(nal, l);
            __threadfence_block();
            if (aligned[a2l]) continue;  // held by an alignment kept earlier in this round
            bool ov = false;
            for (int q = cbl + lane; q < cel; q += 64) ov |= aligned[q] != 0;
            if (__any(ov)) continue;
            if (ns + nall > P.segCap || nh >= P.ncand) {
                overflow = true;
                break;
            }
            for (int q = cbl + lane; q < cel; q += 64) aligned[q] = 1;
            __threadfence_block();
            if (lane == l) {
                seg0 = ns;
                slot = nh;
            }
            ns += nall;
            nh++;
        }
        if (overflow) break;
        if (slot >= 0) {  // kept: trace again, writing the segments
            SegStack st(out + seg0, nal);
            trace(V, X, pre, suf, fwdEv, revEv, c.a1, c.a2, st);
            hits[slot] = PostHit{c.m, c.e, c.a1, c.a2, seg0, nal};
            if (st.full || st.size() != nal) overflow = true;
        }
        if (__any(overflow)) {
            overflow = true;
            break;
        }
    }
    if (lane == 0) {
        R.nhits = nh;
        R.nsegs = ns;
        if (overflow) R.overflow = 1;
    }
}

}  // namespace

hipError_t launch_post_tables(const PostArgs &a, int maxN, int maxM, int64_t maxCells, hipStream_t s) {
    if (a.npairs <= 0) return hipSuccess;
    const int initBlocks = (int)std::min<int64_t>(1024, (maxCells + 255) / 256);
    hipLaunchKernelGGL(post_init_kernel, dim3(std::max(1, initBlocks), a.npairs), dim3(256), 0, s, a);
    const dim3 cgrid((maxN + 255) / 256, a.npairs);
    hipLaunchKernelGGL(post_columns_kernel<1>, cgrid, dim3(256), 0, s, a);
    hipLaunchKernelGGL(post_columns_kernel<2>, cgrid, dim3(256), 0, s, a);
    hipLaunchKernelGGL(post_columns_kernel<3>, cgrid, dim3(256), 0, s, a);
    hipLaunchKernelGGL(post_columns_kernel<4>, cgrid, dim3(256), 0, s, a);
    hipLaunchKernelGGL(post_columns_kernel<5>, cgrid, dim3(256), 0, s, a);
    hipLaunchKernelGGL(post_columns_kernel<6>, cgrid, dim3(256), 0, s, a);
    hipLaunchKernelGGL(post_columns_kernel<7>, cgrid, dim3(256), 0, s, a);
    hipLaunchKernelGGL(post_columns_kernel<8>, cgrid, dim3(256), 0, s, a);
    hipLaunchKernelGGL(post_columns_kernel<12>, cgrid, dim3(256), 0, s, a);
    hipLaunchKernelGGL(post_columns_kernel<16>, cgrid, dim3(256), 0, s, a);
    hipLaunchKernelGGL(post_rowmax_kernel, dim3(a.npairs), dim3(256), 0, s, a);
    (void)maxM;
    hipLaunchKernelGGL(post_flanks_kernel, dim3(a.npairs), dim3(192), 0, s, a);
    hipLaunchKernelGGL(post_mea_kernel<1>, dim3(a.npairs, 2), dim3(64), 0, s, a);
    if (maxN > kMeaShort) hipLaunchKernelGGL(post_mea_kernel<4>, dim3(a.npairs, 2), dim3(256), 0, s, a);
    return hipGetLastError();
}

bool post_rows_supported(int R) { return (R >= 1 && R <= 8) || R == 12 || R == 16; }

hipError_t launch_post_walk(const PostArgs &a, hipStream_t s) {
    if (a.npairs <= 0) return hipSuccess;
    hipLaunchKernelGGL(post_walk_kernel, dim3(a.npairs), dim3(64), 0, s, a);
    return hipGetLastError();
}

}  // namespace dm
