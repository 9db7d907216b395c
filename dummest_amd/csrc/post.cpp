// Post-gate processing of one hit pair on the host, over the w0/W1 matrices
// of the pair as exact ExpScores (the dump-mode kernel's step-major values and
// per-step exponents, converted by the engine).
//
// Follows forwardPass's epilogue: anchor products and per-column best
// (dummest-dp-forward.hh:174-206), the posterior match matrix
// X(i,j) = X_ij * W1(i+1,j) (:126-156), gate (:252-274), flank totals
// (buildFlankEVs, dummest-dp.hh:137-268), the row-scaled MEA tables (:294-452,
// dummest-scalerow.hh), anchor ranking and traceback (:456-545,
// dummest-traceback.hh:6-118). Column totals are summed as exact ExpScores
// instead of the reference's rescaled float accumulators.
#include <algorithm>
#include <climits>
#include <cmath>
#include <functional>

#include "dm_host.h"

namespace dm {

namespace {

inline ES norm(double m, int e) {
    if (m == 0.0) return ES{0.f, 0};
    int ex = 0;
    m = std::frexp(m, &ex);
    return ES{(float)m, e + ex};
}
inline bool less(const ES &a, const ES &b) {
    if (a.m == 0.0f) return b.m != 0.0f;
    if (b.m == 0.0f) return false;
    if (a.e != b.e) return a.e < b.e;
    return a.m < b.m;
}
inline ES scale(const ES &a, double f) {
    if (a.m == 0.0f || !(f > 0.0) || !std::isfinite(f)) return ES{0.f, 0};
    return norm((double)a.m * f, a.e);
}
inline ES add(ES a, ES b) {
    if (a.m == 0.0f) return b;
    if (b.m == 0.0f) return a;
    if (a.e >= b.e) {
        const int d = a.e - b.e;
        if (d > 60) return a;
        return norm((double)a.m + std::ldexp((double)b.m, -d), a.e);
    }
    const int d = b.e - a.e;
    if (d > 60) return b;
    return norm((double)b.m + std::ldexp((double)a.m, -d), b.e);
}
inline ES mul(const ES &a, const ES &b) {
    if (a.m == 0.0f || b.m == 0.0f) return ES{0.f, 0};
    return norm((double)a.m * (double)b.m, a.e + b.e);
}
inline ES bmul(const ES &A, const ES &B) {  // dummest-batch.hh:216-235
    if (A.m == 0.0f || B.m == 0.0f) return ES{0.f, 0};
    double p = (double)A.m * (double)B.m;
    int e = A.e + B.e;
    if (p < 0.5) {
        p *= 2.0;
        e -= 1;
    }
    return ES{(float)p, e};
}
}  // namespace

ES esFromStored(float v, int off) {  // promote a stored value at a power-of-two offset
    if (!(v > 0.f)) return ES{0.f, 0};
    if (!(v < INFINITY)) return ES{std::nextafterf(1.0f, 0.0f), 1023 + off};
    int ex;
    const float m = std::frexp(v, &ex);
    return ES{m, ex + off};
}
namespace {

inline ES fromFloat(float v, int off) { return esFromStored(v, off); }

inline double lnOf(const ES &a) { return a.m == 0.0f ? -INFINITY : std::log(a.m) + (double)a.e * 0.6931471805599453; }

inline float toRow(const ES &s, int E) {  // scoreToRowScale, dummest-scalerow.hh:26-38
    if (s.m == 0.0f) return 0.0f;
    int d = s.e - E;
    d = d > 0 ? 0 : (d < -1000 ? -1000 : d);
    const int ne = 126 + d;
    if (ne <= 0) return 0.0f;
    uint32_t u;
    std::memcpy(&u, &s.m, 4);
    u = (u & 0x807FFFFFu) | ((uint32_t)ne << 23);
    float r;
    std::memcpy(&r, &u, 4);
    return r;
}
inline ES fromRow(float s, int E) {  // scoreFromRowScale, dummest-scalerow.hh:43-55
    uint32_t u;
    std::memcpy(&u, &s, 4);
    const uint32_t ef = (u >> 23) & 0xFFu;
    if (!(s > 0.0f) || ef == 0) return ES{0.f, 0};
    u = (u & 0x807FFFFFu) | (126u << 23);
    float m;
    std::memcpy(&m, &u, 4);
    return ES{m, (int)ef - 126 + E};
}

struct Mat {  // ExpScore matrix with a front pad, zero outside the written range
    int rows = 0, cols = 0, pre = 0;
    std::vector<ES> d;
    void init(int r, int c, int p) {
        rows = r;
        cols = c;
        pre = p;
        d.assign((size_t)r * (c + p), ES{0.f, 0});
    }
    ES &at(int i, int j) { return d[(size_t)i * (cols + pre) + j + pre]; }
    ES get(int i, int j) const { return d[(size_t)i * (cols + pre) + j + pre]; }
};

}  // namespace

void rankAnchors(const ES *colBest, const int *colRow, int n, float minProbRatio, std::vector<Anchor> &out) {
    struct Cand {
        ES pair;
        int a1, a2;
        bool operator>(const Cand &o) const { return less(o.pair, pair); }
        bool operator<(const Cand &o) const { return less(pair, o.pair); }
    };
    out.clear();
    std::vector<Cand> opt(n, Cand{ES{0.f, 0}, 0, 0});
    ES bestOverall{0.f, 0};
    for (int j = 0; j < n; j++)
        if (colBest[j].m != 0.f) {
            if (less(bestOverall, colBest[j])) bestOverall = colBest[j];
            opt[j] = Cand{colBest[j], colRow[j], j};
        }
    if (minProbRatio < 0) {
        const Cand sel = *std::max_element(opt.begin(), opt.end());
        out.push_back(Anchor{sel.pair, sel.a1, sel.a2});
        return;
    }
    const ES minExp = norm((double)minProbRatio, 0);
    if (bestOverall.m == 0.f || less(bestOverall, minExp)) return;
    // Highest first, the anchors at or above the threshold are a prefix of the
    // sorted columns. When no two of them tie, that prefix is the same however
    // it is sorted, so only those are sorted; a tie among them falls back to
    // std::sort over every column, whose order of tied elements is the
    // reference's.
    std::vector<Cand> top;
    for (const Cand &c : opt)
        if (c.pair.m != 0.f && !less(c.pair, minExp)) top.push_back(c);
    std::sort(top.begin(), top.end(), std::greater<>());
    bool tie = false;
    for (size_t k = 1; k < top.size() && !tie; k++) tie = !less(top[k].pair, top[k - 1].pair);
    if (tie) {
        std::sort(opt.begin(), opt.end(), std::greater<>());
        top.clear();
        for (const Cand &c : opt) {
            if (c.pair.m == 0.f || less(c.pair, minExp)) break;
            top.push_back(c);
        }
    }
    for (const Cand &c : top) out.push_back(Anchor{c.pair, c.a1, c.a2});
}

void postProcess(const HostProfile &P, const PostInput &in, float minProbRatio, std::vector<Hit> &out) {
    const int M = in.M, n = in.n;
    auto code = [&](int j) -> int { return (j >= 0 && j < n) ? (int)in.codes[j] : kZeroIdx; };
    auto w0at = [&](int i, int j) -> ES {
        if (j < 0 || j >= n || i < 0 || i > M) return ES{0.f, 0};
        return in.w0[(size_t)i * n + j];
    };
    auto W1at = [&](int i, int j) -> ES {
        if (j < 0 || j >= n || i < 0 || i > M) return ES{0.f, 0};
        return in.W1[(size_t)i * n + j];
    };
    // anchors, X, column totals
    std::vector<ES> colBest(n, ES{0.f, 0});
    std::vector<int> colRow(n, 0);
    std::vector<ES> fwdSide(n, ES{0.f, 0}), revSide(n, ES{0.f, 0});
    std::vector<int> rowMaxExp(M + 2, INT_MIN);
    Mat X;
    X.init(M + 2, n + 4, 0);
    ES gbest{0.f, 0};
    for (int i = 0; i <= M; i++) {
        const float *pr = &P.params[(size_t)i * 11];
        const float enterC = pr[10] * in.sC;
        const float *em = &P.values[(size_t)i * kWidth + 4];
        for (int j = 0; j < n; j++) {
            const ES a = w0at(i, j), b = W1at(i, j);
            const ES wm = bmul(a, b);
            if (less(gbest, wm)) gbest = wm;
            if (less(colBest[j], wm)) {
                colBest[j] = wm;
                colRow[j] = i;
            }
            fwdSide[j] = add(fwdSide[j], a);
            revSide[j] = add(revSide[j], b);
            if (i < M && j >= 3) {
                const ES w3 = w0at(i, j - 3);
                if (w3.m != 0.f) {
                    const float xf = (enterC * em[code(j - 2)]) * w3.m;
                    const ES xs = fromFloat(xf, w3.e);
                    const ES x = bmul(xs, W1at(i + 1, j));
                    X.at(i, j) = x;
                    if (x.m != 0.0f && x.e > rowMaxExp[i]) rowMaxExp[i] = x.e;
                }
            }
        }
    }
    ES minExp{0.f, 0};
    if (minProbRatio >= 0) minExp = norm((double)minProbRatio, 0);
    if (!(minProbRatio < 0 || !less(gbest, minExp))) return;

    // flank totals, dummest-dp.hh:137-268
    const double s1d = in.bgfs1, s2d = in.bgfs2;
    const double keep = 1.0 - s1d - s2d;
    const ES zero{0.f, 0};
    const ES aS = norm((double)in.s1, 0), aD = norm((double)in.s2, 0), aC = norm((double)in.sC, 0);
    const ES bS = scale(aS, 4.0), bD = scale(aD, 16.0);
    auto bgc = [&](int pos) { return norm((double)P.bg[4 + code(pos)], 0); };
    const int W = n, L = n;
    std::vector<ES> pm(W), eS(W, zero), eD(W, zero), eC(W, zero);
    for (int p = 0; p < W; p++) pm[p] = norm((double)in.prefixSeed[p], 0);
    eS[0] = pm[0];
    eD[1] = scale(aD, 1.0 / 3.0);
    eC[2] = mul(scale(aC, 1.0 / 3.0), bgc(0));
    for (int p = 1; p < W; p++) eS[p] = mul(scale(pm[p - 1], s1d), aS);
    for (int p = 2; p < W; p++) eD[p] = mul(scale(pm[p - 2], s2d), aD);
    for (int p = 3; p < W; p++) eC[p] = mul(scale(mul(pm[p - 3], aC), keep), bgc(p - 2));
    std::vector<ES> mid(W + 3, zero);
    for (int p = L - 1; p >= 0; p--) {
        ES c = mul(scale(mid[p + 1], s1d), aS);
        c = add(c, mul(scale(mid[p + 2], s2d), aD));
        c = add(c, mul(scale(mul(mid[p + 3], aC), keep), bgc(p + 1)));
        mid[p] = add(revSide[p], c);
    }
    std::vector<ES> sr(W + 4, zero);
    sr[L] = ES{0.5f, 1};
    for (int p = L - 1; p >= 0; p--) {
        const ES st = mul(scale(sr[p + 1], s1d), aS);
        const ES dt = (p + 1 <= L - 1) ? mul(scale(sr[p + 2], s2d), aD) : scale(bS, s2d * 0.25);
        ES ct;
        if (p + 2 <= L - 1) ct = mul(scale(mul(sr[p + 3], aC), keep), bgc(p));
        else if (p + 1 <= L - 1) ct = scale(bD, keep * 0.0625);
        else ct = scale(bS, keep * 0.25);
        sr[p] = add(st, add(dt, ct));
    }
    std::vector<ES> fa(W, zero);
    for (int p = 0; p < W; p++) {
        ES a = fwdSide[p];
        if (p >= 1) a = add(a, mul(scale(fa[p - 1], s1d), aS));
        if (p >= 2) a = add(a, mul(scale(fa[p - 2], s2d), aD));
        if (p >= 3) a = add(a, mul(scale(mul(fa[p - 3], aC), keep), bgc(p - 2)));
        fa[p] = a;
    }
    std::vector<ES> uni(W, zero);
    for (int p = 0; p < W; p++) {
        const ES we = add(eS[p], add(scale(eD[p], 2.0), scale(eC[p], 3.0)));
        ES tot = mul(mid[p], we);
        if (p >= 1 && p < L) tot = add(tot, mul(fa[p - 1], mul(scale(sr[p + 1], s1d), aS)));
        if (p >= 2 && p < L) tot = add(tot, scale(mul(fa[p - 2], mul(scale(sr[p + 1], s2d), aD)), 2.0));
        if (p >= 3 && p < L)
            tot = add(tot, scale(mul(fa[p - 3], mul(scale(mul(sr[p + 1], aC), keep), bgc(p - 2))), 3.0));
        if (p == L - 1) {
            tot = add(tot, scale(mul(fa[p - 1], scale(bS, s2d * 0.25)), 2.0));
            tot = add(tot, scale(mul(fa[p - 1], scale(bS, keep * 0.25)), 3.0));
            if (p >= 2) tot = add(tot, scale(mul(fa[p - 2], scale(bD, keep * 0.0625)), 3.0));
        }
        uni[p] = scale(tot, 1.0 / 3.0);
    }
    std::vector<ES> revEv(W + 7, zero), fwdEv(W + 4, zero);  // revEv indexed p+3
    ES cs = zero;
    for (int p = 0; p < W; p++) revEv[p + 3] = cs = add(cs, uni[p]);
    cs = zero;
    for (int p = W - 1; p >= 0; p--) fwdEv[p] = cs = add(cs, uni[p]);
    int revMax = INT_MIN, fwdMax = INT_MIN;
    for (int j = 0; j < W; j++) {
        if (revEv[j + 3].m != 0.f && revEv[j + 3].e > revMax) revMax = revEv[j + 3].e;
        if (fwdEv[j].m != 0.f && fwdEv[j].e > fwdMax) fwdMax = fwdEv[j].e;
    }
    // MEA tables, dummest-dp-forward.hh:294-452
    Mat pre, suf;
    pre.init(M + 2, W + 4, 3);
    suf.init(M + 2, W + 4, 0);
    {
        int runMax = INT_MIN;
        for (int i = 0; i <= M; i++) {
            runMax = std::max(runMax, rowMaxExp[i]);
            const int E = std::max(runMax, revMax) + 1;
            float left = 0.f;
            for (int j = 0; j < W; j++) {
                const float comb = toRow(X.get(i, j), E);
                const float rsr = toRow(revEv[j + 3], E);
                float mx;
                if (i == 0) {
                    mx = 0.f;
                    if (left > mx) mx = left;
                    if (comb > mx) mx = comb;
                    if (rsr > mx) mx = rsr;
                } else {
                    const float cps = comb + toRow(pre.get(i - 1, j - 3), E);
                    mx = toRow(pre.get(i - 1, j), E);
                    if (left > mx) mx = left;
                    if (cps > mx) mx = cps;
                    if (rsr > mx) mx = rsr;
                }
                left = mx;
                pre.at(i, j) = fromRow(mx, E);
            }
        }
    }
    {
        int runMax = INT_MIN;
        for (int i = M; i >= 0; i--) {
            runMax = std::max(runMax, rowMaxExp[i]);
            const int E = std::max(runMax, fwdMax) + 1;
            float right = 0.f;
            for (int j = W - 1; j >= 0; j--) {
                const float fsr = toRow(fwdEv[j], E);
                float mx;
                if (i == M) {
                    mx = 0.f;
                    if (right > mx) mx = right;
                    if (fsr > mx) mx = fsr;
                } else {
                    const float cps = toRow(X.get(i, j + 3), E) + toRow(suf.get(i + 1, j + 3), E);
                    mx = toRow(suf.get(i + 1, j), E);
                    if (right > mx) mx = right;
                    if (fsr > mx) mx = fsr;
                    if (cps > mx) mx = cps;
                }
                right = mx;
                suf.at(i, j) = fromRow(mx, E);
            }
        }
    }
    // anchors and traceback, dummest-dp-forward.hh:459-545
    std::vector<Anchor> ranked;
    rankAnchors(colBest.data(), colRow.data(), n, minProbRatio, ranked);
    if (minProbRatio < 0) {
        const Anchor &sel = ranked.front();
        Hit h;
        h.pair = sel.pair;
        h.lnratio = lnOf(sel.pair);
        h.anchor1 = sel.a1;
        h.anchor2 = sel.a2;
        out.push_back(h);
        return;
    }
    std::vector<bool> aligned(n, false);
    const int cols = W + 4, bound = W;
    for (const Anchor &c : ranked) {
        if (aligned[c.a2]) continue;
        Hit h;
        h.pair = c.pair;
        h.lnratio = lnOf(c.pair);
        h.anchor1 = c.a1;
        h.anchor2 = c.a2;
        std::vector<Seg> &al = h.segs;
        // forward walk over suffixOptimal, dummest-traceback.hh:35-66
        {
            int i = c.a1, p = c.a2;
            while (i <= M && p < bound) {
                ES best{0.f, 0};
                int bi = INT_MAX, bj = INT_MAX;
                bool em = false;
                auto consider = [&](const ES &v, int ni, int nj, bool e) {
                    if (less(best, v)) { best = v; bi = ni; bj = nj; em = e; }
                };
                if (p + 3 < cols) consider(add(X.get(i, p + 3), suf.get(i + 1, p + 3)), i + 1, p + 3, true);
                consider(suf.get(i + 1, p), i + 1, p, false);
                consider(suf.get(i, p + 1), i, p + 1, false);
                consider(fwdEv[p], INT_MAX, INT_MAX, false);
                if (em) {
                    const int p1 = i, p2 = p + 1;
                    bool push = true;
                    if (!al.empty()) {
                        Seg &x = al.back();
                        if (x.start1 + x.length == p1 && x.start2 + x.length == p2) ++x.length;
                        if (x.start1 + x.length > p1 || x.start2 + x.length > p2) push = false;
                    }
                    if (push) al.push_back(Seg{p1, p2, 1});
                }
                i = bi;
                p = bj;
            }
        }
        std::reverse(al.begin(), al.end());
        // reverse walk over prefixOptimal, dummest-traceback.hh:68-95
        {
            int i = c.a1 - 1, p = c.a2;
            while (i >= 0 && p >= 0) {
                ES best{0.f, 0};
                int bi = -1, bj = -1;
                bool em = false;
                auto consider = [&](const ES &v, int ni, int nj, bool e) {
                    if (less(best, v)) { best = v; bi = ni; bj = nj; em = e; }
                };
                if (p >= 3) {
                    const ES os = (i >= 1) ? pre.get(i - 1, p - 3) : ES{0.f, 0};
                    consider(add(X.get(i, p), os), i - 1, p - 3, true);
                }
                if (i >= 1) consider(pre.get(i - 1, p), i - 1, p, false);
                if (p > 0) consider(pre.get(i, p - 1), i, p - 1, false);
                consider(revEv[p + 3], -1, -1, false);
                if (em) {
                    const int p1 = i, p2 = p - 2;
                    bool push = true;
                    if (!al.empty()) {
                        Seg &x = al.back();
                        if (x.start1 - 1 == p1 && x.start2 - 1 == p2) {
                            --x.start1;
                            --x.start2;
                            ++x.length;
                        }
                        if (x.start1 <= p1 || x.start2 <= p2) push = false;
                    }
                    if (push) al.push_back(Seg{p1, p2, 1});
                }
                i = bi;
                p = bj;
            }
        }
        std::reverse(al.begin(), al.end());
        const int beg = al.empty() ? c.a2 : al.front().start2;
        const int en = al.empty() ? c.a2 : al.back().start2 + al.back().length;
        const int cb = std::max(beg, 0), ce = std::min(en, n);
        bool ov = false;
        for (int p = cb; p < ce; ++p)
            if (aligned[p]) {
                ov = true;
                break;
            }
        if (ov) continue;
        std::fill(aligned.begin() + cb, aligned.begin() + ce, true);
        out.push_back(std::move(h));
    }
}

}  // namespace dm
