// Gradient of the summed forward log-likelihood on gfx950: for each (profile,
// DNA sequence) pair the total Z of all local alignments,
// Z = sum_{i,j} prefixSeed(j) W1(i, j) = sum_{i,j} w0(i, j) suffixSeed(j),
// and the expected number of uses of every parameter of the profile,
// d ln Z / d ln theta = (1/Z) sum over the uses of theta of
// (forward value into the use) theta (backward value out of it).
//
// The recurrences are the reference's (forward dummest-dp-forward.hh:9-250,
// backward dummest-dp-backward.hh:444-615) in FP64. In place of the reference's
// per-row rescaling (dummest-dp-backward.hh:560-573), every 64-row strip and
// 64-column block of a pass has its own power-of-two exponent. Per use (row i,
// column j):
//   match:   w0(i,j) enter_i sC e_i(code(j+1)) W1(i+1, j+3)   (enter_i, and the
//            emission of the codon at j+1)
//   delete:  w0(i,j) d0_i Yb(i+1,j);  Y(i-1,j) eps_i Yb(i+1,j)
//   frameshift deletes: w0(i,j) d1_i s2 W1(i+1,j+2);  w0(i,j) d2_i s1 W1(i+1,j+1)
//   frameshift inserts: w0(i,j) a1_i s1 W1(i,j+1);  w0(i,j) a2_i s2 W1(i,j+2)
//   insert:  w0(i,j) a0_i sC g(j+1) Zb(i,j+3);  Z0(i,j) b0_i sC g(j+1) Zb(i,j+3)
// where Y, Z0 are the forward delete and insert chains and Yb, Zb their
// backward counterparts (the adjoints of Y and Z0).
//
// Layout: one wavefront per pair for forward and backward, profile rows 64 at
// a time (lane k = one row, one column behind the row it depends on), the
// values from that row arriving by DPP; the first row of a strip reads the
// neighbouring strip's edge row, which every strip keeps, in 64-column chunks.
// The strip's emission rows sit in LDS, the pair's per-column codes and seeds
// in a 256-column LDS ring. A pass stores the matrices of the strips it runs
// diagonal-major (each step's 64 values contiguous), or only their edge rows
// (the checkpoints from which a run of strips is computed again). Z comes
// from the column sums of W1; the count pass gives each (pair, strip) a
// wavefront whose lanes walk the forward's anti-diagonals, so every read is a
// contiguous run, with the codon-indexed emission sums in LDS.
#include <hip/hip_runtime.h>

#include <type_traits>

#include "grad.h"

namespace dm {
namespace {

__device__ __forceinline__ double lane_from_above(double v) {  // lane k gets lane k-1's value (DPP wave_shr:1)
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), 0x138, 0xf, 0xf, true);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), 0x138, 0xf, 0xf, true);
    return __hiloint2double(hi, lo);
}

__device__ __forceinline__ double lane_read(double v, int l) {
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), l);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), l);
    return __hiloint2double(hi, lo);
}

// the strip's emission rows (profile rows rowOf(0..63), 0 past the profile) into LDS
template <class F>
__device__ __forceinline__ void stage_emissions(double *sE, const GradArgs &a, int M, int lane, F rowOf) {
    for (int idx = lane; idx < 64 * kNCode; idx += 64) {
        const int rr = idx / kNCode, c = idx - rr * kNCode;
        const int row = rowOf(rr);
        sE[idx] = (row >= 0 && row <= M) ? (double)a.emis[a.prof->emOff + (int64_t)row * kEmStride + c] : 0.0;
    }
}

struct RowParams {
    double a0, a1, a2, b0, d0, d1, d2, eps, enter;
};
__device__ __forceinline__ RowParams row_params(const GradArgs &a, int row) {
    const float *pr = a.params + a.prof->parOff + (int64_t)row * 11;
    return RowParams{pr[0], pr[1], pr[2], pr[3], pr[6], pr[7], pr[8], pr[9], pr[10]};
}

constexpr int kExpLds = 256;      // 64-column blocks whose exponents the count pass stages in LDS
constexpr int kColRing = 256;        // columns of seeds and codes a pass keeps in LDS (the lanes of a step span 64)
constexpr int kNoExp = -1000000;  // binary exponent given to zero
constexpr int kRaise = 64;        // a block whose largest value passes 2^kRaise gets a higher exponent
constexpr int kLower = 256;       // one whose values all stay below 2^-kLower gets a lower one (not below 0)

__device__ __forceinline__ int exp_of(double v) { return v > 0.0 ? ilogb(v) : kNoExp; }

__device__ __forceinline__ int wave_max(int v) {
    for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o));
    return v;
}

// Exponent of a strip's next 64-column block. When lane 0 enters the block,
// every lane still holds values of the previous block (stored at E); m is the
// largest binary exponent among them and among the block's incoming edge
// values, relative to E. Within 64 columns values grow by less than 2^160 (a
// codon step multiplies by at most the largest emission odds), so stored
// values stay below 2^230 and products of two below 2^460; values more than
// 2^1000 below a block's largest are negligible against Z and may flush to 0.
__device__ __forceinline__ int next_block_exp(int E, int m) {
    if (m > kRaise) return E + m;
    if (m < -kLower && E > 0) return max(0, E + m);
    return E;
}

// Forward: w0, Y, Z0 of strips stripLo..stripHi (stored when kStore), strip by
// strip down, from the edge row of strip stripLo-1. Lane k of a strip runs row
// base+k at column t-k in step t. Each 64-column block of the strip has one
// exponent (expF), chosen when lane 0 enters the block; a lane converts its
// three columns of history when it enters the block itself.
template <bool kStore>
__global__ __launch_bounds__(64) void grad_forward_kernel(GradArgs a) {
#pragma clang fp contract(off)
    __shared__ double sE[64 * kNCode];
    __shared__ double sBg[32];
    // prefixSeed and code(j-2) of column j at slot j mod kColRing: the lanes of a
    // step span 64 columns, and every 64 steps the next 64 are written
    __shared__ float sCol[kColRing];
    __shared__ uint8_t sCode[kColRing];
    const int p = a.pairBase + blockIdx.x, lane = threadIdx.x;
    const PairInfo PR = a.pairs[p];
    const int n = PR.n, M = a.prof->M, NB = (n + 63) >> 6;
    const double s1 = PR.s1, s2 = PR.s2, sC = PR.sC;
    const Rec *rec = a.rec + PR.recOff;
    if (lane < 32) sBg[lane] = lane < kNCode ? (double)a.bg[lane] : 0.0;
    const size_t SP = (size_t)(n + 63) * 64;  // stored values per strip
    double *w0 = a.w0 + a.cellOff[p], *Yf = a.Y + a.cellOff[p], *Z0 = a.Z0 + a.cellOff[p];
    int *expF = a.expF + a.expOff[p];
    for (int strip = a.stripLo; strip <= a.stripHi; strip++, w0 += SP, Yf += SP, Z0 += SP) {
        const int base = 64 * strip, i = base + lane;
        // the edge row this strip reads (strip-1's) and the one it writes
        const int64_t inAt = a.bndOff[p] + (int64_t)(strip > 0 ? strip - 1 : 0) * n;
        const int64_t outAt = a.bndOff[p] + (int64_t)strip * n;
        const double *hA = a.fwdA + inAt, *hB = a.fwdB + inAt;
        const int *hE = a.fwdE + inAt;
        double *oA = a.fwdA + outAt, *oB = a.fwdB + outAt;
        int *oE = a.fwdE + outAt;
        const bool rowOk = i <= M;
        const RowParams P = row_params(a, rowOk ? i : 0);
        const double a1s1 = P.a1 * s1, a2s2 = P.a2 * s2, d1s2 = P.d1 * s2, d2s1 = P.d2 * s1, enterC = P.enter * sC;
        __syncthreads();  // the previous strip is done with sE and the column ring
        stage_emissions(sE, a, M, lane, [&](int rr) { return base + rr; });
        __syncthreads();
        const double *Er = sE + lane * kNCode;
        // row base-1 (the previous strip's last row): its contribution to row
        // base at column j, its Y, and their exponent, in chunks of 64 columns
        const bool edge = base > 0;
        auto chunkA = [&](int c) { return edge ? hA[min(64 * c + lane, n - 1)] : 0.0; };
        auto chunkB = [&](int c) { return edge ? hB[min(64 * c + lane, n - 1)] : 0.0; };
        auto chunkE = [&](int c) { return edge ? hE[min(64 * c + lane, n - 1)] : 0; };
        double cA = chunkA(0), cB = chunkB(0), nA = chunkA(1), nB = chunkB(1);
        int cE = chunkE(0), nE = chunkE(1);
        double w1v = 0.0, w2v = 0.0, w3v = 0.0, z1 = 0.0, z2 = 0.0, z3 = 0.0;
        double lastC = 0.0, lastY = 0.0;
        int Ecur = 0, Eprev = 0, dConv = 0;  // exponents of the newest block and the one before
        auto decide = [&](int c) {
            const double mx = fmax(fmax(fmax(w1v, w2v), fmax(w3v, z1)), fmax(fmax(z2, z3), fmax(lastC, lastY)));
            int m = exp_of(mx);
            if (edge) m = max(m, exp_of(fmax(cA, cB)) + (cE - Ecur));
            m = wave_max(m);
            Eprev = Ecur;
            Ecur = next_block_exp(Ecur, m);
            dConv = Eprev - Ecur;
            if (edge) {
                cA = ldexp(cA, cE - Ecur);
                cB = ldexp(cB, cE - Ecur);
            }
            if (lane == 0) expF[(int64_t)strip * NB + c] = Ecur;
        };
        auto step = [&](int t) {
            if (dConv != 0 && lane == (t & 63)) {  // this lane's first column of the newest block
                w1v = ldexp(w1v, dConv);
                w2v = ldexp(w2v, dConv);
                w3v = ldexp(w3v, dConv);
                z1 = ldexp(z1, dConv);
                z2 = ldexp(z2, dConv);
                z3 = ldexp(z3, dConv);
            }
            const double inC = lane_from_above(lastC), inY = lane_from_above(lastY);
            const double eA = lane_read(cA, t & 63), eB = lane_read(cB, t & 63);
            const int j = t - lane;
            const int El = (t & 63) >= lane ? Ecur : Eprev;  // the exponent of column j's block
            double oC = lastC, oY = lastY, sw = 0.0, sy = 0.0, sz = 0.0;
            if (rowOk && j >= 0 && j < n) {
                const double W0in = lane == 0 ? eA : inC, Yup = lane == 0 ? eB : inY;
                const int code = sCode[j & (kColRing - 1)];
                const double e = Er[code], g = sBg[code];
                const double X = enterC * e * w3v;
                const double z0 = g * sC * (P.a0 * w3v + P.b0 * z3);
                const double w = W0in + (z0 + a1s1 * w1v + a2s2 * w2v + ldexp((double)sCol[j & (kColRing - 1)], -El));
                const double y = P.d0 * w + P.eps * Yup;
                oC = X + y + d1s2 * w2v + d2s1 * w1v;  // row i+1's inflow at column j
                oY = y;
                sw = w;
                sy = y;
                sz = z0;
                if (lane == 63) {
                    oA[j] = oC;
                    oB[j] = oY;
                    oE[j] = El;
                }
                w3v = w2v;
                w2v = w1v;
                w1v = w;
                z3 = z2;
                z2 = z1;
                z1 = z0;
            }
            lastC = oC;
            lastY = oY;
            if constexpr (kStore) {
                const size_t at = (size_t)t * 64 + lane;  // one contiguous run per step
                w0[at] = sw;
                Yf[at] = sy;
                Z0[at] = sz;
            }
        };
        // at step 64c: the edge chunk of block c, then block c's exponent (no
        // block starts at or past column n: the lanes stay at the last one)
        Rec cur = rec[min(lane, n - 1)];  // the records of the next block to enter the ring
        auto boundary = [&](int t) {
            {
                const int jw = t + lane;  // block t/64's columns (past n: unused slots)
                sCol[jw & (kColRing - 1)] = cur.ps;
                sCode[jw & (kColRing - 1)] = (uint8_t)((cur.codes >> 8) & 255);  // code(j-2)
                cur = rec[min(t + 64 + lane, n - 1)];
            }
            if (t > 0) {
                cA = nA;
                cB = nB;
                cE = nE;
                nA = chunkA((t >> 6) + 1);
                nB = chunkB((t >> 6) + 1);
                nE = chunkE((t >> 6) + 1);
            }
            if (t < n) {
                decide(t >> 6);
            } else {
                Eprev = Ecur;
                dConv = 0;
            }
            __syncthreads();  // the ring's new block is visible to every lane
        };
        // one outer iteration per 64 steps: the inner loop issues no loads, so
        // its stores are never waited for
        const int T = n + 63;
        for (int t0 = 0; t0 < T; t0 += 64) {
            boundary(t0);
            const int tEnd = min(t0 + 64, T);
            int t = t0;
            for (; t + 4 <= tEnd; t += 4) {
                step(t);
                step(t + 1);
                step(t + 2);
                step(t + 3);
            }
            for (; t < tEnd; t++) step(t);
        }
        __threadfence();  // the next strip reads this strip's hand-over values
    }
}

// Backward: W1, Yb, Zb of strips stripHi..stripLo (stored when kStore), strip
// by strip up (strip s = rows 64s..64s+63, as in the forward), from the edge
// row of strip stripHi+1; lane k of a strip runs row 64s+63-k at column
// n-1-t+k in step t. Blocks of 64 columns from the right (expB) as in the
// forward. The column sums of W1 run along the lanes with Yb and leave each
// strip in its edge row (bwdC), strip 0's as sum_i W1(i, j).
template <bool kStore>
__global__ __launch_bounds__(64) void grad_backward_kernel(GradArgs a) {
#pragma clang fp contract(off)
    __shared__ double sE[64 * kNCode];
    __shared__ double sBg[32];
    __shared__ float sCol[kColRing];  // suffixSeed and code(j+1) of column j at j mod kColRing
    __shared__ uint8_t sCode[kColRing];
    const int p = a.pairBase + blockIdx.x, lane = threadIdx.x;
    const PairInfo PR = a.pairs[p];
    const int n = PR.n, M = a.prof->M, NB = (n + 63) >> 6;
    const double s1 = PR.s1, s2 = PR.s2, sC = PR.sC;
    const Rec *rec = a.rec + PR.recOff;
    if (lane < 32) sBg[lane] = lane < kNCode ? (double)a.bg[lane] : 0.0;
    const int S = (M + 64) / 64;
    const size_t SP = (size_t)(n + 63) * 64;  // stored values per strip
    int *expB = a.expB + a.expOff[p];
    // strip s = rows 64s..64s+63 from the bottom strip up; lane 0 runs the strip's last row
    for (int strip = a.stripHi; strip >= a.stripLo; strip--) {
        const int top = 64 * strip + 63;
        const size_t at0 = a.cellOff[p] + (size_t)(strip - a.stripLo) * SP;
        double *W1 = a.W1 + at0, *Yb = a.Yb + at0, *Zb = a.Zb + at0;
        // the edge row this strip reads (strip+1's) and the one it writes
        const int64_t inAt = a.bndOff[p] + (int64_t)(strip + 1 < S ? strip + 1 : strip) * n;
        const int64_t outAt = a.bndOff[p] + (int64_t)strip * n;
        const double *hA = a.bwdA + inAt, *hB = a.bwdB + inAt, *hC = a.bwdC + inAt;
        const int *hE = a.bwdE + inAt;
        double *oA = a.bwdA + outAt, *oB = a.bwdB + outAt, *oC = a.bwdC + outAt;
        int *oE = a.bwdE + outAt;
        const int i = top - lane;
        const bool rowOk = i <= M;
        const RowParams P = row_params(a, rowOk ? i : 0);
        const double a0sC = P.a0 * sC, a1s1 = P.a1 * s1, a2s2 = P.a2 * s2, b0sC = P.b0 * sC;
        const double d1s2 = P.d1 * s2, d2s1 = P.d2 * s1, enterC = P.enter * sC;
        __syncthreads();
        stage_emissions(sE, a, M, lane, [&](int rr) { return top - rr; });
        __syncthreads();
        const double *Er = sE + lane * kNCode;
        // row top+1 (the strip below's first row): W1, Yb, the column sum and
        // their exponent at column j = n-1-t for step t, in chunks of 64 columns
        const bool edge = strip < S - 1;
        auto chunkA = [&](int c) { return edge ? hA[max(n - 1 - 64 * c - lane, 0)] : 0.0; };
        auto chunkB = [&](int c) { return edge ? hB[max(n - 1 - 64 * c - lane, 0)] : 0.0; };
        auto chunkC = [&](int c) { return edge ? hC[max(n - 1 - 64 * c - lane, 0)] : 0.0; };
        auto chunkE = [&](int c) { return edge ? hE[max(n - 1 - 64 * c - lane, 0)] : 0; };
        double cA = chunkA(0), cB = chunkB(0), cC = chunkC(0), nA = chunkA(1), nB = chunkB(1), nC = chunkC(1);
        int cE = chunkE(0), nE = chunkE(1);
        double r1 = 0.0, r2 = 0.0, r3 = 0.0;  // W1(i+1, j+1..j+3), received earlier
        double c1 = 0.0, c2 = 0.0;            // W1(i, j+1), W1(i, j+2)
        double zb1 = 0.0, zb2 = 0.0, zb3 = 0.0;  // Zb(i, j+1..j+3)
        double lastW = 0.0, lastY = 0.0, lastS = 0.0;
        int Ecur = 0, Eprev = 0, dConv = 0;
        auto decide = [&](int c) {
            const double mx = fmax(fmax(fmax(fmax(r1, r2), fmax(r3, c1)), fmax(fmax(c2, zb1), fmax(zb2, zb3))),
                                   fmax(fmax(lastW, lastY), lastS));
            int m = exp_of(mx);
            if (edge) m = max(m, exp_of(fmax(fmax(cA, cB), cC)) + (cE - Ecur));
            m = wave_max(m);
            Eprev = Ecur;
            Ecur = next_block_exp(Ecur, m);
            dConv = Eprev - Ecur;
            if (edge) {
                cA = ldexp(cA, cE - Ecur);
                cB = ldexp(cB, cE - Ecur);
                cC = ldexp(cC, cE - Ecur);
            }
            if (lane == 0) expB[(int64_t)strip * NB + c] = Ecur;
        };
        auto step = [&](int t) {
            if (dConv != 0 && lane == (t & 63)) {  // this lane's first column of the newest block
                r1 = ldexp(r1, dConv);
                r2 = ldexp(r2, dConv);
                r3 = ldexp(r3, dConv);
                c1 = ldexp(c1, dConv);
                c2 = ldexp(c2, dConv);
                zb1 = ldexp(zb1, dConv);
                zb2 = ldexp(zb2, dConv);
                zb3 = ldexp(zb3, dConv);
            }
            const double inW = lane_from_above(lastW), inY = lane_from_above(lastY), inS = lane_from_above(lastS);
            const double eA = lane_read(cA, t & 63), eB = lane_read(cB, t & 63), eC = lane_read(cC, t & 63);
            const int j = n - 1 - t + lane;
            const int El = (t & 63) >= lane ? Ecur : Eprev;
            double oW = lastW, oY = lastY, oS = lastS, sw = 0.0, sy = 0.0, sz = 0.0;
            if (rowOk && j >= 0 && j < n) {
                const double Wdn = lane == 0 ? eA : inW, Ynext = lane == 0 ? eB : inY, Sdn = lane == 0 ? eC : inS;
                const int code = sCode[j & (kColRing - 1)];
                const double e = Er[code], g = sBg[code];
                double w = r3 * e * enterC + Ynext * P.d0;
                w = w + r2 * d1s2;
                w = w + r1 * d2s1;
                w = w + zb3 * g * a0sC;
                w = w + c2 * a2s2;
                w = w + c1 * a1s1;
                w = w + ldexp((double)sCol[j & (kColRing - 1)], -El);
                const double y = __fma_rn(P.eps, Ynext, w);
                const double zb = __fma_rn(b0sC, zb3 * g, w);
                const double sum = Sdn + w;
                sw = w;
                sy = y;
                sz = zb;
                if (lane == 63) {
                    oA[j] = w;
                    oB[j] = y;
                    oC[j] = sum;
                    oE[j] = El;
                }
                r3 = r2;
                r2 = r1;
                r1 = Wdn;
                c2 = c1;
                c1 = w;
                zb3 = zb2;
                zb2 = zb1;
                zb1 = zb;
                oW = w;
                oY = y;
                oS = sum;
            }
            lastW = oW;
            lastY = oY;
            lastS = oS;
            if constexpr (kStore) {
                const size_t at = (size_t)t * 64 + lane;  // one contiguous run per step
                W1[at] = sw;
                Yb[at] = sy;
                Zb[at] = sz;
            }
        };
        Rec cur = rec[max(n - 1 - lane, 0)];  // the records of the next block to enter the ring
        auto boundary = [&](int t) {
            {
                const int jw = n - 1 - t - lane;  // block t/64's columns, right to left (below 0: unused)
                sCol[jw & (kColRing - 1)] = cur.ss;
                sCode[jw & (kColRing - 1)] = (uint8_t)(cur.codes & 255);  // code(j+1)
                cur = rec[max(n - 1 - t - 64 - lane, 0)];
            }
            if (t > 0) {
                cA = nA;
                cB = nB;
                cC = nC;
                cE = nE;
                nA = chunkA((t >> 6) + 1);
                nB = chunkB((t >> 6) + 1);
                nC = chunkC((t >> 6) + 1);
                nE = chunkE((t >> 6) + 1);
            }
            if (t < n) {
                decide(t >> 6);
            } else {
                Eprev = Ecur;
                dConv = 0;
            }
            __syncthreads();  // the ring's new block is visible to every lane
        };
        // one outer iteration per 64 steps: the inner loop issues no loads, so
        // its stores are never waited for
        const int T = n + 63;
        for (int t0 = 0; t0 < T; t0 += 64) {
            boundary(t0);
            const int tEnd = min(t0 + 64, T);
            int t = t0;
            for (; t + 4 <= tEnd; t += 4) {
                step(t);
                step(t + 1);
                step(t + 2);
                step(t + 3);
            }
            for (; t < tEnd; t++) step(t);
        }
        __threadfence();
    }
}

__device__ __forceinline__ double wave_sum(double v) {  // the same total in every lane
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

// Per pair (one wavefront): Z = sum_j prefixSeed(j) colSum(j) as a mantissa
// and a binary exponent (zInfo); lane l sums columns l, l+64, ... at its own
// exponent, then the lanes meet at the largest.
__global__ __launch_bounds__(64) void grad_z_kernel(GradArgs a) {
    const int p = a.pairBase + blockIdx.x, lane = threadIdx.x;
    const PairInfo PR = a.pairs[p];
    const int n = PR.n;
    const double *cs = a.bwdC + a.bndOff[p];  // strip 0's edge: the column sums over all rows
    const int *eb = a.expB + a.expOff[p];  // strip 0 holds row 0
    const Rec *rec = a.rec + PR.recOff;
    double zm = 0.0;
    int ze = kNoExp;
    for (int j = lane; j < n; j += 64) {
        const double v = (double)rec[j].ps * cs[j];
        if (!(v > 0.0)) continue;
        const int e = eb[(n - 1 - j) >> 6];
        if (e > ze) {
            zm = ldexp(zm, ze - e) + v;
            ze = e;
        } else {
            zm += ldexp(v, e - ze);
        }
    }
    const int zmax = wave_max(ze);
    const double tot = wave_sum(ldexp(zm, ze - zmax));
    if (lane == 0) {
        double *zi = a.zInfo + 2 * (size_t)a.rawIdx[p];
        zi[0] = tot;
        zi[1] = (double)zmax;
    }
}

// Per (pair, strip s of stripLo..stripHi) one wavefront: lane k takes row
// 64s+k and walks the forward's steps t = 0..n+62 (column j = t-k), so each
// read is one contiguous run of the diagonal-major matrices: the forward's
// cell is at step t, and the backward's cell of row 64s+k or 64s+k+1 at column
// j+d is at step n+62-t-d or n+61-t-d, the same in every lane (the strip's
// edge lanes read the neighbouring strips' edge rows). Cells outside the
// sequence (the first and last 63 steps of a lane) contribute zero. Flows of
// the row's parameters in registers, its match flows by codon in LDS; each
// flow is its stored product times 2^(forward exponent + backward exponent -
// Z's exponent).
template <bool kLdsExp>
__global__ __launch_bounds__(64) void grad_counts_kernel(GradArgs a) {
#pragma clang fp contract(off)
    __shared__ double bins[kGradCodons][64];
    // exponent rows of forward strips s-1, s and backward strips s, s+1: in LDS
    // when every sequence of the launch has at most 64 kExpLds letters (kLdsExp),
    // read in place otherwise
    __shared__ int sExp[4][kLdsExp ? kExpLds : 1];
    const int p = a.pairBase + blockIdx.x, s = a.stripLo + blockIdx.y, lane = threadIdx.x;
    const PairInfo PR = a.pairs[p];
    const int n = PR.n, M = a.prof->M, NB = (n + 63) >> 6, S = (M + 64) / 64;
    const int i = 64 * s + lane;
    for (int c = 0; c < kGradCodons; c++) bins[c][lane] = 0.0;
    const int *expRow[4];
    {
        const int *eF = a.expF + a.expOff[p], *eB = a.expB + a.expOff[p];
        expRow[0] = eF + (int64_t)(s > 0 ? s - 1 : s) * NB;  // read only for rows below 64s
        expRow[1] = eF + (int64_t)s * NB;
        expRow[2] = eB + (int64_t)s * NB;
        expRow[3] = eB + (int64_t)(s + 1 < S ? s + 1 : s) * NB;  // read only for rows above 64s+63
        if constexpr (kLdsExp) {
            for (int r = 0; r < 4; r++)
                for (int c = lane; c < NB; c += 64) sExp[r][c] = expRow[r][c];
        }
    }
    __syncthreads();
    const bool rowOk = i <= M, hasBelow = i < M, hasAbove = i > 0;
    const double s1 = PR.s1, s2 = PR.s2, sC = PR.sC;
    const RowParams P = row_params(a, rowOk ? i : 0);
    const double a0sC = P.a0 * sC, a1s1 = P.a1 * s1, a2s2 = P.a2 * s2, b0sC = P.b0 * sC;
    const double d1s2 = P.d1 * s2, d2s1 = P.d2 * s1, enterC = P.enter * sC;
    const float *Er = a.emis + a.prof->emOff + (int64_t)(rowOk ? i : 0) * kEmStride;
    const Rec *rec = a.rec + PR.recOff;
    const uint8_t *codon = a.codon + PR.seqOff;
    const int64_t SP = (int64_t)(n + 63) * 64;
    const size_t off = a.cellOff[p] + (size_t)(s - a.stripLo) * SP;
    const double *w0 = a.w0 + off, *Z0 = a.Z0 + off, *Yf = a.Y + off;
    const double *W1 = a.W1 + off, *Yb = a.Yb + off, *Zb = a.Zb + off;
    // offsets of column j = 0 (the forward's grow by 64 per column, the
    // backward's fall by 64)
    const int64_t fHere = (int64_t)lane * 65;
    const int64_t bHere = (int64_t)(n - 1) * 64 + 63 * 65 - (int64_t)lane * 65;
    // the neighbour rows: Y of row i-1 and W1, Yb of row i+1 at column c are
    // yU[c yUs], wD[c wDs], ybD[c wDs]; inside the strip the matrices, at its
    // edges the neighbour strips' edge rows (one value per column); a missing
    // neighbour row reads row i and is zeroed
    const double *yU = Yf + fHere - 65, *wD = W1 + bHere - 65, *ybD = Yb + bHere - 65;
    int64_t yUs = 64, wDs = -64;
    if (!hasAbove) {
        yU = Yf + fHere;
    } else if (lane == 0) {
        yU = a.fwdB + a.bndOff[p] + (int64_t)(s - 1) * n;
        yUs = 1;
    }
    if (!hasBelow) {
        wD = W1 + bHere;
        ybD = Yb + bHere;
    } else if (lane == 63) {
        wD = a.bwdA + a.bndOff[p] + (int64_t)(s + 1) * n;
        ybD = a.bwdB + a.bndOff[p] + (int64_t)(s + 1) * n;
        wDs = 1;
    }
    const int *eFh = kLdsExp ? sExp[1] : expRow[1], *eFu = kLdsExp ? (lane > 0 ? sExp[1] : sExp[0])
                                                                   : (lane > 0 ? expRow[1] : expRow[0]);
    const int *eBh = kLdsExp ? sExp[2] : expRow[2], *eBd = kLdsExp ? (lane < 63 ? sExp[2] : sExp[3])
                                                                   : (lane < 63 ? expRow[2] : expRow[3]);
    const double up = hasAbove ? 1.0 : 0.0, dn = hasBelow ? 1.0 : 0.0;
    const int ZE = (int)a.zInfo[2 * (size_t)a.rawIdx[p] + 1];
    double cnt[kGradParams] = {0.0};
    if (rowOk) {
        // step t: lane k at column t - k, so every lane's reads of one matrix
        // form one contiguous run (cells outside the sequence read a clamped
        // column and count zero)
        auto step = [&](int t) {
            const int jr = t - lane;
            const double v = (jr >= 0 && jr < n) ? 1.0 : 0.0;
            const int j = min(max(jr, 0), n - 1);
            const int64_t fb = (int64_t)j * 64;
            const int r1 = min(j + 1, n - 1), r2 = min(j + 2, n - 1), r3 = min(j + 3, n - 1);
            const double m1 = j + 1 < n ? v : 0.0, m2 = j + 2 < n ? v : 0.0, m3 = j + 3 < n ? v : 0.0;
            const int code = (int)(rec[j].codes & 255);  // code(j+1)
            const int cod = codon[r1];
            const double w = v * w0[fHere + fb], z0 = Z0[fHere + fb], yup = up * yU[(int64_t)j * yUs];
            const double wd1 = dn * m1 * wD[(int64_t)r1 * wDs];
            const double wd2 = dn * m2 * wD[(int64_t)r2 * wDs];
            const double wd3 = dn * m3 * wD[(int64_t)r3 * wDs];
            const double ybd = dn * ybD[(int64_t)j * wDs];
            const double wh1 = m1 * W1[bHere - (int64_t)r1 * 64], wh2 = m2 * W1[bHere - (int64_t)r2 * 64];
            const double zb3 = m3 * Zb[bHere - (int64_t)r3 * 64];
            const double e = Er[code], g = a.bg[code];
            const int f = eFh[j >> 6] - ZE, fu = eFu[j >> 6] - ZE;
            auto bx = [&](int x) { return (n - 1 - x) >> 6; };
            const double fm = ldexp(w * enterC * e * wd3, f + eBd[bx(r3)]);
            cnt[10] += fm;
            bins[j + 1 < n ? cod : kGradCodons - 1][lane] += fm;
            cnt[7] += ldexp(w * d1s2 * wd2, f + eBd[bx(r2)]);
            cnt[8] += ldexp(w * d2s1 * wd1, f + eBd[bx(r1)]);
            const int e0 = eBd[bx(j)];
            cnt[6] += ldexp(w * P.d0 * ybd, f + e0);
            cnt[9] += ldexp(v * yup * P.eps * ybd, fu + e0);
            cnt[1] += ldexp(w * a1s1 * wh1, f + eBh[bx(r1)]);
            cnt[2] += ldexp(w * a2s2 * wh2, f + eBh[bx(r2)]);
            const int e3 = eBh[bx(r3)];
            cnt[0] += ldexp(w * a0sC * g * zb3, f + e3);
            cnt[3] += ldexp(v * z0 * b0sC * g * zb3, f + e3);
        };
        const int T = n + 63;
        int t = 0;
        for (; t + 4 <= T; t += 4) {
            step(t);
            step(t + 1);
            step(t + 2);
            step(t + 3);
        }
        for (; t < T; t++) step(t);
        double *out = a.raw + ((size_t)a.rawIdx[p] * (M + 1) + i) * kGradRaw;
        for (int k = 0; k < kGradParams; k++) out[k] = cnt[k];
        for (int c = 0; c < kGradCodons; c++) out[kGradParams + c] = bins[c][lane];
    }
}

// Per pair: ln Z and the factor weight / (Z's mantissa).
__global__ void grad_zfactor_kernel(GradReduceArgs a, double *factor) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= a.npairs) return;
    const double zm = a.zInfo[2 * (size_t)p], ze = a.zInfo[2 * (size_t)p + 1];
    a.lnZ[p] = log(zm) + ze * 0.693147180559945309417;  // ln 2
    factor[p] = a.weight[p] / zm;
}

// counts[r][c] = sum over pairs, in order, of the pair's flow / Z * weight.
__global__ void grad_reduce_kernel(GradReduceArgs a, const double *factor) {
#pragma clang fp contract(off)
    const int id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= (a.M + 1) * kGradRaw) return;
    double acc = 0.0;
    for (int p = 0; p < a.npairs; p++) acc += a.raw[(size_t)p * (a.M + 1) * kGradRaw + id] * factor[p];
    a.counts[id] = acc;
}

}  // namespace


hipError_t launch_grad_forward(const GradArgs &a, int npairs, int maxN, hipStream_t s) {
    if (npairs <= 0) return hipSuccess;
    (void)maxN;
    if (a.store)
        hipLaunchKernelGGL(grad_forward_kernel<true>, dim3(npairs), dim3(64), 0, s, a);
    else
        hipLaunchKernelGGL(grad_forward_kernel<false>, dim3(npairs), dim3(64), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_grad_backward(const GradArgs &a, int npairs, int maxN, hipStream_t s) {
    if (npairs <= 0) return hipSuccess;
    (void)maxN;
    if (a.store)
        hipLaunchKernelGGL(grad_backward_kernel<true>, dim3(npairs), dim3(64), 0, s, a);
    else
        hipLaunchKernelGGL(grad_backward_kernel<false>, dim3(npairs), dim3(64), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_grad_z(const GradArgs &a, int npairs, hipStream_t s) {
    if (npairs <= 0) return hipSuccess;
    hipLaunchKernelGGL(grad_z_kernel, dim3(npairs), dim3(64), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_grad_counts(const GradArgs &a, int npairs, int M, int maxN, hipStream_t s) {
    if (npairs <= 0) return hipSuccess;
    (void)M;
    const dim3 grid(npairs, a.stripHi - a.stripLo + 1);
    if (maxN <= 64 * kExpLds)
        hipLaunchKernelGGL(grad_counts_kernel<true>, grid, dim3(64), 0, s, a);
    else
        hipLaunchKernelGGL(grad_counts_kernel<false>, grid, dim3(64), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_grad_reduce(const GradReduceArgs &a, double *factor, hipStream_t s) {
    if (a.npairs > 0) hipLaunchKernelGGL(grad_zfactor_kernel, dim3((a.npairs + 63) / 64), dim3(64), 0, s, a, factor);
    const int cells = (a.M + 1) * kGradRaw;
    hipLaunchKernelGGL(grad_reduce_kernel, dim3((cells + 255) / 256), dim3(256), 0, s, a, (const double *)factor);
    return hipGetLastError();
}

}  // namespace dm
