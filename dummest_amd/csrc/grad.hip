// Gradient of the summed forward log-likelihood on gfx950: for each (profile,
// DNA sequence) pair the total Z of all local alignments,
// Z = sum_{i,j} prefixSeed(j) W1(i, j) = sum_{i,j} w0(i, j) suffixSeed(j),
// and the expected number of uses of every parameter of the profile,
// d ln Z / d ln theta = (1/Z) sum over the uses of theta of
// (forward value into the use) theta (backward value out of it).
//
// The recurrences are the reference's (forward dummest-dp-forward.hh:9-250,
// backward dummest-dp-backward.hh:444-615) in FP64 without row rescaling. Per
// use (row i, column j):
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
// previous strip's last row from a hand-over buffer in 64-column chunks. The
// pair's per-column codes and seeds sit in LDS with the strip's emission rows.
// Every pass stores its matrices in full (row-major, pitch n); the count pass
// gives each row a lane walking its row, with the codon-indexed emission sums
// in LDS.
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

__global__ __launch_bounds__(64) void grad_forward_kernel(GradArgs a) {
#pragma clang fp contract(off)
    __shared__ double sE[64 * kNCode];
    __shared__ double sBg[32];
    extern __shared__ float sCol[];  // prefixSeed per column, then code(j-2) per column (bytes)
    const int p = a.pairBase + blockIdx.x, lane = threadIdx.x;
    const PairInfo PR = a.pairs[p];
    const int n = PR.n, M = a.prof->M;
    const double s1 = PR.s1, s2 = PR.s2, sC = PR.sC;
    uint8_t *sCode = reinterpret_cast<uint8_t *>(sCol + n);
    const Rec *rec = a.rec + PR.recOff;
    for (int j = lane; j < n; j += 64) {
        const Rec r = rec[j];
        sCol[j] = r.ps;
        sCode[j] = (uint8_t)((r.codes >> 8) & 255);  // code(j-2)
    }
    if (lane < 32) sBg[lane] = lane < kNCode ? (double)a.bg[lane] : 0.0;
    double *w0 = a.w0 + a.cellOff[p], *Yf = a.Y + a.cellOff[p], *Z0 = a.Z0 + a.cellOff[p];
    double *hA = a.bndA + a.bndOff[p], *hB = a.bndB + a.bndOff[p];
    for (int base = 0; base <= M; base += 64) {
        const int i = base + lane;
        const bool rowOk = i <= M;
        const RowParams P = row_params(a, rowOk ? i : 0);
        const double a1s1 = P.a1 * s1, a2s2 = P.a2 * s2, d1s2 = P.d1 * s2, d2s1 = P.d2 * s1, enterC = P.enter * sC;
        __syncthreads();  // columns staged; the previous strip is done with sE
        stage_emissions(sE, a, M, lane, [&](int rr) { return base + rr; });
        __syncthreads();
        const double *Er = sE + lane * kNCode;
        // row base-1 (the previous strip's last row): its contribution to row
        // base at column j and its Y, for step t = j, in chunks of 64 columns
        const bool edge = base > 0;
        auto chunkA = [&](int c) { return edge ? hA[min(64 * c + lane, n - 1)] : 0.0; };
        auto chunkB = [&](int c) { return edge ? hB[min(64 * c + lane, n - 1)] : 0.0; };
        double cA = chunkA(0), cB = chunkB(0), nA = chunkA(1), nB = chunkB(1);
        double w1v = 0.0, w2v = 0.0, w3v = 0.0, z1 = 0.0, z2 = 0.0, z3 = 0.0;
        double lastC = 0.0, lastY = 0.0;
        auto step = [&](int t) {
            const double inC = lane_from_above(lastC), inY = lane_from_above(lastY);
            const double eA = lane_read(cA, t & 63), eB = lane_read(cB, t & 63);
            const int j = t - lane;
            double oC = lastC, oY = lastY;
            if (rowOk && j >= 0 && j < n) {
                const double W0in = lane == 0 ? eA : inC, Yup = lane == 0 ? eB : inY;
                const int code = sCode[j];
                const double e = Er[code], g = sBg[code];
                const double X = enterC * e * w3v;
                const double z0 = g * sC * (P.a0 * w3v + P.b0 * z3);
                const double w = W0in + (z0 + a1s1 * w1v + a2s2 * w2v + (double)sCol[j]);
                const double y = P.d0 * w + P.eps * Yup;
                oC = X + y + d1s2 * w2v + d2s1 * w1v;  // row i+1's inflow at column j
                oY = y;
                const size_t at = (size_t)i * n + j;
                w0[at] = w;
                Yf[at] = y;
                Z0[at] = z0;
                if (lane == 63) {
                    hA[j] = oC;
                    hB[j] = oY;
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
        };
        const int T = n + 63;
        int t = 0;
        for (; t + 4 <= T; t += 4) {
            if ((t & 63) == 0 && t > 0) {
                cA = nA;
                cB = nB;
                nA = chunkA((t >> 6) + 1);
                nB = chunkB((t >> 6) + 1);
            }
            step(t);
            step(t + 1);
            step(t + 2);
            step(t + 3);
        }
        if ((t & 63) == 0 && t > 0) {
            cA = nA;
            cB = nB;
        }
        for (; t < T; t++) step(t);
        __threadfence();  // the next strip reads this strip's hand-over values
    }
}

__global__ __launch_bounds__(64) void grad_backward_kernel(GradArgs a) {
#pragma clang fp contract(off)
    __shared__ double sE[64 * kNCode];
    __shared__ double sBg[32];
    extern __shared__ float sCol[];  // suffixSeed per column, then code(j+1) per column
    const int p = a.pairBase + blockIdx.x, lane = threadIdx.x;
    const PairInfo PR = a.pairs[p];
    const int n = PR.n, M = a.prof->M;
    const double s1 = PR.s1, s2 = PR.s2, sC = PR.sC;
    uint8_t *sCode = reinterpret_cast<uint8_t *>(sCol + n);
    const Rec *rec = a.rec + PR.recOff;
    for (int j = lane; j < n; j += 64) {
        const Rec r = rec[j];
        sCol[j] = r.ss;
        sCode[j] = (uint8_t)(r.codes & 255);  // code(j+1)
    }
    if (lane < 32) sBg[lane] = lane < kNCode ? (double)a.bg[lane] : 0.0;
    double *W1 = a.W1 + a.cellOff[p], *Yb = a.Yb + a.cellOff[p], *Zb = a.Zb + a.cellOff[p];
    double *hA = a.bndA + a.bndOff[p], *hB = a.bndB + a.bndOff[p];
    for (int top = M; top >= 0; top -= 64) {
        const int i = top - lane;
        const bool rowOk = i >= 0;
        const RowParams P = row_params(a, rowOk ? i : 0);
        const double a0sC = P.a0 * sC, a1s1 = P.a1 * s1, a2s2 = P.a2 * s2, b0sC = P.b0 * sC;
        const double d1s2 = P.d1 * s2, d2s1 = P.d2 * s1, enterC = P.enter * sC;
        __syncthreads();
        stage_emissions(sE, a, M, lane, [&](int rr) { return top - rr; });
        __syncthreads();
        const double *Er = sE + lane * kNCode;
        // row top+1 (the previous strip's last row): W1 and Yb at column
        // j = n-1-t for step t, in chunks of 64 columns
        const bool edge = top < M;
        auto chunkA = [&](int c) { return edge ? hA[max(n - 1 - 64 * c - lane, 0)] : 0.0; };
        auto chunkB = [&](int c) { return edge ? hB[max(n - 1 - 64 * c - lane, 0)] : 0.0; };
        double cA = chunkA(0), cB = chunkB(0), nA = chunkA(1), nB = chunkB(1);
        double r1 = 0.0, r2 = 0.0, r3 = 0.0;  // W1(i+1, j+1..j+3), received earlier
        double c1 = 0.0, c2 = 0.0;            // W1(i, j+1), W1(i, j+2)
        double zb1 = 0.0, zb2 = 0.0, zb3 = 0.0;  // Zb(i, j+1..j+3)
        double lastW = 0.0, lastY = 0.0;
        auto step = [&](int t) {
            const double inW = lane_from_above(lastW), inY = lane_from_above(lastY);
            const double eA = lane_read(cA, t & 63), eB = lane_read(cB, t & 63);
            const int j = n - 1 - t + lane;
            double oW = lastW, oY = lastY;
            if (rowOk && j >= 0 && j < n) {
                const double Wdn = lane == 0 ? eA : inW, Ynext = lane == 0 ? eB : inY;
                const int code = sCode[j];
                const double e = Er[code], g = sBg[code];
                double w = r3 * e * enterC + Ynext * P.d0;
                w = w + r2 * d1s2;
                w = w + r1 * d2s1;
                w = w + zb3 * g * a0sC;
                w = w + c2 * a2s2;
                w = w + c1 * a1s1;
                w = w + (double)sCol[j];
                const double y = __fma_rn(P.eps, Ynext, w);
                const double zb = __fma_rn(b0sC, zb3 * g, w);
                const size_t at = (size_t)i * n + j;
                W1[at] = w;
                Yb[at] = y;
                Zb[at] = zb;
                if (lane == 63) {
                    hA[j] = w;
                    hB[j] = y;
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
            }
            lastW = oW;
            lastY = oY;
        };
        const int T = n + 63;
        int t = 0;
        for (; t + 4 <= T; t += 4) {
            if ((t & 63) == 0 && t > 0) {
                cA = nA;
                cB = nB;
                nA = chunkA((t >> 6) + 1);
                nB = chunkB((t >> 6) + 1);
            }
            step(t);
            step(t + 1);
            step(t + 2);
            step(t + 3);
        }
        if ((t & 63) == 0 && t > 0) {
            cA = nA;
            cB = nB;
        }
        for (; t < T; t++) step(t);
        __threadfence();
    }
}

// Per row (one lane): the flows of the row's parameters and codons over all
// columns, and the row's share of Z (unnormalized).
__global__ __launch_bounds__(64) void grad_counts_kernel(GradArgs a) {
#pragma clang fp contract(off)
    __shared__ double emis[kGradCodons][64];
    const int p = a.pairBase + blockIdx.x, lane = threadIdx.x;
    const PairInfo PR = a.pairs[p];
    const int n = PR.n, M = a.prof->M;
    const int i = 64 * blockIdx.y + lane;
    for (int c = 0; c < kGradCodons; c++) emis[c][lane] = 0.0;
    if (i > M) return;
    const double s1 = PR.s1, s2 = PR.s2, sC = PR.sC;
    const RowParams P = row_params(a, i);
    const double a0sC = P.a0 * sC, a1s1 = P.a1 * s1, a2s2 = P.a2 * s2, b0sC = P.b0 * sC;
    const double d1s2 = P.d1 * s2, d2s1 = P.d2 * s1, enterC = P.enter * sC;
    const float *Er = a.emis + a.prof->emOff + (int64_t)i * kEmStride;
    const Rec *rec = a.rec + PR.recOff;
    const uint8_t *codon = a.codon + PR.seqOff;
    const bool hasBelow = i < M, hasAbove = i > 0;
    const size_t off = a.cellOff[p], here = (size_t)i * n, below = hasBelow ? (size_t)(i + 1) * n : here,
                 above = hasAbove ? (size_t)(i - 1) * n : here;
    const double *w0 = a.w0 + off + here, *Z0 = a.Z0 + off + here, *Yup = a.Y + off + above;
    const double *W1 = a.W1 + off + here, *W1d = a.W1 + off + below, *Ybd = a.Yb + off + below;
    const double *Zbh = a.Zb + off + here;
    double cnt[kGradParams] = {0.0};
    double zrow = 0.0;
    for (int j = 0; j < n; j++) {
        const Rec r = rec[j];
        const int code = (int)(r.codes & 255);  // code(j+1)
        const double w = w0[j];
        const double e = Er[code], g = a.bg[code];
        const double wd1 = (hasBelow && j + 1 < n) ? W1d[j + 1] : 0.0;
        const double wd2 = (hasBelow && j + 2 < n) ? W1d[j + 2] : 0.0;
        const double wd3 = (hasBelow && j + 3 < n) ? W1d[j + 3] : 0.0;
        const double ybd = hasBelow ? Ybd[j] : 0.0;
        const double wh1 = (j + 1 < n) ? W1[j + 1] : 0.0, wh2 = (j + 2 < n) ? W1[j + 2] : 0.0;
        const double zb3 = (j + 3 < n) ? Zbh[j + 3] : 0.0;
        const double yup = hasAbove ? Yup[j] : 0.0;
        const double fm = w * enterC * e * wd3;
        cnt[10] += fm;
        emis[(j + 1 < n) ? codon[j + 1] : kGradCodons - 1][lane] += fm;
        cnt[7] += w * d1s2 * wd2;
        cnt[8] += w * d2s1 * wd1;
        cnt[6] += w * P.d0 * ybd;
        cnt[9] += yup * P.eps * ybd;
        cnt[1] += w * a1s1 * wh1;
        cnt[2] += w * a2s2 * wh2;
        cnt[0] += w * a0sC * g * zb3;
        cnt[3] += Z0[j] * b0sC * g * zb3;
        zrow += (double)r.ps * W1[j];
    }
    double *out = a.raw + ((size_t)a.rawIdx[p] * (M + 1) + i) * kGradRaw;
    for (int k = 0; k < kGradParams; k++) out[k] = cnt[k];
    for (int c = 0; c < kGradCodons; c++) out[kGradParams + c] = emis[c][lane];
    out[kGradRaw - 1] = zrow;
}

// Per pair: Z (rows added in order), ln Z, and the factor weight / Z.
__global__ void grad_z_kernel(GradReduceArgs a, double *factor) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= a.npairs) return;
    double z = 0.0;
    for (int i = 0; i <= a.M; i++) z += a.raw[((size_t)p * (a.M + 1) + i) * kGradRaw + kGradRaw - 1];
    a.lnZ[p] = log(z);
    factor[p] = a.weight[p] / z;
}

// counts[r][c] = sum over pairs, in order, of the pair's flow / Z * weight.
__global__ void grad_reduce_kernel(GradReduceArgs a, const double *factor) {
#pragma clang fp contract(off)
    const int W = kGradRaw - 1;
    const int id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= (a.M + 1) * W) return;
    const int r = id / W, c = id - r * W;
    double acc = 0.0;
    for (int p = 0; p < a.npairs; p++) acc += a.raw[((size_t)p * (a.M + 1) + r) * kGradRaw + c] * factor[p];
    a.counts[id] = acc;
}

}  // namespace

static size_t col_lds(int maxN) { return (size_t)maxN * sizeof(float) + (size_t)maxN + 16; }

hipError_t launch_grad_forward(const GradArgs &a, int npairs, int maxN, hipStream_t s) {
    if (npairs <= 0) return hipSuccess;
    hipLaunchKernelGGL(grad_forward_kernel, dim3(npairs), dim3(64), col_lds(maxN), s, a);
    return hipGetLastError();
}

hipError_t launch_grad_backward(const GradArgs &a, int npairs, int maxN, hipStream_t s) {
    if (npairs <= 0) return hipSuccess;
    hipLaunchKernelGGL(grad_backward_kernel, dim3(npairs), dim3(64), col_lds(maxN), s, a);
    return hipGetLastError();
}

hipError_t launch_grad_counts(const GradArgs &a, int npairs, int M, hipStream_t s) {
    if (npairs <= 0) return hipSuccess;
    hipLaunchKernelGGL(grad_counts_kernel, dim3(npairs, (M + 1 + 63) / 64), dim3(64), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_grad_reduce(const GradReduceArgs &a, double *factor, hipStream_t s) {
    if (a.npairs > 0) hipLaunchKernelGGL(grad_z_kernel, dim3((a.npairs + 63) / 64), dim3(64), 0, s, a, factor);
    const int cells = (a.M + 1) * (kGradRaw - 1);
    hipLaunchKernelGGL(grad_reduce_kernel, dim3((cells + 255) / 256), dim3(256), 0, s, a, (const double *)factor);
    return hipGetLastError();
}

}  // namespace dm
