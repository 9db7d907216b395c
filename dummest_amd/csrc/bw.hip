// Baum-Welch E-step of the profile builder on gfx950 (dummest-build.cc:561-736):
// forward (X, Y, Z) and backward (Wb, Yb, Zb) over every ungapped sequence of
// an alignment against the current profile, then the expected event counts
// per profile row. FP64 throughout, no fused multiply-adds, every sum taken in
// the reference's order (j ascending within a row, sequences in order), so the
// counts equal the reference's except through the forward total v, which is
// summed by rows and strips here instead of cell by cell (a rounding-level
// change of the per-sequence factor weight / (v 2^-480)).
//
// Layout: one wavefront per sequence for forward and backward. The profile
// rows are taken 64 at a time (a strip), lane k holding row base + k and
// running one column behind lane k - 1, so the values a lane needs from the row
// above arrive from its neighbour by a DPP shift; the strip's first lane reads
// the previous strip's last row back from the matrices, which every pass
// stores in full because the count pass needs all six (row-major, pitch n+2).
// The count pass gives each row a lane that walks its row left to right, with
// the letter-indexed emission sums in LDS. One wave per sequence makes forward
// and backward latency-bound (the FP64 chains of a step, the LDS reads of its
// letter and likelihood ratio); an alignment's sequences run side by side.
#include <hip/hip_runtime.h>

#include <type_traits>

#include "bw.h"

namespace dm {
namespace {

// 2^-480: the reference keeps forward/backward values scaled down by this.
__device__ __forceinline__ double bw_scale() { return 0x1p-480; }

// lane k receives lane k-1's value, lane 0 receives 0 (DPP wave_shr:1)
__device__ __forceinline__ double from_lane_above(double v) {
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), 0x138, 0xf, 0xf, true);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), 0x138, 0xf, 0xf, true);
    return __hiloint2double(hi, lo);
}

// The sequence's letters are staged in LDS (dynamic shared memory, n+1
// bytes) with the strip's likelihood ratios beside them; each step reads its
// letter one step ahead. The strip's first lane takes the previous strip's last
// row in chunks of 64 columns: the whole wave loads a chunk (coalesced) one
// chunk ahead and the lane picks its value with readlane, so a strip waits on
// global memory once per 64 steps.
__device__ __forceinline__ double lane_value(double v, int l) {
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), l);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), l);
    return __hiloint2double(hi, lo);
}

__device__ __forceinline__ void stage_letters(uint8_t *sL, const uint8_t *seq, int n, int lane) {
    for (int j = lane; j <= n; j += 64) sL[j] = seq[j];
}

// The strip's 64 rows of likelihood ratios into LDS ([row][letter]); rowOf(rr)
// gives the profile row of strip row rr, or -1 past the profile. All loads are
// issued before the first LDS write.
template <class F>
__device__ __forceinline__ void stage_ratios(double *sS, const double *ratio, int A, int lane, F rowOf) {
    double v[21];
#pragma unroll
    for (int q = 0; q < 21; q++) {
        const int idx = lane + 64 * q;
        const int rr = idx / (A + 1), x = idx - rr * (A + 1);
        const int row = idx < 64 * (A + 1) ? rowOf(rr) : -1;
        v[q] = row >= 0 ? ratio[(size_t)row * (A + 1) + x] : 0.0;
    }
#pragma unroll
    for (int q = 0; q < 21; q++)
        if (lane + 64 * q < 64 * (A + 1)) sS[lane + 64 * q] = v[q];
}

__global__ __launch_bounds__(64) void bw_forward_kernel(BWArgs a) {
#pragma clang fp contract(off)
    __shared__ double sS[64 * 21];  // the strip's match likelihood ratios [row][letter]
    extern __shared__ uint8_t sL[];  // the sequence's letters (and the one after)
    const int s = a.seqBase + blockIdx.x, lane = threadIdx.x;
    const int n = a.len[s], L = a.L, A = a.A, P = n + 2;
    double *X = a.X + a.cellOff[s], *Y = a.Y + a.cellOff[s], *Z = a.Z + a.cellOff[s];
    auto at = [&](int ii, int jj) { return (size_t)ii * P + jj; };
    stage_letters(sL, a.letters + a.letterOff[s], n, lane);
    for (int j = lane; j <= n + 1; j += 64) {  // row 0
        if (j <= n) X[at(0, j)] = 0.0;
        if (j >= 1) Y[at(0, j)] = 0.0;
    }
    double vsum = 0.0;  // forward total, strip by strip
    for (int base = 1; base <= L + 1; base += 64) {
        const int i = base + lane;
        const bool rowOk = i <= L + 1;
        const int r = rowOk ? i - 1 : 0;  // profile row of this DP row's transitions and emissions
        const double ra = a.rates[4 * r], rb = a.rates[4 * r + 1], rd = a.rates[4 * r + 2], re = a.rates[4 * r + 3];
        __syncthreads();  // letters staged; the previous strip is done with sS
        stage_ratios(sS, a.ratio, A, lane, [&](int rr) { return base - 1 + rr <= L ? base - 1 + rr : -1; });
        __syncthreads();
        const double *Srow = sS + lane * (A + 1);
        if (rowOk) {
            X[at(i - 1, 0)] = 0.0;
            Z[at(i, 0)] = 0.0;
        }
        // previous strip's last row: step t needs X[base-1][t] and Y[base-1][t+1]
        const bool edge = base > 1;
        auto chunkX = [&](int c) { return edge ? X[at(base - 1, min(64 * c + lane, n + 1))] : 0.0; };
        auto chunkY = [&](int c) { return edge ? Y[at(base - 1, min(64 * c + lane + 1, n + 1))] : 0.0; };
        double cx = chunkX(0), cy = chunkY(0), nxC = chunkX(1), nyC = chunkY(1);
        int x = sL[min(max(-lane, 0), n)];  // letter of step 0
        double z = 0.0, tot = 0.0, lastX = 0.0, lastX2 = 0.0, lastY = 0.0;
        auto step = [&](int t) {
            const double upXn = from_lane_above(lastX2), upYn = from_lane_above(lastY);
            const int xn = sL[min(max(t + 1 - lane, 0), n)];  // next step's letter
            const int j = t - lane + 1;
            const double ex = lane_value(cx, t & 63), ey = lane_value(cy, t & 63);
            double nx = lastX, ny = lastY;
            if (rowOk && j >= 1 && j <= n + 1) {
                double upX = upXn, upY = upYn;
                if (lane == 0) {  // column 0 of every row is 0; row 0 is 0
                    upX = (!edge || j == 1) ? 0.0 : ex;
                    upY = edge ? ey : 0.0;
                }
                const double v = upX + upY + z + bw_scale();
                z = ra * v + rb * z;
                tot += v;
                nx = Srow[x] * v;
                ny = rd * v + re * upY;
                const size_t o = at(i, j);
                X[o] = nx;
                Y[o] = ny;
                Z[o] = z;
            }
            lastX2 = lastX;
            lastX = nx;
            lastY = ny;
            x = xn;
        };
        const int T = n + 64;
        int t = 0;
        for (; t + 4 <= T; t += 4) {
            if ((t & 63) == 0 && t > 0) {  // next chunk of the row above, and load the one after
                cx = nxC;
                cy = nyC;
                nxC = chunkX((t >> 6) + 1);
                nyC = chunkY((t >> 6) + 1);
            }
            step(t);
            step(t + 1);
            step(t + 2);
            step(t + 3);
        }
        if ((t & 63) == 0 && t > 0) {
            cx = nxC;
            cy = nyC;
        }
        for (; t < T; t++) step(t);
        // the strip's row sums, added in a fixed order
        double rs = rowOk ? tot : 0.0;
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) rs += __shfl_xor(rs, o, 64);
        vsum += rs;
        __threadfence();  // the next strip reads this strip's last row
    }
    if (lane == 0) {
        a.total[s] = vsum;
        a.wt[s] = a.weights[s] / (vsum * bw_scale());
    }
}

__global__ __launch_bounds__(64) void bw_backward_kernel(BWArgs a) {
#pragma clang fp contract(off)
    __shared__ double sS[64 * 21];
    extern __shared__ uint8_t sL[];
    const int s = a.seqBase + blockIdx.x, lane = threadIdx.x;
    const int n = a.len[s], L = a.L, A = a.A, P = n + 2;
    double *Wb = a.Wb + a.cellOff[s], *Yb = a.Yb + a.cellOff[s], *Zb = a.Zb + a.cellOff[s];
    auto at = [&](int ii, int jj) { return (size_t)ii * P + jj; };
    stage_letters(sL, a.letters + a.letterOff[s], n, lane);
    for (int j = lane; j <= n + 1; j += 64) {  // row L+1
        if (j >= 1) Wb[at(L + 1, j)] = 0.0;
        if (j <= n) Yb[at(L + 1, j)] = 0.0;
    }
    for (int top = L; top >= 0; top -= 64) {
        const int i = top - lane;
        const bool rowOk = i >= 0;
        const int r = rowOk ? i : 0;
        const double ra = a.rates[4 * r], rb = a.rates[4 * r + 1], rd = a.rates[4 * r + 2], re = a.rates[4 * r + 3];
        __syncthreads();
        stage_ratios(sS, a.ratio, A, lane, [&](int rr) { return top - rr >= 0 ? top - rr : -1; });
        __syncthreads();
        const double *Srow = sS + lane * (A + 1);
        if (rowOk) {
            Wb[at(i + 1, n + 1)] = 0.0;
            Zb[at(i, n + 1)] = 0.0;
        }
        // previous strip's last row: step t needs Wb[top+1][n-t+1] and Yb[top+1][n-t]
        const bool edge = top < L;
        auto chunkW = [&](int c) { return edge ? Wb[at(top + 1, min(max(n - 64 * c - lane + 1, 0), n + 1))] : 0.0; };
        auto chunkY = [&](int c) { return edge ? Yb[at(top + 1, min(max(n - 64 * c - lane, 0), n))] : 0.0; };
        double cw = chunkW(0), cy = chunkY(0), nwC = chunkW(1), nyC = chunkY(1);
        int x = sL[min(max(n + lane, 0), n)];  // letter of step 0
        double z = 0.0, lastW = 0.0, lastW2 = 0.0, lastY = 0.0;
        auto step = [&](int t) {
            const double dnWn = from_lane_above(lastW2), dnYn = from_lane_above(lastY);
            const int xn = sL[min(max(n - (t + 1) + lane, 0), n)];
            const int j = n - t + lane;
            const double ew = lane_value(cw, t & 63), ey = lane_value(cy, t & 63);
            double nw = lastW, ny = lastY;
            if (rowOk && j >= 0 && j <= n) {
                double dnW = dnWn, dnY = dnYn;
                if (lane == 0) {  // column n+1 of every backward row is 0; row L+1 is 0
                    dnW = (!edge || j == n) ? 0.0 : ew;
                    dnY = edge ? ey : 0.0;
                }
                const double xv = Srow[x] * dnW;
                const double w = xv + rd * dnY + ra * z + bw_scale();
                z = w + rb * z;
                nw = w;
                ny = w + re * dnY;
                const size_t o = at(i, j);
                Wb[o] = nw;
                Yb[o] = ny;
                Zb[o] = z;
            }
            lastW2 = lastW;
            lastW = nw;
            lastY = ny;
            x = xn;
        };
        const int T = n + 64;
        int t = 0;
        for (; t + 4 <= T; t += 4) {
            if ((t & 63) == 0 && t > 0) {
                cw = nwC;
                cy = nyC;
                nwC = chunkW((t >> 6) + 1);
                nyC = chunkY((t >> 6) + 1);
            }
            step(t);
            step(t + 1);
            step(t + 2);
            step(t + 3);
        }
        if ((t & 63) == 0 && t > 0) {
            cw = nwC;
            cy = nyC;
        }
        for (; t < T; t++) step(t);
        __threadfence();
    }
}

// Per DP row i (one lane), the sums of calculateTransitionCounts over j.
__global__ __launch_bounds__(64) void bw_counts_kernel(BWArgs a) {
#pragma clang fp contract(off)
    __shared__ double emis[21][64];
    const int s = a.seqBase + blockIdx.x, lane = threadIdx.x;
    const int n = a.len[s], P = n + 2, L = a.L, A = a.A;
    const int i = 1 + 64 * blockIdx.y + lane;
    for (int k = 0; k <= A; k++) emis[k][lane] = 0.0;
    if (i > L + 1) return;
    const uint8_t *seq = a.letters + a.letterOff[s];
    const size_t off = a.cellOff[s], up = (size_t)(i - 1) * P, here = (size_t)i * P;
    const double *X = a.X + off + here, *Z = a.Z + off + here, *Yu = a.Y + off + up;
    const double *Wu = a.Wb + off + up, *Yu_b = a.Yb + off + up, *Zu = a.Zb + off + up, *Wh = a.Wb + off + here;
    double betap = 0.0, beta = 0.0, epsp = 0.0, eps = 0.0, etap = 0.0;
    for (int j = 1; j <= n + 1; j++) {
        const double zf = Z[j - 1], wu = Wu[j - 1], yf = Yu[j];
        betap += zf * wu;
        beta += zf * Zu[j - 1];
        epsp += yf * wu;
        eps += yf * Yu_b[j - 1];
        etap += wu;
        emis[seq[j - 1]][lane] += X[j] * Wh[j];
    }
    beta -= betap;
    eps -= epsp;
    double gamma = 0.0;
    for (int k = 0; k <= A; k++) gamma += emis[k][lane];
    double *out = a.raw + ((size_t)s * (L + 1) + (i - 1)) * (kBwRaw + A);
    out[0] = gamma;
    out[1] = etap;
    out[2] = betap;
    out[3] = beta;
    out[4] = epsp;
    out[5] = eps;
    for (int k = 0; k < A; k++) out[kBwRaw + k] = emis[k][lane];
}

// counts[r][c] = sum over sequences, in order, of term * weight (the
// reference's counts[...] += term * wt), with notIns = etap 2^-480 + gamma of
// the row above + epsilonp.
__global__ void bw_reduce_kernel(BWReduceArgs a) {
#pragma clang fp contract(off)
    const int width = 7 + a.A, RW = kBwRaw + a.A;
    const int id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= (a.L + 1) * width) return;
    const int r = id / width, c = id % width;
    double acc = 0.0;
    if (c != 2 && !(c >= 7 && r == a.L)) {
        for (int s = 0; s < a.nseq; s++) {
            const double *R = a.raw + ((size_t)s * (a.L + 1) + r) * RW;
            double term;
            switch (c) {
            case 0: term = R[0]; break;
            case 1: term = R[1] * bw_scale() + (r > 0 ? R[0 - RW] : 0.0) + R[4]; break;
            case 3: term = R[2]; break;
            case 4: term = R[3]; break;
            case 5: term = R[4]; break;
            case 6: term = R[5]; break;
            default: term = R[kBwRaw + c - 7]; break;
            }
            acc += term * a.wt[s];
        }
    }
    a.counts[id] = acc;
}

}  // namespace

hipError_t launch_bw_forward(const BWArgs &a, int nseq, int maxLen, hipStream_t s) {
    if (nseq <= 0) return hipSuccess;
    const size_t lds = ((size_t)maxLen + 1 + 15) & ~(size_t)15;
    hipLaunchKernelGGL(bw_forward_kernel, dim3(nseq), dim3(64), lds, s, a);
    return hipGetLastError();
}

hipError_t launch_bw_backward(const BWArgs &a, int nseq, int maxLen, hipStream_t s) {
    if (nseq <= 0) return hipSuccess;
    const size_t lds = ((size_t)maxLen + 1 + 15) & ~(size_t)15;
    hipLaunchKernelGGL(bw_backward_kernel, dim3(nseq), dim3(64), lds, s, a);
    return hipGetLastError();
}

hipError_t launch_bw_counts(const BWArgs &a, int nseq, hipStream_t s) {
    if (nseq <= 0) return hipSuccess;
    hipLaunchKernelGGL(bw_counts_kernel, dim3(nseq, (a.L + 1 + 63) / 64), dim3(64), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_bw_reduce(const BWReduceArgs &a, hipStream_t s) {
    const int cells = (a.L + 1) * (7 + a.A);
    hipLaunchKernelGGL(bw_reduce_kernel, dim3((cells + 255) / 256), dim3(256), 0, s, a);
    return hipGetLastError();
}

}  // namespace dm
