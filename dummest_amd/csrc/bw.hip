// Baum-Welch E-step of the profile builder on gfx950 (dummest-build.cc:561-736):
// forward (X, Y, Z) and backward (Wb, Yb, Zb) over every ungapped sequence of
// an alignment against the current profile, then the expected event counts
// per profile row. FP64 throughout, no fused multiply-adds, every sum taken in
// the reference's order (j ascending within a row, sequences in order), so the
// counts equal the reference's except through the forward total v, which is
// summed row by row here instead of cell by cell (a rounding-level change of
// the per-sequence factor weight / (v 2^-480)).
//
// Layout: one wavefront per sequence for forward and backward. The profile
// rows are taken 64 at a time (a strip), lane k holding row base + k and
// running one column behind lane k - 1, so the values a lane needs from the row
// above arrive from its neighbour by a DPP shift; the strip's first lane reads
// the previous strip's last row back from the matrices, which every pass
// stores in full because the count pass needs all six. The count pass gives
// each row a lane that walks its row left to right, with the letter-indexed
// emission sums in LDS.
#include <hip/hip_runtime.h>

#include "dm_layout.h"

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

__global__ __launch_bounds__(64) void bw_forward_kernel(BWArgs a) {
#pragma clang fp contract(off)
    const int s = a.seqBase + blockIdx.x, lane = threadIdx.x;
    const int n = a.len[s], P = n + 2, L = a.L, A = a.A;
    const uint8_t *seq = a.letters + a.letterOff[s];
    double *X = a.X + a.cellOff[s], *Y = a.Y + a.cellOff[s], *Z = a.Z + a.cellOff[s];
    double *rowTot = a.rowTotal + (size_t)s * (L + 2);
    for (int j = lane; j <= n + 1; j += 64) {  // row 0
        if (j <= n) X[j] = 0.0;
        if (j >= 1) Y[j] = 0.0;
    }
    for (int base = 1; base <= L + 1; base += 64) {
        const int i = base + lane;
        const bool rowOk = i <= L + 1;
        const int r = rowOk ? i - 1 : 0;  // profile row of this DP row's transitions and emissions
        const double ra = a.rates[4 * r], rb = a.rates[4 * r + 1], rd = a.rates[4 * r + 2], re = a.rates[4 * r + 3];
        const double *S = a.ratio + (size_t)r * (A + 1);
        if (rowOk) {
            X[(size_t)(i - 1) * P] = 0.0;
            Z[(size_t)i * P] = 0.0;
        }
        const double *aboveX = X + (size_t)(base - 1) * P, *aboveY = Y + (size_t)(base - 1) * P;
        double z = 0.0, tot = 0.0, lastX = 0.0, lastX2 = 0.0, lastY = 0.0;
        for (int t = 0; t < n + 64; t++) {
            const double upXn = from_lane_above(lastX2), upYn = from_lane_above(lastY);
            const int j = t - lane + 1;
            double nx = lastX, ny = lastY;
            if (rowOk && j >= 1 && j <= n + 1) {
                double upX = upXn, upY = upYn;
                if (lane == 0) {  // column 0 of every row is 0
                    upX = (base == 1 || j == 1) ? 0.0 : aboveX[j - 1];
                    upY = base == 1 ? 0.0 : aboveY[j];
                }
                const double v = upX + upY + z + bw_scale();
                z = ra * v + rb * z;
                tot += v;
                nx = S[seq[j - 1]] * v;
                ny = rd * v + re * upY;
                const size_t at = (size_t)i * P + j;
                X[at] = nx;
                Y[at] = ny;
                Z[at] = z;
            }
            lastX2 = lastX;
            lastX = nx;
            lastY = ny;
        }
        if (rowOk) rowTot[i] = tot;
        __threadfence();  // the next strip's first lane reads this strip's last row
    }
}

__global__ __launch_bounds__(64) void bw_backward_kernel(BWArgs a) {
#pragma clang fp contract(off)
    const int s = a.seqBase + blockIdx.x, lane = threadIdx.x;
    const int n = a.len[s], P = n + 2, L = a.L, A = a.A;
    const uint8_t *seq = a.letters + a.letterOff[s];
    double *Wb = a.Wb + a.cellOff[s], *Yb = a.Yb + a.cellOff[s], *Zb = a.Zb + a.cellOff[s];
    for (int j = lane; j <= n + 1; j += 64) {  // row L+1
        if (j >= 1) Wb[(size_t)(L + 1) * P + j] = 0.0;
        if (j <= n) Yb[(size_t)(L + 1) * P + j] = 0.0;
    }
    for (int top = L; top >= 0; top -= 64) {
        const int i = top - lane;
        const bool rowOk = i >= 0;
        const int r = rowOk ? i : 0;
        const double ra = a.rates[4 * r], rb = a.rates[4 * r + 1], rd = a.rates[4 * r + 2], re = a.rates[4 * r + 3];
        const double *S = a.ratio + (size_t)r * (A + 1);
        if (rowOk) {
            Wb[(size_t)(i + 1) * P + n + 1] = 0.0;
            Zb[(size_t)i * P + n + 1] = 0.0;
        }
        const double *belowW = Wb + (size_t)(top + 1) * P, *belowY = Yb + (size_t)(top + 1) * P;
        double z = 0.0, lastW = 0.0, lastW2 = 0.0, lastY = 0.0;
        for (int t = 0; t < n + 64; t++) {
            const double dnWn = from_lane_above(lastW2), dnYn = from_lane_above(lastY);
            const int j = n - t + lane;
            double nw = lastW, ny = lastY;
            if (rowOk && j >= 0 && j <= n) {
                double dnW = dnWn, dnY = dnYn;
                if (lane == 0) {  // column n+1 of every backward row is 0
                    dnW = (top == L || j == n) ? 0.0 : belowW[j + 1];
                    dnY = top == L ? 0.0 : belowY[j];
                }
                const double x = S[seq[j]] * dnW;
                const double w = x + rd * dnY + ra * z + bw_scale();
                z = w + rb * z;
                nw = w;
                ny = w + re * dnY;
                const size_t at = (size_t)i * P + j;
                Wb[at] = nw;
                Yb[at] = ny;
                Zb[at] = z;
            }
            lastW2 = lastW;
            lastW = nw;
            lastY = ny;
        }
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

__global__ void bw_weights_kernel(BWReduceArgs a) {
#pragma clang fp contract(off)
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= a.nseq) return;
    const double *rt = a.rowTotal + (size_t)s * (a.L + 2);
    double v = 0.0;
    for (int i = 1; i <= a.L + 1; i++) v += rt[i];
    a.total[s] = v;
    a.wt[s] = a.weights[s] / (v * bw_scale());
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

hipError_t launch_bw_forward(const BWArgs &a, int nseq, hipStream_t s) {
    if (nseq <= 0) return hipSuccess;
    hipLaunchKernelGGL(bw_forward_kernel, dim3(nseq), dim3(64), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_bw_backward(const BWArgs &a, int nseq, hipStream_t s) {
    if (nseq <= 0) return hipSuccess;
    hipLaunchKernelGGL(bw_backward_kernel, dim3(nseq), dim3(64), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_bw_counts(const BWArgs &a, int nseq, hipStream_t s) {
    if (nseq <= 0) return hipSuccess;
    hipLaunchKernelGGL(bw_counts_kernel, dim3(nseq, (a.L + 1 + 63) / 64), dim3(64), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_bw_reduce(const BWReduceArgs &a, hipStream_t s) {
    if (a.nseq > 0) hipLaunchKernelGGL(bw_weights_kernel, dim3((a.nseq + 63) / 64), dim3(64), 0, s, a);
    const int cells = (a.L + 1) * (7 + a.A);
    hipLaunchKernelGGL(bw_reduce_kernel, dim3((cells + 255) / 256), dim3(256), 0, s, a);
    return hipGetLastError();
}

}  // namespace dm
