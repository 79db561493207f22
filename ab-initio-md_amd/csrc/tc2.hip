// Trace-correcting purification step (Niklasson, Phys. Rev. B 66, 155115
// (2002)) of the SCF density: Xout = X2 if tr X > nocc, else 2 X - X2, with
// X2 = X X formed by the caller (a dense GEMM). Every block recomputes
// tr X and tr X2 (= sum X_ij^2 for symmetric X) from the n diagonal elements
// itself, so one launch needs no reduction pass and no memory clearing (it can
// be captured in a graph).
//
// Past convergence the step is not harmless in floating point: each branch
// doubles the distance from {0, 1} of the eigenvalues on one side, and a run
// of steps beyond convergence drives the iterate to overflow. With freeze > 0
// an iterate whose idempotency error |tr X - tr X2| is below freeze is copied
// unchanged, so any number of extra steps leaves a converged iterate as it is.

#include <algorithm>

#include "aimdhf_internal.h"

namespace aimdhf {

namespace {
__global__ __launch_bounds__(256) void k_tc2_step(int n, const double* __restrict__ X, const double* __restrict__ X2,
                                                  double* __restrict__ Xout, double nocc, double freeze) {
    __shared__ double part[2][4];
    double t = 0.0, t2 = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        t += X[(size_t)i * n + i];
        t2 += X2[(size_t)i * n + i];
    }
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) {
        t += __shfl_xor(t, m, 64);
        t2 += __shfl_xor(t2, m, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        part[0][threadIdx.x >> 6] = t;
        part[1][threadIdx.x >> 6] = t2;
    }
    __syncthreads();
    const double tr = part[0][0] + part[0][1] + part[0][2] + part[0][3];
    const double tr2 = part[1][0] + part[1][1] + part[1][2] + part[1][3];
    const bool hold = freeze > 0.0 && fabs(tr - tr2) < freeze;
    const bool square = tr > nocc;
    const size_t nn = (size_t)n * n;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < nn; i += (size_t)gridDim.x * blockDim.x)
        Xout[i] = hold ? X[i] : square ? X2[i] : 2.0 * X[i] - X2[i];
}
// X0 = (e_max - F) / (e_max - e_min) with the Gershgorin bounds of F (rows i:
// F_ii -/+ sum_{j != i} |F_ij|), in two launches: one wavefront per row writes
// the row's bounds, then every block of the fill reduces the bounds itself (min
// and max do not depend on the order) and writes its rows of X0. The row bounds
// and the partial sums of the error below live in a file-scope device array, so
// calls on one device must not overlap; matrices above TC2_ROWS rows take the
// one-workgroup kernels.
constexpr int TC2_ROWS = 16384;
constexpr int TC2_WAVES = 16;   // rows per block of the bounds and error passes
__device__ double tc2_rows[2 * TC2_ROWS];

__device__ __forceinline__ double wave_sum(double r) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) r += __shfl_xor(r, m, 64);
    return r;
}

__global__ __launch_bounds__(1024) void k_tc2_bounds(int n, const double* __restrict__ F) {
    const int lane = threadIdx.x & 63, i = blockIdx.x * TC2_WAVES + (threadIdx.x >> 6);
    if (i >= n) return;
    const double* row = F + (size_t)i * n;
    double r = 0.0;
    for (int j = lane; j < n; j += 64) r += j == i ? 0.0 : fabs(row[j]);
    r = wave_sum(r);
    if (lane == 0) {
        tc2_rows[i] = row[i] - r;
        tc2_rows[TC2_ROWS + i] = row[i] + r;
    }
}

__global__ __launch_bounds__(256) void k_tc2_fill(int n, const double* __restrict__ F, double* __restrict__ X0) {
    __shared__ double lo[4], hi[4];
    double mn = 1e300, mx = -1e300;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        mn = fmin(mn, tc2_rows[i]);
        mx = fmax(mx, tc2_rows[TC2_ROWS + i]);
    }
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) {
        mn = fmin(mn, __shfl_xor(mn, m, 64));
        mx = fmax(mx, __shfl_xor(mx, m, 64));
    }
    if ((threadIdx.x & 63) == 0) {
        lo[threadIdx.x >> 6] = mn;
        hi[threadIdx.x >> 6] = mx;
    }
    __syncthreads();
    const double emin = fmin(fmin(lo[0], lo[1]), fmin(lo[2], lo[3]));
    const double emax = fmax(fmax(hi[0], hi[1]), fmax(hi[2], hi[3]));
    const double inv = 1.0 / (emax - emin);
    for (int i = blockIdx.x; i < n; i += gridDim.x) {
        const size_t r0 = (size_t)i * n;
        for (int j = threadIdx.x; j < n; j += blockDim.x) X0[r0 + j] = ((i == j ? emax : 0.0) - F[r0 + j]) * inv;
    }
}

// the same bounds in one workgroup (matrices above TC2_ROWS rows)
__global__ __launch_bounds__(1024) void k_tc2_init(int n, const double* __restrict__ F, double* __restrict__ X0) {
    __shared__ double lo[16], hi[16];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    double mn = 1e300, mx = -1e300;
    for (int i = w; i < n; i += nw) {
        const double* row = F + (size_t)i * n;
        double r = 0.0;
        for (int j = lane; j < n; j += 64) r += j == i ? 0.0 : fabs(row[j]);
        r = wave_sum(r);
        mn = fmin(mn, row[i] - r);
        mx = fmax(mx, row[i] + r);
    }
    if (lane == 0) {
        lo[w] = mn;
        hi[w] = mx;
    }
    __syncthreads();
    double emin = lo[0], emax = hi[0];
    for (int k = 1; k < nw; ++k) {
        emin = fmin(emin, lo[k]);
        emax = fmax(emax, hi[k]);
    }
    const double inv = 1.0 / (emax - emin);
    for (int i = w; i < n; i += nw) {
        const size_t r0 = (size_t)i * n;
        for (int j = lane; j < n; j += 64) X0[r0 + j] = ((i == j ? emax : 0.0) - F[r0 + j]) * inv;
    }
}

// |tr X - sum_ij X_ij^2| (the idempotency error of a symmetric X): one
// wavefront per row writes the row's part, one wavefront adds the rows in a
// fixed order (lane k sums rows k, k + 64, ..., then a fixed butterfly), so the
// value does not depend on scheduling
__global__ __launch_bounds__(1024) void k_tc2_error_rows(int n, const double* __restrict__ X) {
    const int lane = threadIdx.x & 63, i = blockIdx.x * TC2_WAVES + (threadIdx.x >> 6);
    if (i >= n) return;
    const double* row = X + (size_t)i * n;
    double s = 0.0;
    for (int j = lane; j < n; j += 64) {
        const double x = row[j];
        s += (i == j ? x : 0.0) - x * x;
    }
    s = wave_sum(s);
    if (lane == 0) tc2_rows[i] = s;
}

__global__ __launch_bounds__(64) void k_tc2_error_sum(int n, double* out) {
    double s = 0.0;
    for (int i = threadIdx.x; i < n; i += 64) s += tc2_rows[i];
    s = wave_sum(s);
    if (threadIdx.x == 0) out[0] = fabs(s);
}

// the same error in one workgroup (matrices above TC2_ROWS rows)
__global__ __launch_bounds__(1024) void k_tc2_error(int n, const double* __restrict__ X, double* out) {
    __shared__ double part[16];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    double s = 0.0;
    for (int i = w; i < n; i += nw) {
        const double* row = X + (size_t)i * n;
        for (int j = lane; j < n; j += 64) {
            const double x = row[j];
            s += (i == j ? x : 0.0) - x * x;
        }
    }
    s = wave_sum(s);
    if (lane == 0) part[w] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int k = 0; k < nw; ++k) t += part[k];
        out[0] = fabs(t);
    }
}
}  // namespace

void launch_tc2_init(int n, const double* F, double* X0, hipStream_t st) {
    if (n <= 0) return;
    if (n > TC2_ROWS) {
        k_tc2_init<<<dim3(1), dim3(1024), 0, st>>>(n, F, X0);
        return;
    }
    k_tc2_bounds<<<dim3((n + TC2_WAVES - 1) / TC2_WAVES), dim3(64 * TC2_WAVES), 0, st>>>(n, F);
    k_tc2_fill<<<dim3((unsigned)std::min(n, 1024)), dim3(256), 0, st>>>(n, F, X0);
}

void launch_tc2_error(int n, const double* X, double* out, hipStream_t st) {
    if (n <= 0) return;
    if (n > TC2_ROWS) {
        k_tc2_error<<<dim3(1), dim3(1024), 0, st>>>(n, X, out);
        return;
    }
    k_tc2_error_rows<<<dim3((n + TC2_WAVES - 1) / TC2_WAVES), dim3(64 * TC2_WAVES), 0, st>>>(n, X);
    k_tc2_error_sum<<<dim3(1), dim3(64), 0, st>>>(n, out);
}

void launch_tc2_step(int n, const double* X, const double* X2, double* Xout, double nocc, double freeze,
                     hipStream_t st) {
    const size_t nn = (size_t)n * n;
    int blocks = (int)std::min<size_t>((nn + 255) / 256, 256);
    if (nn) k_tc2_step<<<dim3(blocks), dim3(256), 0, st>>>(n, X, X2, Xout, nocc, freeze);
}

}  // namespace aimdhf
