// Trace-correcting purification step (Niklasson, Phys. Rev. B 66, 155115
// (2002)) of the SCF density: Xout = X2 if tr X > nocc, else 2 X - X2, with
// X2 = X X formed by the caller (a dense GEMM). Every block recomputes
// tr X from the n diagonal elements itself, so one launch needs no
// reduction pass and no memory clearing (it can be captured in a graph).

#include <algorithm>

#include "aimdhf_internal.h"

namespace aimdhf {

namespace {
__global__ __launch_bounds__(256) void k_tc2_step(int n, const double* __restrict__ X, const double* __restrict__ X2,
                                                  double* __restrict__ Xout, double nocc) {
    __shared__ double part[4];
    double t = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) t += X[(size_t)i * n + i];
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) t += __shfl_xor(t, m, 64);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = t;
    __syncthreads();
    const double tr = part[0] + part[1] + part[2] + part[3];
    const bool square = tr > nocc;
    const size_t nn = (size_t)n * n;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < nn; i += (size_t)gridDim.x * blockDim.x)
        Xout[i] = square ? X2[i] : 2.0 * X[i] - X2[i];
}
}  // namespace

void launch_tc2_step(int n, const double* X, const double* X2, double* Xout, double nocc, hipStream_t st) {
    const size_t nn = (size_t)n * n;
    int blocks = (int)std::min<size_t>((nn + 255) / 256, 256);
    if (nn) k_tc2_step<<<dim3(blocks), dim3(256), 0, st>>>(n, X, X2, Xout, nocc);
}

}  // namespace aimdhf
