// One-electron integrals (S, T, V, dipole) and the one-electron gradient on
// gfx950: one thread per shell pair (s1 >= s2, the reference's pair order).
// Gradient contributions are accumulated per workgroup in LDS and flushed
// once, both as exact fixed-point integers (order-independent).

#include "aimdhf_internal.h"
#include "md_oneel.h"

namespace aimdhf {
using namespace md;

__device__ __forceinline__ void pair_of(int P, int& s1, int& s2) {
    int r = (int)((sqrt(8.0 * P + 1.0) - 1.0) * 0.5);
    while ((r + 1) * (r + 2) / 2 <= P) ++r;
    while (r * (r + 1) / 2 > P) --r;
    s1 = r;
    s2 = P - r * (r + 1) / 2;
}

__global__ void k_oneel(OneElArgs g, double* S, double* T, double* M, double* V) {
    int P = blockIdx.x * blockDim.x + threadIdx.x;
    if (P >= g.npair) return;
    int s1, s2;
    pair_of(P, s1, s2);
    ShellView B{g.shell_atom, g.shell_l, g.shell_prim, g.shell_ao, g.prim_exp, g.prim_coef, g.pos};
    const double O[3] = {0.0, 0.0, 0.0};
    overlap_type_pair(B, g.nao, s1, s2, O, S, T, M);
    nuclear_pair(B, g.nao, s1, s2, g.natm, g.charges, g.pos, g.boys, V);
}

// fixed-point gradient accumulator in LDS: lds[2 * (3 * atom + d) + {0: hi, 1: lo}]
__device__ __forceinline__ void lds_add(long long* lds, int atom, int d, double x) {
    long long h, l;
    fx_split(x, h, l);
    if (h) atomicAdd((unsigned long long*)&lds[2 * (3 * atom + d)], (unsigned long long)h);
    if (l) atomicAdd((unsigned long long*)&lds[2 * (3 * atom + d) + 1], (unsigned long long)l);
}

__device__ void lds_flush(long long* lds, int natm, long long* out) {
    __syncthreads();
    for (int i = threadIdx.x; i < 6 * natm; i += blockDim.x)
        if (lds[i]) atomicAdd((unsigned long long*)&out[i], (unsigned long long)lds[i]);
}

__global__ void k_grad_one(OneElArgs g, const double* D, const double* W, long long* out) {
    extern __shared__ long long lds[];
    for (int i = threadIdx.x; i < 6 * g.natm; i += blockDim.x) lds[i] = 0;
    __syncthreads();
    ShellView B{g.shell_atom, g.shell_l, g.shell_prim, g.shell_ao, g.prim_exp, g.prim_coef, g.pos};
    for (int P = blockIdx.x * blockDim.x + threadIdx.x; P < g.npair; P += gridDim.x * blockDim.x) {
        int s1, s2;
        pair_of(P, s1, s2);
        int a1 = g.shell_atom[s1], a2 = g.shell_atom[s2];
        if (s1 != s2 && a1 != a2) {
            // d/dR [tr(D T) - tr(W S)]: the overlap term enters with -W
            double gg[3];
            ovlkin_grad_pair(B, g.nao, s1, s2, W, D, gg);
            for (int d = 0; d < 3; ++d) {
                lds_add(lds, a1, d, gg[d]);
                lds_add(lds, a2, d, -gg[d]);
            }
        }
        nuclear_grad_pair(B, g.nao, s1, s2, g.natm, g.charges, D, g.boys,
                          [&](int atom, int d, double x) { lds_add(lds, atom, d, x); });
    }
    lds_flush(lds, g.natm, out);
}

__global__ void k_negate(const double* W, size_t n, double* out) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = -W[i];
}

__global__ void k_fx_to_double(const long long* fx, int n, double* out) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = fx_value(fx[2 * i], fx[2 * i + 1]);
}

void launch_oneel(const OneElArgs& g, double* S, double* T, double* M, double* V, hipStream_t st) {
    if (g.npair) k_oneel<<<dim3((g.npair + 63) / 64), dim3(64), 0, st>>>(g, S, T, M, V);
}

void launch_grad_one(const OneElArgs& g, const double* D, const double* Wneg, long long* out, hipStream_t st) {
    int blocks = std::min(1024, (g.npair + 63) / 64);
    if (g.npair) k_grad_one<<<dim3(blocks), dim3(64), sizeof(long long) * 6 * g.natm, st>>>(g, D, Wneg, out);
}

void launch_negate(const double* W, size_t n, double* out, hipStream_t st) {
    if (n) k_negate<<<dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st>>>(W, n, out);
}

void launch_fx_to_double(const long long* fx, int n, double* out, hipStream_t st) {
    if (n) k_fx_to_double<<<dim3((n + 255) / 256), dim3(256), 0, st>>>(fx, n, out);
}

}  // namespace aimdhf
