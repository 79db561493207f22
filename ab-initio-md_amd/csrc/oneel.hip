// One-electron integrals (S, T, V, dipole) and the one-electron gradient on
// gfx950, per shell-pair class (l_A >= l_B, compile-time shapes):
//   S, T, dipole   one thread per shell pair;
//   V              one wavefront per pair, lanes over the nuclei, summed by a
//                  fixed butterfly (deterministic, no atomics);
//   gradient       one wavefront per pair, lanes over primitive pairs
//                  (overlap/kinetic) and nuclei (nuclear attraction); per-atom
//                  sums in exact fixed point (order-independent), staged in
//                  LDS per workgroup and flushed once.

#include "aimdhf_internal.h"
#include "md_oneel.h"

namespace aimdhf {
using namespace md;

namespace {
constexpr int la_of(int pc) { return pc == 0 ? 0 : pc <= 2 ? 1 : 2; }
constexpr int lb_of(int pc) { return pc - la_of(pc) * (la_of(pc) + 1) / 2; }

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
    return v;
}

__device__ __forceinline__ ShellView shells(const OneElArgs& g) {
    return ShellView{g.shell_atom, g.shell_l, g.shell_prim, g.shell_ao, g.prim_exp, g.prim_coef, g.pos};
}

// first element of the matrices of shell s's replica
__device__ __forceinline__ size_t mat_base(const OneElArgs& g, int s) {
    return g.shell_rep ? (size_t)g.shell_rep[s] * g.ld * g.ld : 0;
}

// the nuclei that act on shell s's electrons: its replica's atoms
__device__ __forceinline__ void nuclei_of(const OneElArgs& g, int s, int& c0, int& c1) {
    if (g.rep_atom) {
        const int r = g.shell_rep[s];
        c0 = g.rep_atom[r];
        c1 = g.rep_atom[r + 1];
    } else {
        c0 = 0;
        c1 = g.natm;
    }
}

template <int LI, int LJ>
__global__ __launch_bounds__(64) void k_oneel_st(OneElArgs g, const int* pairs, int np, double* S, double* T,
                                                 double* M) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= np) return;
    const int P = pairs[t];
    const int si = g.pair_a[P], sj = g.pair_b[P];
    constexpr int NI = NC(LI), NJ = NC(LJ);
    double s[NI * NJ] = {}, tk[NI * NJ] = {}, m[3 * NI * NJ] = {};
    const double O[3] = {0.0, 0.0, 0.0};
    overlap_type_acc<LI, LJ>(shells(g), si, sj, O, 0, 1, s, tk, m);
    const size_t n = g.ld, nn = g.mstride, base = mat_base(g, si);
    const int oi = g.shell_ao[si], oj = g.shell_ao[sj];
#pragma unroll
    for (int q = 0; q < NI * NJ; ++q) {
        const size_t u = base + (size_t)(oi + q / NJ) * n + oj + q % NJ,
                     v = base + (size_t)(oj + q % NJ) * n + oi + q / NJ;
        S[u] = S[v] = s[q];
        T[u] = T[v] = tk[q];
#pragma unroll
        for (int d = 0; d < 3; ++d) M[d * nn + u] = M[d * nn + v] = m[d * NI * NJ + q];
    }
}

template <int LI, int LJ>
__global__ __launch_bounds__(64) void k_oneel_v(OneElArgs g, const int* pairs, double* V) {
    const int P = pairs[blockIdx.x];
    const int si = g.pair_a[P], sj = g.pair_b[P];
    constexpr int NI = NC(LI), NJ = NC(LJ);
    double v[NI * NJ] = {};
    int c0, c1;
    nuclei_of(g, si, c0, c1);
    nuclear_acc<LI, LJ>(shells(g), si, sj, c0 + (int)threadIdx.x, 64, c1, g.charges, g.pos, g.boys, v);
    const size_t n = g.ld, base = mat_base(g, si);
    const int oi = g.shell_ao[si], oj = g.shell_ao[sj];
#pragma unroll
    for (int q = 0; q < NI * NJ; ++q) {
        const double x = wave_sum(v[q]);
        if (threadIdx.x == 0) {
            V[base + (size_t)(oi + q / NJ) * n + oj + q % NJ] = x;
            V[base + (size_t)(oj + q % NJ) * n + oi + q / NJ] = x;
        }
    }
}

// fixed-point accumulator: acc[2 * (3 * atom + d) + {0: hi, 1: lo}]
__device__ __forceinline__ void fx_add(long long* acc, int atom, int d, double x) {
    long long h, l;
    fx_split(x, h, l);
    if (h) atomicAdd((unsigned long long*)&acc[2 * (3 * atom + d)], (unsigned long long)h);
    if (l) atomicAdd((unsigned long long*)&acc[2 * (3 * atom + d) + 1], (unsigned long long)l);
}

// 4 wavefronts per workgroup, grid-stride over the pairs of the class; LDS
// staging when use_lds (6 * natm * 8 bytes of dynamic shared memory)
template <int LI, int LJ>
__global__ __launch_bounds__(256) void k_grad_one_cls(OneElArgs g, const int* pairs, int np, const double* D,
                                                      const double* Wneg, long long* out, int use_lds) {
    extern __shared__ long long lds[];
    long long* acc = use_lds ? lds : out;
    if (use_lds) {
        for (int i = threadIdx.x; i < 6 * g.natm; i += blockDim.x) lds[i] = 0;
        __syncthreads();
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const ShellView B = shells(g);
    for (int t = blockIdx.x * 4 + wave; t < np; t += gridDim.x * 4) {
        const int P = pairs[t];
        const int si = g.pair_a[P], sj = g.pair_b[P];
        const int ai = g.shell_atom[si], aj = g.shell_atom[sj];
        const size_t base = mat_base(g, si);
        double gA[3] = {0.0, 0.0, 0.0}, gB[3] = {0.0, 0.0, 0.0};
        if ((g.parts & 1) && si != sj && ai != aj) {
            // d/dR [tr(D T) - tr(W S)]: the overlap term enters with -W
            double gk[3] = {0.0, 0.0, 0.0};
            ovlkin_grad_acc<LI, LJ>(B, g.ld, si, sj, Wneg + base, D + base, lane, 64, gk);
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                gA[d] += gk[d];
                gB[d] -= gk[d];
            }
        }
        if (g.parts & 2) {
            int c0, c1;
            nuclei_of(g, si, c0, c1);
            nuclear_grad_acc<LI, LJ>(B, g.ld, si, sj, c0 + lane, 64, c1, g.charges, D + base, g.boys,
                                     [&](int atom, int d, double x) { fx_add(acc, atom, d, x); }, gA, gB);
        }
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const double xa = wave_sum(gA[d]), xb = wave_sum(gB[d]);
            if (lane == 0) {
                fx_add(acc, ai, d, xa);
                fx_add(acc, aj, d, xb);
            }
        }
    }
    if (use_lds) {
        __syncthreads();
        for (int i = threadIdx.x; i < 6 * g.natm; i += blockDim.x)
            if (lds[i]) atomicAdd((unsigned long long*)&out[i], (unsigned long long)lds[i]);
    }
}

__global__ void k_negate(const double* W, size_t n, double* out) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = -W[i];
}

__global__ void k_fx_to_double(const long long* fx, int n, double* out) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = fx_value(fx[2 * i], fx[2 * i + 1]);
}
}  // namespace

void launch_oneel(const OneElArgs& g, double* S, double* T, double* M, double* V, hipStream_t st) {
    static_for<0, 6>([&](auto cc) {
        constexpr int C = decltype(cc)::value, LI = la_of(C), LJ = lb_of(C);
        const int np = g.ncls_host[C];
        if (!np) return;
        const int* pairs = g.cls_pairs + g.cls_off_host[C];
        k_oneel_st<LI, LJ><<<dim3((np + 63) / 64), dim3(64), 0, st>>>(g, pairs, np, S, T, M);
        k_oneel_v<LI, LJ><<<dim3(np), dim3(64), 0, st>>>(g, pairs, V);
    });
}

void launch_grad_one(const OneElArgs& g, const double* D, const double* Wneg, long long* out, hipStream_t st) {
    const int use_lds = g.natm <= 1024;
    const size_t shm = use_lds ? sizeof(long long) * 6 * g.natm : 0;
    static_for<0, 6>([&](auto cc) {
        constexpr int C = decltype(cc)::value, LI = la_of(C), LJ = lb_of(C);
        const int np = g.ncls_host[C];
        if (!np) return;
        const int blocks = std::min(512, (np + 3) / 4);
        k_grad_one_cls<LI, LJ><<<dim3(blocks), dim3(256), shm, st>>>(g, g.cls_pairs + g.cls_off_host[C], np, D, Wneg,
                                                                      out, use_lds);
    });
}

void launch_negate(const double* W, size_t n, double* out, hipStream_t st) {
    if (n) k_negate<<<dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st>>>(W, n, out);
}

void launch_fx_to_double(const long long* fx, int n, double* out, hipStream_t st) {
    if (n) k_fx_to_double<<<dim3((n + 255) / 256), dim3(256), 0, st>>>(fx, n, out);
}

}  // namespace aimdhf
