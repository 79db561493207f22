// Per-iteration helpers of the owner-computes J/K build: densities carried
// into Hermite space for the J-engine ket streams, and the fixed-order
// assembly of the per-class partial J and K blocks into full matrices.

#include "aimdhf_internal.h"
#include "md_jk2.h"

namespace aimdhf {
using namespace md;

namespace {
constexpr int la_of(int pc) { return pc == 0 ? 0 : pc <= 2 ? 1 : 2; }
constexpr int lb_of(int pc) { return pc - la_of(pc) * (la_of(pc) + 1) / 2; }

template <int CQ>
__global__ void k_dh(DhArgs a, long long n_entries) {
    constexpr int LC = la_of(CQ), LD = lb_of(CQ), NHK = NH(LC + LD);
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_entries) return;
    const long long e = a.pbase[CQ] + i;
    const double* K = a.recs + a.krec[e];
    const int Q = a.kpair[e];
    const int C = a.pair_a[Q], D = a.pair_b[Q];
    const double w = (C != D) ? 2.0 : 1.0;
    const size_t n = a.nao, nn = n * n;
    double* out = a.dh + a.dh_base[CQ] + i * (a.nset * NHK);
    for (int s = 0; s < a.nset; ++s)
        density_hermite<LC, LD>(K, a.dens + s * nn, n, a.shell_ao[C], a.shell_ao[D], w, out + s * NHK);
}

__global__ void k_j_assemble(JAsmArgs a) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.npair * a.nset) return;
    const int P = t / a.nset, s = t % a.nset;
    const int A = a.pair_a[P], B = a.pair_b[P];
    const int nA = a.shell_ao[A + 1] - a.shell_ao[A], nB = a.shell_ao[B + 1] - a.shell_ao[B];
    const int oA = a.shell_ao[A], oB = a.shell_ao[B];
    const size_t n = a.nao;
    double* J = a.J + (size_t)s * n * n;
    for (int ab = 0; ab < nA * nB; ++ab) {
        double v = 0.0;
        for (int c = 0; c < 6; ++c)
            if (a.cq_valid[c]) v += a.jpart[(((size_t)c * a.npair + P) * a.nset + s) * 36 + ab];
        const int i = ab / nB, j = ab % nB;
        J[(size_t)(oA + i) * n + oB + j] = v;
        J[(size_t)(oB + j) * n + oA + i] = v;
    }
}

__global__ void k_k_assemble(const double* kpart, int nset, int n, const int* ao_shell, double* K) {
    const size_t nn = (size_t)n * n;
    size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)nset * nn) return;
    const size_t s = t / nn, r = (t % nn) / n, c = t % n;
    const bool own = ao_shell[r] >= ao_shell[c];
    const size_t idx = s * nn + (own ? r * n + c : c * n + r);
    double v = 0.0;
    for (int sl = 0; sl < 9; ++sl) v += kpart[(size_t)sl * nset * nn + idx];
    K[t] = v;
}

__global__ void k_zero_d(double* p, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0;
}
}  // namespace

void launch_dh(const DhArgs& a, const long long* n_per_class, hipStream_t st) {
    static_for<0, 6>([&](auto cc) {
        constexpr int CQ = decltype(cc)::value;
        long long m = n_per_class[CQ];
        if (m > 0) k_dh<CQ><<<dim3((unsigned)((m + 127) / 128)), dim3(128), 0, st>>>(a, m);
    });
}

void launch_j_assemble(const JAsmArgs& a, hipStream_t st) {
    int m = a.npair * a.nset;
    if (m > 0) k_j_assemble<<<dim3((m + 127) / 128), dim3(128), 0, st>>>(a);
}

void launch_k_assemble(const double* kpart, int nset, int n, const int* ao_shell, double* K, hipStream_t st) {
    size_t m = (size_t)nset * n * n;
    if (m) k_k_assemble<<<dim3((unsigned)((m + 255) / 256)), dim3(256), 0, st>>>(kpart, nset, n, ao_shell, K);
}

void launch_zero_d(double* p, size_t n, hipStream_t st) {
    if (n) k_zero_d<<<dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st>>>(p, n);
}

}  // namespace aimdhf
