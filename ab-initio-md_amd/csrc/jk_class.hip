// Integral-direct J/K kernel of one quartet class; this file is compiled once
// per class with -DAIMDHF_QCLASS=k (k = hi (hi+1)/2 + lo over the pair
// classes ss, ps, pp, ds, dp, dd) so the 21 instantiations build in parallel.

#include "aimdhf_internal.h"
#include "md_digest.h"

#ifndef AIMDHF_QCLASS
#error "compile with -DAIMDHF_QCLASS=<0..20>"
#endif

namespace aimdhf {
using namespace md;

namespace {
constexpr int QC = AIMDHF_QCLASS;
constexpr int pc_hi() { int h = 0; while ((h + 1) * (h + 2) / 2 <= QC) ++h; return h; }
constexpr int pc_lo() { return QC - pc_hi() * (pc_hi() + 1) / 2; }
constexpr int la_of(int pc) { return pc == 0 ? 0 : pc <= 2 ? 1 : 2; }
constexpr int lb_of(int pc) { return pc - la_of(pc) * (la_of(pc) + 1) / 2; }
constexpr int LA = la_of(pc_hi()), LB = lb_of(pc_hi()), LC = la_of(pc_lo()), LD = lb_of(pc_lo());
using Qt = Quartet<LA, LB, LC, LD>;
constexpr bool SPLIT = split_by_cd<LA, LB, LC, LD>();

__device__ __forceinline__ void fx_atomic(long long* hi, long long* lo, double x) {
    long long h, l;
    fx_split(x, h, l);
    if (h) atomicAdd((unsigned long long*)hi, (unsigned long long)h);
    if (l) atomicAdd((unsigned long long*)lo, (unsigned long long)l);
}

struct Sink {
    long long *Jh, *Jl, *Kh, *Kl;
    __device__ void j(size_t i, double x) const { fx_atomic(Jh + i, Jl + i, x); }
    __device__ void k(size_t i, double x) const { fx_atomic(Kh + i, Kl + i, x); }
};

__device__ __forceinline__ Sink sink_of(const JKArgs& a, int s) {
    size_t nn = (size_t)a.nao * a.nao;
    return Sink{a.fx + (4 * s + 0) * nn, a.fx + (4 * s + 1) * nn, a.fx + (4 * s + 2) * nn, a.fx + (4 * s + 3) * nn};
}

template <int Q>
__global__ __launch_bounds__(64) void k_jk_class(JKArgs a, const int2* __restrict__ list, long long nq) {
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    constexpr int NSPLIT = SPLIT ? Qt::NCD : 1;
    if (t >= nq * NSPLIT) return;
    const int2 pq = list[t / NSPLIT];
    const int sA = a.pair_a[pq.x], sB = a.pair_b[pq.x], sC = a.pair_a[pq.y], sD = a.pair_b[pq.y];
    double w = 1.0;
    if (sA == sB) w *= 0.5;
    if (sC == sD) w *= 0.5;
    if (pq.x == pq.y) w *= 0.5;
    const int oA = a.shell_ao[sA], oB = a.shell_ao[sB], oC = a.shell_ao[sC], oD = a.shell_ao[sD];
    const double* rb = a.recs + a.rec_off[pq.x];
    const double* rk = a.recs + a.rec_off[pq.y];
    if constexpr (SPLIT) {
        const int cd = (int)(t % NSPLIT);
        double acc[Qt::NAB];
        eri_quartet_cd<LA, LB, LC, LD>(rb, a.npp[pq.x], rk, a.npp[pq.y], a.boys, cd, acc);
        digest_cd<LA, LB, LC, LD>(acc, a.dens, a.nao, a.nset, [&](int s) { return sink_of(a, s); }, oA, oB, oC, oD, cd, w);
    } else {
        double acc[Qt::NAB * Qt::NCD];
        eri_quartet<LA, LB, LC, LD>(rb, a.npp[pq.x], rk, a.npp[pq.y], a.boys, acc);
        digest_full<LA, LB, LC, LD>(acc, a.dens, a.nao, a.nset, [&](int s) { return sink_of(a, s); }, oA, oB, oC, oD, w);
    }
}

}  // namespace

#define CAT2(a, b) a##b
#define CAT(a, b) CAT2(a, b)

void CAT(launch_jk_class_, AIMDHF_QCLASS)(const JKArgs& a, const int2* list, long long nq, hipStream_t st) {
    constexpr int NSPLIT = SPLIT ? Qt::NCD : 1;
    long long nt = nq * NSPLIT;
    if (nt > 0) k_jk_class<QC><<<dim3((unsigned)((nt + 63) / 64)), dim3(64), 0, st>>>(a, list, nq);
}

}  // namespace aimdhf
