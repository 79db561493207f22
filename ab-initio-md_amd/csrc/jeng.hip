// J-engine kernels for bra pair class AIMDHF_PCLASS (compiled once per pair
// class): one wave per bra pair, lanes striding over the screened prefix of
// each ket class's primitive-pair stream; the per-lane Coulomb blocks are
// summed by a fixed butterfly and written to the ket class's partial slice.

#include "aimdhf_internal.h"
#include "md_jk2.h"

#ifndef AIMDHF_PCLASS
#error "compile with -DAIMDHF_PCLASS=<0..5>"
#endif

namespace aimdhf {
using namespace md;

namespace {
constexpr int CP = AIMDHF_PCLASS;
constexpr int la_of(int pc) { return pc == 0 ? 0 : pc <= 2 ? 1 : 2; }
constexpr int lb_of(int pc) { return pc - la_of(pc) * (la_of(pc) + 1) / 2; }
constexpr int LA = la_of(CP), LB = lb_of(CP);
constexpr int NAB = NC(LA) * NC(LB);
constexpr int RSB = rec_size(LA, LB);

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
    return v;
}

template <int CP_, int CQ, int NSET>
__global__ __launch_bounds__(64) void k_jeng(JEArgs a) {
    constexpr int LC = la_of(CQ), LD = lb_of(CQ);
    constexpr int NHB = NH(LA + LB);
    const int P = a.cls_pairs[a.cls_pair_off[CP] + blockIdx.x];
    const int lane = threadIdx.x;
    const double qP = a.qscr[P];
    const double* qs = a.ket_q + a.ket_off[CQ];
    int lo = 0, hi = a.ket_off[CQ + 1] - a.ket_off[CQ];
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (qP * qs[mid] < a.thr) hi = mid;
        else lo = mid + 1;
    }
    const long long pb = a.kpre[a.ket_off[CQ]];
    const long long k1 = a.kpre[a.ket_off[CQ] + lo] - pb;
    const long long* krec = a.krec + pb;
    const double* dh = a.dh + a.dh_base[CQ];
    const double* brec = a.recs + a.rec_off[P];
    const int nb = a.npp[P];
    double Jacc[NSET * NAB];
#pragma unroll
    for (int i = 0; i < NSET * NAB; ++i) Jacc[i] = 0.0;
    for (int bp = 0; bp < nb; ++bp) {
        const double* B = brec + bp * RSB;
        double X[NSET * NHB];
#pragma unroll
        for (int i = 0; i < NSET * NHB; ++i) X[i] = 0.0;
        jeng_bra_prim<LA, LB, LC, LD, NSET>(B, a.recs, krec, dh, lane, k1, 64, a.boys, X);
        jeng_bra_transform<LA, LB, NSET>(B, X, Jacc);
    }
    double* out = a.jpart + ((size_t)CQ * a.npair + P) * (NSET * 36);
#pragma unroll
    for (int i = 0; i < NSET * NAB; ++i) {
        double v = wave_sum(Jacc[i]);
        if (lane == 0) out[(i / NAB) * 36 + i % NAB] = v;
    }
}

}  // namespace

#define CAT2(a, b) a##b
#define CAT(a, b) CAT2(a, b)

// one launch per ket class for the bra pairs of this class; ket classes without kets are skipped
void CAT(launch_jeng_, AIMDHF_PCLASS)(const JEArgs& a, int nset, const int* nket_host, hipStream_t st) {
    const int nbra = a.ncls_host[CP];
    if (nbra <= 0) return;
    static_for<0, 6>([&](auto cc) {
        constexpr int CQ = decltype(cc)::value;
        if (nket_host[CQ] <= 0) return;
        if (nset == 1) k_jeng<AIMDHF_PCLASS, CQ, 1><<<dim3(nbra), dim3(64), 0, st>>>(a);
        else k_jeng<AIMDHF_PCLASS, CQ, 2><<<dim3(nbra), dim3(64), 0, st>>>(a);
    });
}

}  // namespace aimdhf
