// Coulomb from the integral store for bra pairs of class AIMDHF_PCLASS
// (compiled once per class): one wave per pair P, lanes over its gather list
// (every partner Q with q_P q_Q >= threshold, built once per geometry in the
// order of the q-sorted partners of each class). J_P[ab] = sum_Q w_Q
// sum_cd (ab|cd) D_cd, w_Q = 2 for C != D; the per-lane blocks are summed by
// a fixed butterfly, so J is bitwise reproducible run to run.

#include <type_traits>

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
constexpr int qcls(int c1, int c2) { return c1 > c2 ? c1 * (c1 + 1) / 2 + c2 : c2 * (c2 + 1) / 2 + c1; }

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
    return v;
}

template <int CQ, bool SW, int NSET>
__device__ __forceinline__ void j_item(const JSArgs& a, const double* __restrict__ S, int cs, int oC, int oD, double w,
                                       double* Jacc) {
    constexpr int LC = la_of(CQ), LD = lb_of(CQ);
    constexpr int NDc = NC(LD), NCD = NC(LC) * NDc;
    const size_t n = a.nao, nn = n * n;
    constexpr int BA = SW ? LC : LA, BB = SW ? LD : LB, KA = SW ? LA : LC, KB = SW ? LB : LD;
    if constexpr (store_split<BA, BB, KA, KB>()) {
        constexpr int NBRA = NC(BA) * NC(BB), NKET = NC(KA) * NC(KB);
        double v[NBRA];
#pragma unroll
        for (int i = 0; i < NBRA; ++i) v[i] = S[i * NKET + cs];
        if constexpr (!SW) {
            // column cs = cd of the partner
            const int c = cs / NDc, d = cs % NDc;
#pragma unroll
            for (int s = 0; s < NSET; ++s) {
                const double dv = w * a.dens[s * nn + (size_t)(oC + c) * n + oD + d];
#pragma unroll
                for (int ab = 0; ab < NAB; ++ab) Jacc[s * NAB + ab] += v[ab] * dv;
            }
        } else {
            // column cs = ab of the owner; rows over the partner's cd
#pragma unroll
            for (int s = 0; s < NSET; ++s) {
                double t = 0.0;
#pragma unroll
                for (int cd = 0; cd < NCD; ++cd)
                    t += v[cd] * (w * a.dens[s * nn + (size_t)(oC + cd / NDc) * n + oD + cd % NDc]);
#pragma unroll
                for (int ab = 0; ab < NAB; ++ab)
                    if (ab == cs) Jacc[s * NAB + ab] += t;
            }
        }
    } else {
        double blk[NAB * NCD];
#pragma unroll
        for (int i = 0; i < NAB * NCD; ++i) blk[i] = S[i];
        double dq[NSET * NCD];
#pragma unroll
        for (int s = 0; s < NSET; ++s)
#pragma unroll
            for (int cd = 0; cd < NCD; ++cd)
                dq[s * NCD + cd] = w * a.dens[s * nn + (size_t)(oC + cd / NDc) * n + oD + cd % NDc];
        j_store_full<LA, LB, LC, LD, SW, NSET>(blk, dq, Jacc);
    }
}

template <int CP_, int NSET>
__global__ __launch_bounds__(64) void k_jstore(JSArgs a) {
    const int P = a.cls_pairs[a.cls_off + blockIdx.x];
    const int lane = threadIdx.x;
    double Jacc[NSET * NAB];
#pragma unroll
    for (int i = 0; i < NSET * NAB; ++i) Jacc[i] = 0.0;
    static_for<0, 6>([&](auto qq) {
        constexpr int CQ = decltype(qq)::value;
        const long long e0 = a.jseg[(size_t)P * 6 + CQ], e1 = a.jseg[(size_t)P * 6 + CQ + 1];
        if (e0 == e1) return;
        constexpr int k = qcls(CP, CQ);
        constexpr int LC = la_of(CQ), LD = lb_of(CQ);
        constexpr int BA = CP >= CQ ? LA : LC, BB = CP >= CQ ? LB : LD;
        constexpr int KA = CP >= CQ ? LC : LA, KB = CP >= CQ ? LD : LB;
        constexpr int BLK = NC(BA) * NC(BB) * NC(KA) * NC(KB);
        constexpr int NS = store_split<BA, BB, KA, KB>() ? NC(KA) * NC(KB) : 1;
        const double* base = a.store + a.store_base[k];
        const long long nt = (e1 - e0) * NS;
        for (long long t = lane; t < nt; t += 64) {
            const unsigned long long ent = (unsigned long long)a.jlist[e0 + t / NS];
            const int cs = (int)(t % NS);
            const long long qi = (long long)(ent & 0xffffffffull);
            const int oC = (int)((ent >> 32) & 0x3fff), oD = (int)((ent >> 46) & 0x3fff);
            const bool sw = (ent >> 60) & 1;
            const double w = ((ent >> 61) & 1) ? 2.0 : 1.0;
            const double* S = base + qi * BLK;
            if constexpr (CP == CQ) {
                if (sw) j_item<CQ, true, NSET>(a, S, cs, oC, oD, w, Jacc);
                else j_item<CQ, false, NSET>(a, S, cs, oC, oD, w, Jacc);
            } else {
                j_item<CQ, (CP < CQ), NSET>(a, S, cs, oC, oD, w, Jacc);
            }
        }
    });
    const int A = a.pair_a[P], B = a.pair_b[P];
    const int oA = a.shell_ao[A], oB = a.shell_ao[B];
    const size_t n = a.nao, nn = n * n;
    constexpr int NBc = NC(LB);
#pragma unroll
    for (int i = 0; i < NSET * NAB; ++i) {
        const double v = wave_sum(Jacc[i]);
        const int s = i / NAB, x = (i % NAB) / NBc, y = i % NBc;
        if (lane == 0) {
            a.J[s * nn + (size_t)(oA + x) * n + oB + y] = v;
            if (A != B) a.J[s * nn + (size_t)(oB + y) * n + oA + x] = v;
        }
    }
}

}  // namespace

#define CAT2(a, b) a##b
#define CAT(a, b) CAT2(a, b)

void CAT(launch_jstore_, AIMDHF_PCLASS)(const JSArgs& a, int npairs, hipStream_t st) {
    if (npairs <= 0) return;
    if (a.nset == 1) k_jstore<CP, 1><<<dim3(npairs), dim3(64), 0, st>>>(a);
    else k_jstore<CP, 2><<<dim3(npairs), dim3(64), 0, st>>>(a);
}

}  // namespace aimdhf
