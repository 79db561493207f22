// J/K digestion of one contracted quartet block, shared by the gfx950
// kernels (atomic fixed-point sink) and the host check harness.
#pragma once
#include "md_eri.h"

namespace md {

// Digest a full quartet block acc[ab*NCD+cd] into J' and K' (J = J'+J'^T, K = K'+K'^T):
//   J'_ab += 2w sum_cd (ab|cd) D_cd     J'_cd += 2w sum_ab (ab|cd) D_ab
//   K'_ac += w sum (ab|cd) D_bd         K'_bc += w sum (ab|cd) D_ad
//   K'_ad += w sum (ab|cd) D_bc         K'_bd += w sum (ab|cd) D_ac
// w = 1/2 for each of A=B, C=D, P=Q (permutational multiplicity of the quartet).
template <int LA, int LB, int LC, int LD, class SinkFn>
MD_HD void digest_full(const double* acc, const double* dens, int nao, int nset, SinkFn sink_of,
                        int oA, int oB, int oC, int oD, double w) {
    using Qt = Quartet<LA, LB, LC, LD>;
    constexpr int NA = Qt::NA, NB = Qt::NB, NCc = Qt::NCc, ND = Qt::ND, NCD = Qt::NCD;
    const size_t n = nao;
    for (int s = 0; s < nset; ++s) {
        const double* D = dens + s * n * n;
        const auto S = sink_of(s);
        double Dab[NA * NB], Dcd[NCc * ND], Dac[NA * NCc], Dad[NA * ND], Dbc[NB * NCc], Dbd[NB * ND];
#pragma unroll
        for (int i = 0; i < NA; ++i) {
#pragma unroll
            for (int j = 0; j < NB; ++j) Dab[i * NB + j] = D[(oA + i) * n + oB + j];
#pragma unroll
            for (int c = 0; c < NCc; ++c) Dac[i * NCc + c] = D[(oA + i) * n + oC + c];
#pragma unroll
            for (int d = 0; d < ND; ++d) Dad[i * ND + d] = D[(oA + i) * n + oD + d];
        }
#pragma unroll
        for (int j = 0; j < NB; ++j) {
#pragma unroll
            for (int c = 0; c < NCc; ++c) Dbc[j * NCc + c] = D[(oB + j) * n + oC + c];
#pragma unroll
            for (int d = 0; d < ND; ++d) Dbd[j * ND + d] = D[(oB + j) * n + oD + d];
        }
#pragma unroll
        for (int c = 0; c < NCc; ++c)
#pragma unroll
            for (int d = 0; d < ND; ++d) Dcd[c * ND + d] = D[(oC + c) * n + oD + d];
#pragma unroll
        for (int ab = 0; ab < NA * NB; ++ab) {
            double t = 0.0;
#pragma unroll
            for (int cd = 0; cd < NCD; ++cd) t += acc[ab * NCD + cd] * Dcd[cd];
            S.j((oA + ab / NB) * n + oB + ab % NB, 2.0 * w * t);
        }
#pragma unroll
        for (int cd = 0; cd < NCD; ++cd) {
            double t = 0.0;
#pragma unroll
            for (int ab = 0; ab < NA * NB; ++ab) t += acc[ab * NCD + cd] * Dab[ab];
            S.j((oC + cd / ND) * n + oD + cd % ND, 2.0 * w * t);
        }
#pragma unroll
        for (int i = 0; i < NA; ++i)
#pragma unroll
            for (int c = 0; c < NCc; ++c) {
                double t = 0.0;
#pragma unroll
                for (int j = 0; j < NB; ++j)
#pragma unroll
                    for (int d = 0; d < ND; ++d) t += acc[(i * NB + j) * NCD + c * ND + d] * Dbd[j * ND + d];
                S.k((oA + i) * n + oC + c, w * t);
            }
#pragma unroll
        for (int j = 0; j < NB; ++j)
#pragma unroll
            for (int c = 0; c < NCc; ++c) {
                double t = 0.0;
#pragma unroll
                for (int i = 0; i < NA; ++i)
#pragma unroll
                    for (int d = 0; d < ND; ++d) t += acc[(i * NB + j) * NCD + c * ND + d] * Dad[i * ND + d];
                S.k((oB + j) * n + oC + c, w * t);
            }
#pragma unroll
        for (int i = 0; i < NA; ++i)
#pragma unroll
            for (int d = 0; d < ND; ++d) {
                double t = 0.0;
#pragma unroll
                for (int j = 0; j < NB; ++j)
#pragma unroll
                    for (int c = 0; c < NCc; ++c) t += acc[(i * NB + j) * NCD + c * ND + d] * Dbc[j * NCc + c];
                S.k((oA + i) * n + oD + d, w * t);
            }
#pragma unroll
        for (int j = 0; j < NB; ++j)
#pragma unroll
            for (int d = 0; d < ND; ++d) {
                double t = 0.0;
#pragma unroll
                for (int i = 0; i < NA; ++i)
#pragma unroll
                    for (int c = 0; c < NCc; ++c) t += acc[(i * NB + j) * NCD + c * ND + d] * Dac[i * NCc + c];
                S.k((oB + j) * n + oD + d, w * t);
            }
    }
}

// Same for one ket component pair cd = (c, d): acc[ab] = (ab|cd)
template <int LA, int LB, int LC, int LD, class SinkFn>
MD_HD void digest_cd(const double* acc, const double* dens, int nao, int nset, SinkFn sink_of,
                      int oA, int oB, int oC, int oD, int cd, double w) {
    using Qt = Quartet<LA, LB, LC, LD>;
    constexpr int NA = Qt::NA, NB = Qt::NB, ND = Qt::ND;
    const size_t n = nao;
    const int c = cd / ND, d = cd % ND;
    for (int s = 0; s < nset; ++s) {
        const double* D = dens + s * n * n;
        const auto S = sink_of(s);
        const double Dcd = D[(oC + c) * n + oD + d];
        double tcd = 0.0;
#pragma unroll
        for (int i = 0; i < NA; ++i)
#pragma unroll
            for (int j = 0; j < NB; ++j) {
                S.j((oA + i) * n + oB + j, 2.0 * w * acc[i * NB + j] * Dcd);
                tcd += acc[i * NB + j] * D[(oA + i) * n + oB + j];
            }
        S.j((oC + c) * n + oD + d, 2.0 * w * tcd);
#pragma unroll
        for (int i = 0; i < NA; ++i) {
            double tac = 0.0, tad = 0.0;
#pragma unroll
            for (int j = 0; j < NB; ++j) {
                tac += acc[i * NB + j] * D[(oB + j) * n + oD + d];
                tad += acc[i * NB + j] * D[(oB + j) * n + oC + c];
            }
            S.k((oA + i) * n + oC + c, w * tac);
            S.k((oA + i) * n + oD + d, w * tad);
        }
#pragma unroll
        for (int j = 0; j < NB; ++j) {
            double tbc = 0.0, tbd = 0.0;
#pragma unroll
            for (int i = 0; i < NA; ++i) {
                tbc += acc[i * NB + j] * D[(oA + i) * n + oD + d];
                tbd += acc[i * NB + j] * D[(oA + i) * n + oC + c];
            }
            S.k((oB + j) * n + oC + c, w * tbc);
            S.k((oB + j) * n + oD + d, w * tbd);
        }
    }
}

}  // namespace md
