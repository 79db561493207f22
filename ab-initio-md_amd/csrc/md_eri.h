// Contracted ERI block of one shell quartet and its J/K digestion, templated
// on the angular momenta of the four shells (l <= 2 on gfx950 builds).
//
// (ab|cd) = sum_{bra prim pairs} sum_{ket prim pairs} 2 pi^(5/2) / (p q sqrt(p+q))
//           sum_h E^{ab}_h sum_k (-1)^{|k|} E^{cd}_k R_{h+k}(pq/(p+q), P - Q)
//
// (aimd/qc/integrals.py:630-842), in arithmetic precision Real (records are
// FP64). For every bra primitive pair the ket sum is
// contracted into X[h][cd] first; the bra Hermite-to-Cartesian transform
// then adds E^{ab}_h X[h][cd] into the contracted block. Entries of the 3-D
// Hermite tables outside the Hermite box of a Cartesian component are zero
// and are skipped at compile time.

#pragma once
#include "md_core.h"

namespace md {

template <int LA, int LB, int LC, int LD>
struct Quartet {
    static constexpr int LAB = LA + LB, LCD = LC + LD, L = LAB + LCD;
    static constexpr int NA = NC(LA), NB = NC(LB), NCc = NC(LC), ND = NC(LD);
    static constexpr int NAB = NA * NB, NCD = NCc * ND;
    static constexpr int NHB = NH(LAB), NHK = NH(LCD);
    static constexpr int RSB = rec_size(LA, LB), RSK = rec_size(LC, LD);
};

// Hermite box test: is E^{ab}_{tuv} structurally nonzero for component pair ab of (la, lb)?
constexpr bool in_box(int la, int lb, int ab, int t, int u, int v) {
    const int nb = NC(lb);
    const int ia = ab / nb, ib = ab % nb;
    return t <= cart_x(la, ia) + cart_x(lb, ib) && u <= cart_y(la, ia) + cart_y(lb, ib) &&
           v <= cart_z(la, ia) + cart_z(lb, ib);
}

// acc[ab * NCD + cd] = (ab|cd) for bra records rb[0..nb), ket records rk[0..nk)
template <int LA, int LB, int LC, int LD, bool SLICE = false, typename Real = double>
MD_HD void eri_quartet(const double* __restrict__ rb, int nb, const double* __restrict__ rk, int nk,
                       const double* __restrict__ tab, Real* __restrict__ acc) {
    using Q = Quartet<LA, LB, LC, LD>;
    constexpr int NAB = Q::NAB, NCD = Q::NCD, NHB = Q::NHB, NHK = Q::NHK, L = Q::L;
#pragma unroll
    for (int i = 0; i < NAB * NCD; ++i) acc[i] = Real(0);
    for (int ib = 0; ib < nb; ++ib) {
        const double* B = rb + ib * Q::RSB;
        const Real p = Real(B[R_P]), px = Real(B[R_PX]), py = Real(B[R_PY]), pz = Real(B[R_PZ]);
        Real X[NHB * NCD];
#pragma unroll
        for (int i = 0; i < NHB * NCD; ++i) X[i] = Real(0);
        for (int ik = 0; ik < nk; ++ik) {
            const double* K = rk + ik * Q::RSK;
            const Real q = Real(K[R_P]);
            Real alpha, pref;
            if constexpr (SLICE) {
                quartet_factors(p, q, Real(B[R_IP]), Real(K[R_IP]), alpha, pref);
            } else {
                const Real s = p + q;
                alpha = p * q / s;
                pref = Real(TWO_PI_2_5) / (p * q * md_sqrt(s));
            }
            Real R[NH(L)];
            hermite_r<L, SLICE, Real>(alpha, px - Real(K[R_PX]), py - Real(K[R_PY]), pz - Real(K[R_PZ]), pref, tab, R);
            static_for<0, NHK>([&](auto kk) {
                constexpr int k = decltype(kk)::value;
                constexpr int tk = HT.t[k], uk = HT.u[k], vk = HT.v[k];
                constexpr bool odd = (tk + uk + vk) & 1;
                static_for<0, NCD>([&](auto cc) {
                    constexpr int cd = decltype(cc)::value;
                    if constexpr (in_box(LC, LD, cd, tk, uk, vk)) {
                        const Real e = Real(odd ? -K[REC_HDR + k * NCD + cd] : K[REC_HDR + k * NCD + cd]);
                        static_for<0, NHB>([&](auto hh) {
                            constexpr int h = decltype(hh)::value;
                            constexpr int m = hidx(HT.t[h] + tk, HT.u[h] + uk, HT.v[h] + vk);
                            X[h * NCD + cd] += R[m] * e;
                        });
                    }
                });
            });
        }
        static_for<0, NHB>([&](auto hh) {
            constexpr int h = decltype(hh)::value;
            static_for<0, NAB>([&](auto aa) {
                constexpr int ab = decltype(aa)::value;
                if constexpr (in_box(LA, LB, ab, HT.t[h], HT.u[h], HT.v[h])) {
                    const Real eb = Real(B[REC_HDR + h * NAB + ab]);
#pragma unroll
                    for (int cd = 0; cd < NCD; ++cd) acc[ab * NCD + cd] += eb * X[h * NCD + cd];
                }
            });
        });
    }
}

// Register footprint (doubles) of eri_quartet; classes above the budget use
// the per-component variant below.
template <int LA, int LB, int LC, int LD>
constexpr int full_footprint() {
    using Q = Quartet<LA, LB, LC, LD>;
    return Q::NAB * Q::NCD + Q::NHB * Q::NCD + NH(Q::L);
}
template <int LA, int LB, int LC, int LD>
constexpr bool split_by_cd() { return full_footprint<LA, LB, LC, LD>() > 140; }

// acc[ab] = (ab|cd) for one ket Cartesian component pair cd (runtime); used
// with one thread per (quartet, cd) for the heavy classes.
template <int LA, int LB, int LC, int LD, bool SLICE = false, typename Real = double>
MD_HD void eri_quartet_cd(const double* __restrict__ rb, int nb, const double* __restrict__ rk, int nk,
                          const double* __restrict__ tab, int cd, Real* __restrict__ acc) {
    using Q = Quartet<LA, LB, LC, LD>;
    constexpr int NAB = Q::NAB, NCD = Q::NCD, NHB = Q::NHB, NHK = Q::NHK, L = Q::L;
#pragma unroll
    for (int i = 0; i < NAB; ++i) acc[i] = Real(0);
    for (int ib = 0; ib < nb; ++ib) {
        const double* B = rb + ib * Q::RSB;
        const Real p = Real(B[R_P]), px = Real(B[R_PX]), py = Real(B[R_PY]), pz = Real(B[R_PZ]);
        Real X[NHB];
#pragma unroll
        for (int i = 0; i < NHB; ++i) X[i] = Real(0);
        for (int ik = 0; ik < nk; ++ik) {
            const double* K = rk + ik * Q::RSK;
            const Real q = Real(K[R_P]);
            Real alpha, pref;
            if constexpr (SLICE) {
                quartet_factors(p, q, Real(B[R_IP]), Real(K[R_IP]), alpha, pref);
            } else {
                const Real s = p + q;
                alpha = p * q / s;
                pref = Real(TWO_PI_2_5) / (p * q * md_sqrt(s));
            }
            Real R[NH(L)];
            hermite_r<L, SLICE, Real>(alpha, px - Real(K[R_PX]), py - Real(K[R_PY]), pz - Real(K[R_PZ]), pref, tab, R);
            static_for<0, NHK>([&](auto kk) {
                constexpr int k = decltype(kk)::value;
                constexpr int tk = HT.t[k], uk = HT.u[k], vk = HT.v[k];
                constexpr bool odd = (tk + uk + vk) & 1;
                const Real e = Real(odd ? -K[REC_HDR + k * NCD + cd] : K[REC_HDR + k * NCD + cd]);
                static_for<0, NHB>([&](auto hh) {
                    constexpr int h = decltype(hh)::value;
                    X[h] += R[hidx(HT.t[h] + tk, HT.u[h] + uk, HT.v[h] + vk)] * e;
                });
            });
        }
        static_for<0, NHB>([&](auto hh) {
            constexpr int h = decltype(hh)::value;
            static_for<0, NAB>([&](auto aa) {
                constexpr int ab = decltype(aa)::value;
                if constexpr (in_box(LA, LB, ab, HT.t[h], HT.u[h], HT.v[h]))
                    acc[ab] += Real(B[REC_HDR + h * NAB + ab]) * X[h];
            });
        });
    }
}

}  // namespace md
