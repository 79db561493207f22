// Derivative-ERI contraction of one shell quartet with its two-particle
// density Gamma (aimd/qc/integrals.py:1049-1369), recomputing the derivative
// integrals on the fly.
//
// A "pass" differentiates with respect to the two centres of the bra pair:
//   d(ab|cd)/dA_x = sum_prim [2a (a+1_x b|cd) - a_x (a-1_x b|cd)],
// i.e. the bra Hermite expansion is replaced by the shifted expansion
// 2a E^{a_x+1,b_x}_t - a_x E^{a_x-1,b_x}_t (t <= a_x+b_x+1), contracted with
// the ket-summed Hermite integrals X[h] = sum_k (-1)^|k| E^{cd}_k R_{h+k} of
// degree up to l_A+l_B+1. The ket centres come from a second pass with the
// pair roles swapped, or from translational invariance when a pair sits on
// one atom. A pass covers NCOL consecutive Cartesian component pairs
// kc0 .. kc0+NCOL-1 of its ket, with gam[ab * NCOL + j] = Gamma[ab][kc0 + j]:
// the ket-summed Hermite integrals X[h][j] are built once per bra primitive
// pair for all NCOL columns and then contracted with Gamma, so R is evaluated
// once per primitive quartet and item.

#pragma once
#include "md_eri.h"
#include "md_pairs.h"

namespace md {

// columns per pass item: the largest divisor of the ket's component count
// keeping the X block and the R table within 112 doubles
template <int LA, int LB, int LC, int LD>
MD_HD constexpr int grad_ncol() {
    constexpr int nkc = NC(LC) * NC(LD), nh = NH(LA + LB + 1), nr = NH(LA + LB + LC + LD + 1);
    int best = 1;
    for (int c = 1; c <= nkc; ++c)
        if (nkc % c == 0 && nh * c + nr <= 112) best = c;
    return best;
}

// Ket columns per folded item (grad_pass_c) of class (LA LB | LC LD) in FP64 or
// FP32, 0 where the class is not folded: kets (p|s) with a light bra ((ps|,
// (pp|, (ds|) take all three columns in one item, other (s|s) and (p|s) kets one
// item per column; (p|p) and (d|s) kets under a bra up to (dp| one item per
// column. Heavier classes keep the separate bra and ket passes: their folded
// items need far more registers than one wave has and multiply the compiled code
// by the ket's column count. The FP64 items of (pp|pp) disagree with the oracle
// on gfx950 (1.3e-3 Eh/bohr on (H2O)8) while the same code agrees in FP32 there,
// in FP64 for (ds|pp) and (dp|pp), and in FP64 on the host; (pp|pp) stays
// unfolded in FP64.
template <int LA, int LB, int LC, int LD, bool FP64>
MD_HD constexpr int fold_ncol() {
    if (LC == 0 && LD == 0) return 1;
    if (LC == 1 && LD == 0) return 3 * (NH(LA + LB) + NH(LA + LB + 1)) <= 90 ? 3 : 1;
    if (LC + LD == 2 && LA + LB <= 3) return FP64 && LA == 1 && LC == 1 ? 0 : 1;
    return 0;
}

template <int LA, int LB, int LC, int LD, int NCOL, bool SLICE = false, typename Real = double>
MD_HD void grad_pass(const double* __restrict__ rb, int nb, const double* __restrict__ rk, int nk,
                     const double* __restrict__ tab, const double* Apos, const double* Bpos, int kc0,
                     const Real* gam, double* gA, double* gB) {
    using Q = Quartet<LA, LB, LC, LD>;
    constexpr int NAB = Q::NAB, NKC = Q::NCD, NHK = Q::NHK;
    constexpr int NHB1 = NH(Q::LAB + 1), L1 = Q::L + 1;
    constexpr int NA = Q::NA, NB = Q::NB;
    constexpr int IS = LA + 2, JS0 = LB + 1, TS0 = LA + LB + 2;
    Real xab[3];
    for (int d = 0; d < 3; ++d) { gA[d] = 0.0; gB[d] = 0.0; xab[d] = Real(Apos[d] - Bpos[d]); }
    Real fcart[NAB];
    static_for<0, NAB>([&](auto aa) {
        constexpr int ab = decltype(aa)::value;
        constexpr int ia = ab / NB, ib = ab % NB;
        fcart[ab] = Real(cart_factor(cart_x(LA, ia), cart_y(LA, ia), cart_z(LA, ia)) *
                         cart_factor(cart_x(LB, ib), cart_y(LB, ib), cart_z(LB, ib)));
    });
    for (int ibp = 0; ibp < nb; ++ibp) {
        const double* Bp = rb + ibp * Q::RSB;
        const Real p = Real(Bp[R_P]), a = Real(Bp[R_A]), b = Real(Bp[R_B]), cab = Real(Bp[R_C]);
        const Real px = Real(Bp[R_PX]), py = Real(Bp[R_PY]), pz = Real(Bp[R_PZ]);
        Real X[NHB1 * NCOL];
#pragma unroll
        for (int i = 0; i < NHB1 * NCOL; ++i) X[i] = Real(0);
        for (int ik = 0; ik < nk; ++ik) {
            const double* K = rk + ik * Q::RSK;
            const Real q = Real(K[R_P]);
            Real alpha, pref;
            if constexpr (SLICE) {
                quartet_factors(p, q, Real(Bp[R_IP]), Real(K[R_IP]), alpha, pref);
            } else {
                const Real s = p + q;
                alpha = p * q / s;
                pref = Real(TWO_PI_2_5) / (p * q * md_sqrt(s));
            }
            Real R[NH(L1)];
            hermite_r<L1, SLICE, Real>(alpha, px - Real(K[R_PX]), py - Real(K[R_PY]), pz - Real(K[R_PZ]), pref, tab, R);
            static_for<0, NHK>([&](auto kk) {
                constexpr int k = decltype(kk)::value;
                constexpr int tk = HT.t[k], uk = HT.u[k], vk = HT.v[k];
                constexpr bool odd = (tk + uk + vk) & 1;
#pragma unroll
                for (int j = 0; j < NCOL; ++j) {
                    const Real e = odd ? -Real(K[REC_HDR + k * NKC + kc0 + j]) : Real(K[REC_HDR + k * NKC + kc0 + j]);
                    static_for<0, NHB1>([&](auto hh) {
                        constexpr int h = decltype(hh)::value;
                        X[h * NCOL + j] += R[hidx(HT.t[h] + tk, HT.u[h] + uk, HT.v[h] + vk)] * e;
                    });
                }
            });
        }
        const Real mu = a * b / p;
        // E^{ij}_t for i <= l_A + 1, j <= l_B: the A derivative needs the row
        // above; the B derivative follows from translational invariance of the
        // pair distribution, (d/dA + d/dB) sum_t E_t L_t = sum_t E_t L_{t+1},
        // i.e. the plain coefficients against the block shifted by one degree
        Real ex[IS * JS0 * TS0], ey[IS * JS0 * TS0], ez[IS * JS0 * TS0];
        hermite_e<LA + 1, LB, Real>(a, b, xab[0], md_exp(-mu * xab[0] * xab[0]), ex);
        hermite_e<LA + 1, LB, Real>(a, b, xab[1], md_exp(-mu * xab[1] * xab[1]), ey);
        hermite_e<LA + 1, LB, Real>(a, b, xab[2], md_exp(-mu * xab[2] * xab[2]), ez);
        const Real ta = Real(2.0) * a;
        Real sA[3] = {0, 0, 0}, sP[3] = {0, 0, 0};
        static_for<0, NAB>([&](auto aa) {
            constexpr int ab = decltype(aa)::value;
            constexpr int ia = ab / NB, ib = ab % NB;
            constexpr int ax = cart_x(LA, ia), ay = cart_y(LA, ia), az = cart_z(LA, ia);
            constexpr int bx = cart_x(LB, ib), by = cart_y(LB, ib), bz = cart_z(LB, ib);
            const Real f = fcart[ab] * cab;
            auto E = [](const Real* e, int i, int j, int tt) { return e[(i * JS0 + j) * TS0 + tt]; };
            auto yv = [&](int h) {
                Real xg = 0;
#pragma unroll
                for (int j = 0; j < NCOL; ++j) xg += gam[ab * NCOL + j] * X[h * NCOL + j];
                return f * xg;
            };
            static_for<0, NHB1>([&](auto hh) {
                constexpr int h = decltype(hh)::value;
                constexpr int t = HT.t[h], u = HT.u[h], v = HT.v[h];
                constexpr int ox = t > ax + bx, oy = u > ay + by, oz = v > az + bz;
                // A derivative: one index may exceed the pair's degree by one
                if constexpr (ox + oy + oz <= 1 && t <= ax + bx + 1 && u <= ay + by + 1 && v <= az + bz + 1) {
                    const Real y = yv(h);
                    const Real e0x = E(ex, ax, bx, t), e0y = E(ey, ay, by, u), e0z = E(ez, az, bz, v);
                    Real dax = ta * E(ex, ax + 1, bx, t), day = ta * E(ey, ay + 1, by, u), daz = ta * E(ez, az + 1, bz, v);
                    if constexpr (ax > 0) dax -= Real(ax) * E(ex, ax - 1, bx, t);
                    if constexpr (ay > 0) day -= Real(ay) * E(ey, ay - 1, by, u);
                    if constexpr (az > 0) daz -= Real(az) * E(ez, az - 1, bz, v);
                    sA[0] += y * dax * e0y * e0z;
                    sA[1] += y * e0x * day * e0z;
                    sA[2] += y * e0x * e0y * daz;
                }
                // A + B derivative: coefficient E_{tuv} against the block at h + e_d
                if constexpr (t <= ax + bx && u <= ay + by && v <= az + bz) {
                    const Real e = E(ex, ax, bx, t) * E(ey, ay, by, u) * E(ez, az, bz, v);
                    sP[0] += e * yv(hidx(t + 1, u, v));
                    sP[1] += e * yv(hidx(t, u + 1, v));
                    sP[2] += e * yv(hidx(t, u, v + 1));
                }
            });
        });
        Real sB[3] = {sP[0] - sA[0], sP[1] - sA[1], sP[2] - sA[2]};
        for (int d = 0; d < 3; ++d) { gA[d] += double(sA[d]); gB[d] += double(sB[d]); }
    }
}

// The bra pass of a quartet with the derivative on the ket's first centre C
// folded in: one R per primitive quartet serves the A, B and C derivatives (D
// follows from translational invariance). For ket component (c, d),
// d/dC_x phi_c = 2 gamma (x - C_x) phi_c - c_x phi_{c-1x}, and both functions'
// products with phi_d expand about Q from the stored table E_k of (c, d):
//   raised:  U_k = E_{k-1x} / (2q) + X_QC E_k + (k_x + 1) E_{k+1x}
//   lowered: L_k = E^{(c-1x) d}_k, from the same recursion run downwards along x,
//            L_{top} = 2q E_{top+1},  L_{s-1} = 2q (E_s - X_QC L_s - (s+1) L_{s+1}),
// and (ab| d_C(cd)) = sum_t E^{ab}_t sum_k (-1)^|k| (2 gamma U_k - c_x L_k) R_{t+k}.
// Gamma is contracted with the stored bra expansion first (Z[t][j] per bra
// primitive pair), so the C part costs one pass over t per ket term. The item
// covers ket columns KC0 .. KC0+NCOL-1 (their share of the C derivative); gA, gB
// as grad_pass for the same columns; with need_c false (C and D on one atom) the
// C part is skipped and gC is zero.
template <int LA, int LB, int LC, int LD, int NCOL, int KC0, bool SLICE = false, typename Real = double>
MD_HD void grad_pass_c(const double* __restrict__ rb, int nb, const double* __restrict__ rk, int nk,
                       const double* __restrict__ tab, const double* Apos, const double* Bpos, const double* Cpos,
                       bool need_c, const Real* gam, double* gA, double* gB, double* gC) {
    using Q = Quartet<LA, LB, LC, LD>;
    constexpr int NAB = Q::NAB, NHB = Q::NHB, NKC = Q::NCD, NHK = Q::NHK, ND = Q::ND;
    constexpr int NHB1 = NH(Q::LAB + 1), L1 = Q::L + 1;
    constexpr int NA = Q::NA, NB = Q::NB;
    constexpr int JS0 = LB + 1, TS0 = LA + LB + 2;
    Real xab[3];
    for (int d = 0; d < 3; ++d) {
        gA[d] = 0.0;
        gB[d] = 0.0;
        gC[d] = 0.0;
        xab[d] = Real(Apos[d] - Bpos[d]);
    }
    Real fcart[NAB];
    static_for<0, NAB>([&](auto aa) {
        constexpr int ab = decltype(aa)::value;
        constexpr int ia = ab / NB, ib = ab % NB;
        fcart[ab] = Real(cart_factor(cart_x(LA, ia), cart_y(LA, ia), cart_z(LA, ia)) *
                         cart_factor(cart_x(LB, ib), cart_y(LB, ib), cart_z(LB, ib)));
    });
    for (int ibp = 0; ibp < nb; ++ibp) {
        const double* Bp = rb + ibp * Q::RSB;
        const Real p = Real(Bp[R_P]), a = Real(Bp[R_A]), b = Real(Bp[R_B]), cab = Real(Bp[R_C]);
        const Real px = Real(Bp[R_PX]), py = Real(Bp[R_PY]), pz = Real(Bp[R_PZ]);
        // Z[t][j] = sum_ab Gamma[ab][j] E^{ab}_t (stored bra expansion)
        Real Z[NHB * NCOL];
        if (need_c) static_for<0, NHB>([&](auto hh) {
            constexpr int h = decltype(hh)::value;
#pragma unroll
            for (int j = 0; j < NCOL; ++j) Z[h * NCOL + j] = Real(0);
            static_for<0, NAB>([&](auto aa) {
                constexpr int ab = decltype(aa)::value;
                if constexpr (in_box(LA, LB, ab, HT.t[h], HT.u[h], HT.v[h])) {
                    const Real e = Real(Bp[REC_HDR + h * NAB + ab]);
#pragma unroll
                    for (int j = 0; j < NCOL; ++j) Z[h * NCOL + j] += gam[ab * NCOL + j] * e;
                }
            });
        });
        Real X[NHB1 * NCOL];
#pragma unroll
        for (int i = 0; i < NHB1 * NCOL; ++i) X[i] = Real(0);
        Real sC[3] = {0, 0, 0};
        for (int ik = 0; ik < nk; ++ik) {
            const double* K = rk + ik * Q::RSK;
            const Real q = Real(K[R_P]);
            Real alpha, pref;
            if constexpr (SLICE) {
                quartet_factors(p, q, Real(Bp[R_IP]), Real(K[R_IP]), alpha, pref);
            } else {
                const Real s = p + q;
                alpha = p * q / s;
                pref = Real(TWO_PI_2_5) / (p * q * md_sqrt(s));
            }
            Real R[NH(L1)];
            hermite_r<L1, SLICE, Real>(alpha, px - Real(K[R_PX]), py - Real(K[R_PY]), pz - Real(K[R_PZ]), pref, tab, R);
            // A, B: the ket-summed block X of grad_pass
            static_for<0, NHK>([&](auto kk) {
                constexpr int k = decltype(kk)::value;
                constexpr int tk = HT.t[k], uk = HT.u[k], vk = HT.v[k];
                constexpr bool odd = (tk + uk + vk) & 1;
#pragma unroll
                for (int j = 0; j < NCOL; ++j) {
                    const Real e = odd ? -Real(K[REC_HDR + k * NKC + KC0 + j]) : Real(K[REC_HDR + k * NKC + KC0 + j]);
                    static_for<0, NHB1>([&](auto hh) {
                        constexpr int h = decltype(hh)::value;
                        X[h * NCOL + j] += R[hidx(HT.t[h] + tk, HT.u[h] + uk, HT.v[h] + vk)] * e;
                    });
                }
            });
            // C: raised and lowered ket expansions of each column, line by line along
            // the derivative's direction, against Z and R
            if (!need_c) continue;
            const Real g2 = Real(2.0) * Real(K[R_A]), h2q = Real(0.5) * Real(K[R_IP]), q2 = Real(2.0) * q;
            const Real qc[3] = {Real(K[R_PX]) - Real(Cpos[0]), Real(K[R_PY]) - Real(Cpos[1]),
                                Real(K[R_PZ]) - Real(Cpos[2])};
            static_for<0, NCOL>([&](auto jj) {
                constexpr int j = decltype(jj)::value, cj = KC0 + j, ci = cj / ND, di = cj % ND;
                constexpr int c3[3] = {cart_x(LC, ci), cart_y(LC, ci), cart_z(LC, ci)};
                constexpr int b3[3] = {c3[0] + cart_x(LD, di), c3[1] + cart_y(LD, di), c3[2] + cart_z(LD, di)};
                static_for<0, 3>([&](auto dd) {
                    constexpr int d = decltype(dd)::value;
                    constexpr int o1 = d == 0 ? 1 : 0, o2 = d == 2 ? 1 : 2;
                    constexpr int bd = b3[d], cd = c3[d];
                    Real acc = 0;
                    static_for<0, b3[o1] + 1>([&](auto ii1) {
                        static_for<0, b3[o2] + 1>([&](auto ii2) {
                            constexpr int i1 = decltype(ii1)::value, i2 = decltype(ii2)::value;
                            // the line's stored coefficients E_s, s = 0 .. bd
                            Real E[bd + 1];
                            static_for<0, bd + 1>([&](auto ss) {
                                constexpr int s = decltype(ss)::value;
                                constexpr int t = d == 0 ? s : i1, u = d == 1 ? s : (d == 0 ? i1 : i2),
                                              v = d == 2 ? s : i2;
                                E[s] = Real(K[REC_HDR + hidx(t, u, v) * NKC + cj]);
                            });
                            Real Lw[bd + 2];
#pragma unroll
                            for (int s = 0; s < bd + 2; ++s) Lw[s] = Real(0);
                            if constexpr (cd > 0) {
                                Lw[bd - 1] = q2 * E[bd];
#pragma unroll
                                for (int s = bd - 1; s >= 1; --s)
                                    Lw[s - 1] = q2 * (E[s] - qc[d] * Lw[s] - Real(s + 1) * Lw[s + 1]);
                            }
                            static_for<0, bd + 2>([&](auto ss) {
                                constexpr int s = decltype(ss)::value;
                                constexpr int t = d == 0 ? s : i1, u = d == 1 ? s : (d == 0 ? i1 : i2),
                                              v = d == 2 ? s : i2;
                                Real w = 0;
                                if constexpr (s <= bd) w += qc[d] * E[s];
                                if constexpr (s + 1 <= bd) w += Real(s + 1) * E[s + 1];
                                if constexpr (s >= 1) w += h2q * E[s - 1];
                                w *= g2;
                                if constexpr (cd > 0 && s < bd) w -= Real(cd) * Lw[s];
                                if constexpr ((t + u + v) & 1) w = -w;
                                Real zr = 0;
                                static_for<0, NHB>([&](auto hh) {
                                    constexpr int h = decltype(hh)::value;
                                    zr += Z[h * NCOL + j] * R[hidx(HT.t[h] + t, HT.u[h] + u, HT.v[h] + v)];
                                });
                                acc += w * zr;
                            });
                        });
                    });
                    sC[d] += acc;
                });
            });
        }
        for (int d = 0; d < 3; ++d) gC[d] += double(sC[d]);
        // A and B: as grad_pass
        const Real mu = a * b / p;
        Real ex[(LA + 2) * JS0 * TS0], ey[(LA + 2) * JS0 * TS0], ez[(LA + 2) * JS0 * TS0];
        hermite_e<LA + 1, LB, Real>(a, b, xab[0], md_exp(-mu * xab[0] * xab[0]), ex);
        hermite_e<LA + 1, LB, Real>(a, b, xab[1], md_exp(-mu * xab[1] * xab[1]), ey);
        hermite_e<LA + 1, LB, Real>(a, b, xab[2], md_exp(-mu * xab[2] * xab[2]), ez);
        const Real ta = Real(2.0) * a;
        Real sA[3] = {0, 0, 0}, sP[3] = {0, 0, 0};
        static_for<0, NAB>([&](auto aa) {
            constexpr int ab = decltype(aa)::value;
            constexpr int ia = ab / NB, ib = ab % NB;
            constexpr int ax = cart_x(LA, ia), ay = cart_y(LA, ia), az = cart_z(LA, ia);
            constexpr int bx = cart_x(LB, ib), by = cart_y(LB, ib), bz = cart_z(LB, ib);
            const Real f = fcart[ab] * cab;
            auto E = [](const Real* e, int i, int j, int tt) { return e[(i * JS0 + j) * TS0 + tt]; };
            auto yv = [&](int h) {
                Real xg = 0;
#pragma unroll
                for (int j = 0; j < NCOL; ++j) xg += gam[ab * NCOL + j] * X[h * NCOL + j];
                return f * xg;
            };
            static_for<0, NHB1>([&](auto hh) {
                constexpr int h = decltype(hh)::value;
                constexpr int t = HT.t[h], u = HT.u[h], v = HT.v[h];
                constexpr int ox = t > ax + bx, oy = u > ay + by, oz = v > az + bz;
                if constexpr (ox + oy + oz <= 1 && t <= ax + bx + 1 && u <= ay + by + 1 && v <= az + bz + 1) {
                    const Real y = yv(h);
                    const Real e0x = E(ex, ax, bx, t), e0y = E(ey, ay, by, u), e0z = E(ez, az, bz, v);
                    Real dax = ta * E(ex, ax + 1, bx, t), day = ta * E(ey, ay + 1, by, u), daz = ta * E(ez, az + 1, bz, v);
                    if constexpr (ax > 0) dax -= Real(ax) * E(ex, ax - 1, bx, t);
                    if constexpr (ay > 0) day -= Real(ay) * E(ey, ay - 1, by, u);
                    if constexpr (az > 0) daz -= Real(az) * E(ez, az - 1, bz, v);
                    sA[0] += y * dax * e0y * e0z;
                    sA[1] += y * e0x * day * e0z;
                    sA[2] += y * e0x * e0y * daz;
                }
                if constexpr (t <= ax + bx && u <= ay + by && v <= az + bz) {
                    const Real e = E(ex, ax, bx, t) * E(ey, ay, by, u) * E(ez, az, bz, v);
                    sP[0] += e * yv(hidx(t + 1, u, v));
                    sP[1] += e * yv(hidx(t, u + 1, v));
                    sP[2] += e * yv(hidx(t, u, v + 1));
                }
            });
        });
        for (int d = 0; d < 3; ++d) {
            gA[d] += double(sA[d]);
            gB[d] += double(sP[d] - sA[d]);
        }
    }
}

// single ket component kc, gam[ab] = Gamma[ab][kc]
template <int LA, int LB, int LC, int LD>
MD_HD void grad_pass_kc(const double* __restrict__ rb, int nb, const double* __restrict__ rk, int nk,
                        const double* __restrict__ tab, const double* Apos, const double* Bpos, int kc,
                        const double* gam, double* gA, double* gB) {
    grad_pass<LA, LB, LC, LD, 1>(rb, nb, rk, nk, tab, Apos, Bpos, kc, gam, gA, gB);
}

}  // namespace md
