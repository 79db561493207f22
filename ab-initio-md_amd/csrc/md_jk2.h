// Owner-computes J and K (no atomics, fixed summation order).
//
// J ("J-engine"): for a bra pair P = (A, B) the Coulomb block is
//   J_ab = sum_bp sum_h E^{ab}_h(bp) sum_Q sum_kp sum_k R_{h+k}(bp, kp) Dh_k(kp),
//   Dh_k(kp) = (-1)^{|k|} w_Q sum_cd E^{cd}_k(kp) D_cd,   w_Q = 2 for C != D, 1 for C = D,
// i.e. the density is carried into Hermite space once per iteration and no
// Cartesian ket integrals are formed (McMurchie-Davidson's Coulomb shortcut).
// The ket sum runs over every pair Q with q_P q_Q >= threshold; kets are
// stored sorted by q, so that set is a prefix of each class's stream.
//
// K: for an output block (X, Y) the exchange block is
//   K_xy = sum_{J, L} sum_{j in J, l in L} (xj|yl) D_jl,
// evaluated from the contracted quartet (XJ|YL) for every partner pair with
// q_XJ q_YL >= threshold (the same screened set of integrals as the stored
// tensor of the reference, each unique quartet recomputed for each of the
// four exchange blocks it feeds).

#pragma once
#include "md_eri.h"

namespace md {

// X[s][h] contributions of ket primitive pairs i = k0, k0+ks, ... < k1 for one bra primitive pair
template <int LA, int LB, int LC, int LD, int NSET>
MD_HD void jeng_bra_prim(const double* __restrict__ B, const double* __restrict__ recs,
                         const long long* __restrict__ krec, const double* __restrict__ dh, long long k0,
                         long long k1, long long ks, const double* __restrict__ tab, double* X) {
    using Q = Quartet<LA, LB, LC, LD>;
    constexpr int NHB = Q::NHB, NHK = Q::NHK, L = Q::L;
    const double p = B[R_P], px = B[R_PX], py = B[R_PY], pz = B[R_PZ];
    for (long long i = k0; i < k1; i += ks) {
        const double* K = recs + krec[i];
        const double q = K[R_P];
        const double s = p + q;
        const double alpha = p * q / s;
        const double pref = TWO_PI_2_5 / (p * q * sqrt(s));
        double R[NH(L)];
        hermite_r<L>(alpha, px - K[R_PX], py - K[R_PY], pz - K[R_PZ], pref, tab, R);
        const double* d = dh + i * (NSET * NHK);
        static_for<0, NSET>([&](auto ss) {
            constexpr int st = decltype(ss)::value;
            static_for<0, NHK>([&](auto kk) {
                constexpr int k = decltype(kk)::value;
                const double e = d[st * NHK + k];
                static_for<0, NHB>([&](auto hh) {
                    constexpr int h = decltype(hh)::value;
                    X[st * NHB + h] += R[hidx(HT.t[h] + HT.t[k], HT.u[h] + HT.u[k], HT.v[h] + HT.v[k])] * e;
                });
            });
        });
    }
}

// Jacc[s][ab] += sum_h E^{ab}_h X[s][h] for one bra primitive pair record B
template <int LA, int LB, int NSET>
MD_HD void jeng_bra_transform(const double* __restrict__ B, const double* X, double* Jacc) {
    constexpr int NAB = NC(LA) * NC(LB), NHB = NH(LA + LB);
    static_for<0, NHB>([&](auto hh) {
        constexpr int h = decltype(hh)::value;
        static_for<0, NAB>([&](auto aa) {
            constexpr int ab = decltype(aa)::value;
            if constexpr (in_box(LA, LB, ab, HT.t[h], HT.u[h], HT.v[h])) {
                const double eb = B[REC_HDR + h * NAB + ab];
#pragma unroll
                for (int s = 0; s < NSET; ++s) Jacc[s * NAB + ab] += eb * X[s * NHB + h];
            }
        });
    });
}

// Dh_k of one ket primitive pair record K (pair class LC >= LD) for density D (row stride n)
template <int LC, int LD>
MD_HD void density_hermite(const double* __restrict__ K, const double* __restrict__ D, size_t n, int oC, int oD,
                           double w, double* out) {
    constexpr int NCc = NC(LC), ND = NC(LD), NCD = NCc * ND, NHK = NH(LC + LD);
    static_for<0, NHK>([&](auto kk) {
        constexpr int k = decltype(kk)::value;
        constexpr bool odd = (HT.t[k] + HT.u[k] + HT.v[k]) & 1;
        double acc = 0.0;
        static_for<0, NCD>([&](auto cc) {
            constexpr int cd = decltype(cc)::value;
            if constexpr (in_box(LC, LD, cd, HT.t[k], HT.u[k], HT.v[k]))
                acc += K[REC_HDR + k * NCD + cd] * D[(size_t)(oC + cd / ND) * n + oD + cd % ND];
        });
        out[k] = odd ? -w * acc : w * acc;
    });
}

// K_acc[s][x*NY + y] += sum_{j,l} (xj|yl) D_s[j][l] from a full block acc[ab*NCD+cd]
// of the canonical records; FB / FK: the bra / ket record is stored as (J, X) / (L, Y).
template <int LX, int LJ, int LY, int LL, bool FB, bool FK>
MD_HD void k_digest_full(const double* acc, const double* __restrict__ D, size_t n, size_t nn, int nset, int oJ,
                         int oL, double* Kacc) {
    constexpr int NX = NC(LX), NJ = NC(LJ), NY = NC(LY), NL = NC(LL);
    constexpr int NB = FB ? NX : NJ;     // second index of the bra record
    constexpr int ND = FK ? NY : NL;     // second index of the ket record
    constexpr int NCD = NY * NL;
    for (int s = 0; s < nset; ++s) {
        const double* Ds = D + s * nn;
        double Djl[NJ * NL];
#pragma unroll
        for (int j = 0; j < NJ; ++j)
#pragma unroll
            for (int l = 0; l < NL; ++l) Djl[j * NL + l] = Ds[(size_t)(oJ + j) * n + oL + l];
#pragma unroll
        for (int x = 0; x < NX; ++x)
#pragma unroll
            for (int y = 0; y < NY; ++y) {
                double t = 0.0;
#pragma unroll
                for (int j = 0; j < NJ; ++j)
#pragma unroll
                    for (int l = 0; l < NL; ++l) {
                        const int ab = FB ? (j * NB + x) : (x * NB + j);
                        const int cd = FK ? (l * ND + y) : (y * ND + l);
                        t += acc[ab * NCD + cd] * Djl[j * NL + l];
                    }
                Kacc[s * NX * NY + x * NY + y] += t;
            }
    }
}

// Same for one ket component pair cd of the canonical ket record: acc[ab] = (ab|cd)
template <int LX, int LJ, int LY, int LL, bool FB, bool FK>
MD_HD void k_digest_cd(const double* acc, int cd, const double* __restrict__ D, size_t n, size_t nn, int nset, int oJ,
                       int oL, double* Kacc) {
    constexpr int NX = NC(LX), NJ = NC(LJ), NY = NC(LY), NL = NC(LL);
    constexpr int NB = FB ? NX : NJ;
    constexpr int ND = FK ? NY : NL;
    const int c = cd / ND, d = cd % ND;
    const int y = FK ? d : c, l = FK ? c : d;
    for (int s = 0; s < nset; ++s) {
        const double* Ds = D + s * nn;
#pragma unroll
        for (int x = 0; x < NX; ++x) {
            double t = 0.0;
#pragma unroll
            for (int j = 0; j < NJ; ++j) {
                const int ab = FB ? (j * NB + x) : (x * NB + j);
                t += acc[ab] * Ds[(size_t)(oJ + j) * n + oL + l];
            }
#pragma unroll
            for (int yy = 0; yy < NY; ++yy)
                if (yy == y) Kacc[s * NX * NY + x * NY + yy] += t;
        }
    }
}

// ---------------------------------------------------------------------------
// Digestion from the per-geometry integral store. A stored block holds the
// contracted quartet with the heavier pair (ties: the higher pair index) as
// bra, S[ab * NCD + cd], ab / cd over the bra / ket pair records (l_first >=
// l_second, ties: higher shell index first).

// component index of (x, j) in the record of pair (X, J); F: record stored as (J, X)
template <int NX, int NJ, bool F>
MD_HD constexpr int rec_idx(int x, int j) { return F ? j * NX + x : x * NJ + j; }

// K_acc[s][x*NY+y] += sum_{j,l} (xj|yl) D_s[j][l] for a whole stored block S;
// SW: the (Y, L) record is the stored bra. Djl[s][j*NL+l] is the density block.
template <int LX, int LJ, int LY, int LL, bool FB, bool FK, bool SW, int NSET>
MD_HD void k_store_full(const double* S, const double* Djl, double* Kacc) {
    constexpr int NX = NC(LX), NJ = NC(LJ), NY = NC(LY), NL = NC(LL);
    constexpr int NXJ = NX * NJ, NYL = NY * NL;
#pragma unroll
    for (int s = 0; s < NSET; ++s)
#pragma unroll
        for (int x = 0; x < NX; ++x)
#pragma unroll
            for (int y = 0; y < NY; ++y) {
                double t = 0.0;
#pragma unroll
                for (int j = 0; j < NJ; ++j)
#pragma unroll
                    for (int l = 0; l < NL; ++l) {
                        const int i1 = rec_idx<NX, NJ, FB>(x, j), i2 = rec_idx<NY, NL, FK>(y, l);
                        t += S[SW ? i2 * NXJ + i1 : i1 * NYL + i2] * Djl[s * NJ * NL + j * NL + l];
                    }
                Kacc[s * NX * NY + x * NY + y] += t;
            }
}

// Same for one stored-ket column cs of S (values v[i] = S[i * NCD + cs] over the stored bra)
template <int LX, int LJ, int LY, int LL, bool FB, bool FK, bool SW, int NSET>
MD_HD void k_store_col(const double* v, int cs, const double* __restrict__ D, size_t n, size_t nn, int oJ, int oL,
                       double* Kacc) {
    constexpr int NX = NC(LX), NJ = NC(LJ), NY = NC(LY), NL = NC(LL);
    if constexpr (!SW) {
        // column = (y, l) of the (Y, L) record; rows over (x, j)
        const int y = FK ? cs % NY : cs / NL, l = FK ? cs / NY : cs % NL;
#pragma unroll
        for (int s = 0; s < NSET; ++s) {
            double dj[NJ];
#pragma unroll
            for (int j = 0; j < NJ; ++j) dj[j] = D[s * nn + (size_t)(oJ + j) * n + oL + l];
#pragma unroll
            for (int x = 0; x < NX; ++x) {
                double t = 0.0;
#pragma unroll
                for (int j = 0; j < NJ; ++j) t += v[rec_idx<NX, NJ, FB>(x, j)] * dj[j];
#pragma unroll
                for (int yy = 0; yy < NY; ++yy)
                    if (yy == y) Kacc[s * NX * NY + x * NY + yy] += t;
            }
        }
    } else {
        // column = (x, j) of the (X, J) record; rows over (y, l)
        const int x = FB ? cs % NX : cs / NJ, j = FB ? cs / NX : cs % NJ;
#pragma unroll
        for (int s = 0; s < NSET; ++s) {
            double dl[NL];
#pragma unroll
            for (int l = 0; l < NL; ++l) dl[l] = D[s * nn + (size_t)(oJ + j) * n + oL + l];
#pragma unroll
            for (int y = 0; y < NY; ++y) {
                double t = 0.0;
#pragma unroll
                for (int l = 0; l < NL; ++l) t += v[rec_idx<NY, NL, FK>(y, l)] * dl[l];
#pragma unroll
                for (int xx = 0; xx < NX; ++xx)
                    if (xx == x) Kacc[s * NX * NY + xx * NY + y] += t;
            }
        }
    }
}

// Jacc[s][ab] += sum_cd (ab|cd) dq[s][cd] for a whole stored block; (LA, LB) is the
// owner pair's record, (LC, LD) the partner's; SW: the partner is the stored bra.
template <int LA, int LB, int LC, int LD, bool SW, int NSET>
MD_HD void j_store_full(const double* S, const double* dq, double* Jacc) {
    constexpr int NAB = NC(LA) * NC(LB), NCD = NC(LC) * NC(LD);
#pragma unroll
    for (int s = 0; s < NSET; ++s)
#pragma unroll
        for (int ab = 0; ab < NAB; ++ab) {
            double t = 0.0;
#pragma unroll
            for (int cd = 0; cd < NCD; ++cd) t += S[SW ? cd * NAB + ab : ab * NCD + cd] * dq[s * NCD + cd];
            Jacc[s * NAB + ab] += t;
        }
}

// Same for one stored-ket column cs (values v[i] = S[i * NCD_stored + cs])
template <int LA, int LB, int LC, int LD, bool SW, int NSET>
MD_HD void j_store_col(const double* v, int cs, const double* dq, double* Jacc) {
    constexpr int NAB = NC(LA) * NC(LB), NCD = NC(LC) * NC(LD);
    if constexpr (!SW) {
        // column = cd of the partner; rows over the owner's ab
#pragma unroll
        for (int s = 0; s < NSET; ++s) {
            double d = 0.0;
#pragma unroll
            for (int cd = 0; cd < NCD; ++cd)
                if (cd == cs) d = dq[s * NCD + cd];
#pragma unroll
            for (int ab = 0; ab < NAB; ++ab) Jacc[s * NAB + ab] += v[ab] * d;
        }
    } else {
        // column = ab of the owner; rows over the partner's cd
#pragma unroll
        for (int s = 0; s < NSET; ++s) {
            double t = 0.0;
#pragma unroll
            for (int cd = 0; cd < NCD; ++cd) t += v[cd] * dq[s * NCD + cd];
#pragma unroll
            for (int ab = 0; ab < NAB; ++ab)
                if (ab == cs) Jacc[s * NAB + ab] += t;
        }
    }
}

// block size and split rule of the digestion of stored class (LA LB | LC LD)
template <int LA, int LB, int LC, int LD>
MD_HD constexpr bool store_split() { return NC(LA) * NC(LB) * NC(LC) * NC(LD) > 36; }

}  // namespace md
