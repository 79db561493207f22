// One-electron integrals and their nuclear derivatives for one shell pair,
// McMurchie-Davidson as aimd/qc/integrals.py:146-501:
//   S, T, dipole          overlap-type (E^{ij}_0, kinetic via E^{i,j+2}, E^{ij}_1)
//   V                     -sum_C Z_C (2 pi/p) sum_h E_h R_h(p, P - C)
//   d/dR [W S + D T]      shifted expansions 2a E^{i+1,j} - i E^{i-1,j}
//   d/dR [D V]            basis centres + the Hellmann-Feynman term (-R_{h+e_d})
// Templated on the two angular momenta so every table has a compile-time
// shape (register-resident on the GPU); the runtime-l entry points at the end
// dispatch to them. The nuclear terms take a strided subset of the point
// charges (c = c0, c0 + cs, ...) so a wavefront can split the sum over its
// lanes. __host__ __device__ so the check harness runs the same code.

#pragma once
#include "md_core.h"
#include "md_pairs.h"

namespace md {

// E^{ij}_t for i <= IM, j <= JM (zero for t > i + j)
template <int IM, int JM>
struct E1T {
    double e[IM + 1][JM + 1][IM + JM + 2];
    MD_HD double operator()(int i, int j, int t) const { return e[i][j][t]; }
};

template <int IM, int JM>
MD_HD void herm_e_t(double a, double b, double xab, double e00, E1T<IM, JM>& E) {
    const double p = a + b;
    const double xpa = -b * xab / p, xpb = a * xab / p, h = 0.5 / p;
#pragma unroll
    for (int i = 0; i <= IM; ++i)
#pragma unroll
        for (int j = 0; j <= JM; ++j)
#pragma unroll
            for (int t = 0; t <= IM + JM + 1; ++t) E.e[i][j][t] = 0.0;
    E.e[0][0][0] = e00;
#pragma unroll
    for (int j = 1; j <= JM; ++j)
#pragma unroll
        for (int t = 0; t <= j; ++t) {
            double v = xpb * E.e[0][j - 1][t] + (t + 1) * E.e[0][j - 1][t + 1];
            if (t > 0) v += h * E.e[0][j - 1][t - 1];
            E.e[0][j][t] = v;
        }
#pragma unroll
    for (int i = 1; i <= IM; ++i)
#pragma unroll
        for (int j = 0; j <= JM; ++j)
#pragma unroll
            for (int t = 0; t <= i + j; ++t) {
                double v = xpa * E.e[i - 1][j][t] + (t + 1) * E.e[i - 1][j][t + 1];
                if (t > 0) v += h * E.e[i - 1][j][t - 1];
                E.e[i][j][t] = v;
            }
}

template <class E>
MD_HD double kin1d(const E& Et, int i, int j, double b) {
    double v = -2.0 * b * (2 * j + 1) * Et(i, j, 0) + 4.0 * b * b * Et(i, j + 2, 0);
    if (j >= 2) v += j * (j - 1) * Et(i, j - 2, 0);
    return -0.5 * v;
}

MD_HD constexpr int cart_comp(int l, int i, int d) {
    return d == 0 ? cart_x(l, i) : (d == 1 ? cart_y(l, i) : cart_z(l, i));
}

// R_{tuv}(alpha, x, y, z) for t+u+v <= L, in place
template <int L>
MD_HD void hermite_r_t(double alpha, double x, double y, double z, const double* tab, double* R) {
    const double T = alpha * (x * x + y * y + z * z);
    double F[L + 1];
    F[L] = boys_top(L, T, tab);
    const double ex = L > 0 ? exp(-T) : 0.0;
#pragma unroll
    for (int n = L - 1; n >= 0; --n) F[n] = (2.0 * T * F[n + 1] + ex) * (1.0 / (2 * n + 1));
    const double m2a = -2.0 * alpha;
    double pw[L + 1];
    pw[0] = 1.0;
#pragma unroll
    for (int n = 1; n <= L; ++n) pw[n] = pw[n - 1] * m2a;
    R[0] = pw[L] * F[L];
#pragma unroll
    for (int n = L - 1; n >= 0; --n) {
#pragma unroll
        for (int h = NH(L - n) - 1; h >= 1; --h) {
            const int t = HT.t[h], u = HT.u[h], v = HT.v[h];
            double r;
            if (t > 0) {
                r = x * R[hidx(t - 1, u, v)];
                if (t > 1) r += (t - 1) * R[hidx(t - 2, u, v)];
            } else if (u > 0) {
                r = y * R[hidx(0, u - 1, v)];
                if (u > 1) r += (u - 1) * R[hidx(0, u - 2, v)];
            } else {
                r = z * R[hidx(0, 0, v - 1)];
                if (v > 1) r += (v - 1) * R[hidx(0, 0, v - 2)];
            }
            R[h] = r;
        }
        R[0] = pw[n] * F[n];
    }
}

struct ShellView {
    const int *atom, *l, *prim, *ao;
    const double *ex, *coef, *pos;
};

// S, T and dipole (origin O) of shells si (l = LI), sj (l = LJ) for primitive pairs
// k = k0, k0 + ks, ... (row-major over si's and sj's primitives), accumulated into
// s, t, m[3] (NC(LI) x NC(LJ) each)
template <int LI, int LJ>
MD_HD void overlap_type_acc(const ShellView& B, int si, int sj, const double* O, int k0, int ks, double* s, double* t,
                            double* m) {
    constexpr int NI = NC(LI), NJ = NC(LJ);
    const int ai = B.atom[si], aj = B.atom[sj];
    const int pi0 = B.prim[si], npj = B.prim[sj + 1] - B.prim[sj], npp = (B.prim[si + 1] - pi0) * npj;
    const double sq_pi = sqrt(MD_PI);
    for (int k = k0; k < npp; k += ks) {
        const int ka = pi0 + k / npj, kb = B.prim[sj] + k % npj;
        const double a = B.ex[ka], b = B.ex[kb];
        const double p = a + b, mu = a * b / p, sq = sq_pi / sqrt(p);
        E1T<LI, LJ + 2> E[3];
        double P[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const double xab = B.pos[3 * ai + d] - B.pos[3 * aj + d];
            double e00 = exp(-mu * xab * xab) * sq;
            if (d == 0) e00 *= B.coef[ka] * B.coef[kb];
            herm_e_t<LI, LJ + 2>(a, b, xab, e00, E[d]);
            P[d] = (a * B.pos[3 * ai + d] + b * B.pos[3 * aj + d]) / p - O[d];
        }
#pragma unroll
        for (int ia = 0; ia < NI; ++ia) {
            const int ix = cart_x(LI, ia), iy = cart_y(LI, ia), iz = cart_z(LI, ia);
            const double fa = cart_factor(ix, iy, iz);
#pragma unroll
            for (int ib = 0; ib < NJ; ++ib) {
                const int jx = cart_x(LJ, ib), jy = cart_y(LJ, ib), jz = cart_z(LJ, ib);
                const double f = fa * cart_factor(jx, jy, jz);
                const double sx = E[0](ix, jx, 0), sy = E[1](iy, jy, 0), sz = E[2](iz, jz, 0);
                const double kx = kin1d(E[0], ix, jx, b), ky = kin1d(E[1], iy, jy, b), kz = kin1d(E[2], iz, jz, b);
                const int q = ia * NJ + ib;
                s[q] += f * sx * sy * sz;
                t[q] += f * (kx * sy * sz + sx * ky * sz + sx * sy * kz);
                m[0 * NI * NJ + q] += f * (E[0](ix, jx, 1) + P[0] * sx) * sy * sz;
                m[1 * NI * NJ + q] += f * sx * (E[1](iy, jy, 1) + P[1] * sy) * sz;
                m[2 * NI * NJ + q] += f * sx * sy * (E[2](iz, jz, 1) + P[2] * sz);
            }
        }
    }
}

// V block contributions of point charges c = c0, c0 + cs, ... < nc into v (NC(LI) x NC(LJ))
template <int LI, int LJ>
MD_HD void nuclear_acc(const ShellView& B, int si, int sj, int c0, int cs, int nc, const double* Z, const double* C,
                       const double* tab, double* v) {
    constexpr int NI = NC(LI), NJ = NC(LJ), L = LI + LJ, NHL = NH(L);
    const int ai = B.atom[si], aj = B.atom[sj];
    for (int ka = B.prim[si]; ka < B.prim[si + 1]; ++ka) {
        const double a = B.ex[ka];
        for (int kb = B.prim[sj]; kb < B.prim[sj + 1]; ++kb) {
            const double b = B.ex[kb];
            const double p = a + b, mu = a * b / p;
            E1T<LI, LJ> E[3];
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                const double xab = B.pos[3 * ai + d] - B.pos[3 * aj + d];
                double e00 = exp(-mu * xab * xab);
                if (d == 0) e00 *= B.coef[ka] * B.coef[kb];
                herm_e_t<LI, LJ>(a, b, xab, e00, E[d]);
            }
            const double bp = b / p;
            double Rs[NHL];
#pragma unroll
            for (int h = 0; h < NHL; ++h) Rs[h] = 0.0;
            for (int c = c0; c < nc; c += cs) {
                const double X = (B.pos[3 * ai + 0] - C[3 * c + 0]) - bp * (B.pos[3 * ai + 0] - B.pos[3 * aj + 0]);
                const double Y = (B.pos[3 * ai + 1] - C[3 * c + 1]) - bp * (B.pos[3 * ai + 1] - B.pos[3 * aj + 1]);
                const double Zc = (B.pos[3 * ai + 2] - C[3 * c + 2]) - bp * (B.pos[3 * ai + 2] - B.pos[3 * aj + 2]);
                double R[NHL];
                hermite_r_t<L>(p, X, Y, Zc, tab, R);
                const double w = -Z[c] * 2.0 * MD_PI / p;
#pragma unroll
                for (int h = 0; h < NHL; ++h) Rs[h] += w * R[h];
            }
#pragma unroll
            for (int ia = 0; ia < NI; ++ia) {
                const int ix = cart_x(LI, ia), iy = cart_y(LI, ia), iz = cart_z(LI, ia);
                const double fa = cart_factor(ix, iy, iz);
#pragma unroll
                for (int ib = 0; ib < NJ; ++ib) {
                    const int jx = cart_x(LJ, ib), jy = cart_y(LJ, ib), jz = cart_z(LJ, ib);
                    double acc = 0.0;
#pragma unroll
                    for (int h = 0; h < NHL; ++h)
                        if (HT.t[h] <= ix + jx && HT.u[h] <= iy + jy && HT.v[h] <= iz + jz)
                            acc += E[0](ix, jx, HT.t[h]) * E[1](iy, jy, HT.u[h]) * E[2](iz, jz, HT.v[h]) * Rs[h];
                    v[ia * NJ + ib] += fa * cart_factor(jx, jy, jz) * acc;
                }
            }
        }
    }
}

// d/dR_{atom(si)} of [sum Wm S + sum D T] over the primitive pairs k = k0, k0 + ks, ...
// of shells si, sj on different atoms (atom(sj) gets -g) (integrals.py:332-404)
template <int LI, int LJ>
MD_HD void ovlkin_grad_acc(const ShellView& B, int n, int si, int sj, const double* Wm, const double* D, int k0, int ks,
                           double* g) {
    constexpr int NI = NC(LI), NJ = NC(LJ);
    const int ai = B.atom[si], aj = B.atom[sj];
    const int oi = B.ao[si], oj = B.ao[sj];
    const int pi0 = B.prim[si], npj = B.prim[sj + 1] - B.prim[sj], npp = (B.prim[si + 1] - pi0) * npj;
    const double sq_pi = sqrt(MD_PI);
    double wmn[NI * NJ], dmn[NI * NJ];
#pragma unroll
    for (int ia = 0; ia < NI; ++ia)
#pragma unroll
        for (int ib = 0; ib < NJ; ++ib) {
            const double f = 2.0 * cart_factor(cart_x(LI, ia), cart_y(LI, ia), cart_z(LI, ia)) *
                             cart_factor(cart_x(LJ, ib), cart_y(LJ, ib), cart_z(LJ, ib));
            wmn[ia * NJ + ib] = f * Wm[(size_t)(oi + ia) * n + oj + ib];
            dmn[ia * NJ + ib] = f * D[(size_t)(oi + ia) * n + oj + ib];
        }
    for (int k = k0; k < npp; k += ks) {
        const int ka = pi0 + k / npj, kb = B.prim[sj] + k % npj;
        const double a = B.ex[ka], b = B.ex[kb];
        const double p = a + b, mu = a * b / p, sq = sq_pi / sqrt(p);
        E1T<LI + 1, LJ + 2> E[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const double xab = B.pos[3 * ai + d] - B.pos[3 * aj + d];
            double e00 = exp(-mu * xab * xab) * sq;
            if (d == 0) e00 *= B.coef[ka] * B.coef[kb];
            herm_e_t<LI + 1, LJ + 2>(a, b, xab, e00, E[d]);
        }
#pragma unroll
        for (int ia = 0; ia < NI; ++ia)
#pragma unroll
            for (int ib = 0; ib < NJ; ++ib) {
                double S1[3], K1[3], dS[3], dK[3];
#pragma unroll
                for (int d = 0; d < 3; ++d) {
                    const int i = cart_comp(LI, ia, d), j = cart_comp(LJ, ib, d);
                    S1[d] = E[d](i, j, 0);
                    K1[d] = kin1d(E[d], i, j, b);
                    dS[d] = 2.0 * a * E[d](i + 1, j, 0);
                    dK[d] = 2.0 * a * kin1d(E[d], i + 1, j, b);
                    if (i > 0) {
                        dS[d] -= i * E[d](i - 1, j, 0);
                        dK[d] -= i * kin1d(E[d], i - 1, j, b);
                    }
                }
                const double w = wmn[ia * NJ + ib], dm = dmn[ia * NJ + ib];
                g[0] += w * dS[0] * S1[1] * S1[2] + dm * (dK[0] * S1[1] * S1[2] + dS[0] * (K1[1] * S1[2] + S1[1] * K1[2]));
                g[1] += w * S1[0] * dS[1] * S1[2] + dm * (dK[1] * S1[0] * S1[2] + dS[1] * (K1[0] * S1[2] + S1[0] * K1[2]));
                g[2] += w * S1[0] * S1[1] * dS[2] + dm * (dK[2] * S1[0] * S1[1] + dS[2] * (K1[0] * S1[1] + S1[0] * K1[1]));
            }
    }
}

// d/dR sum D V of shells si, sj for operator centres c = c0, c0 + cs, ... < natm
// (integrals.py:407-501): add(c, d, x) for each centre, gA / gB += the basis-centre
// terms of atom(si) / atom(sj)
template <int LI, int LJ, class Add>
MD_HD void nuclear_grad_acc(const ShellView& B, int n, int si, int sj, int c0, int cs, int natm, const double* Z,
                            const double* D, const double* tab, Add add, double* gA, double* gB) {
    constexpr int NI = NC(LI), NJ = NC(LJ), L = LI + LJ, NHL = NH(L), NH1 = NH(L + 1);
    const int ai = B.atom[si], aj = B.atom[sj];
    const int oi = B.ao[si], oj = B.ao[sj];
    const double sym = (si != sj) ? 2.0 : 1.0;
    double wd[NI * NJ];
#pragma unroll
    for (int ia = 0; ia < NI; ++ia)
#pragma unroll
        for (int ib = 0; ib < NJ; ++ib)
            wd[ia * NJ + ib] = sym * cart_factor(cart_x(LI, ia), cart_y(LI, ia), cart_z(LI, ia)) *
                               cart_factor(cart_x(LJ, ib), cart_y(LJ, ib), cart_z(LJ, ib)) *
                               D[(size_t)(oi + ia) * n + oj + ib];
    for (int c = c0; c < natm; c += cs) {
        double gc[3] = {0.0, 0.0, 0.0};
        for (int ka = B.prim[si]; ka < B.prim[si + 1]; ++ka) {
            const double a = B.ex[ka];
            for (int kb = B.prim[sj]; kb < B.prim[sj + 1]; ++kb) {
                const double b = B.ex[kb];
                const double p = a + b, mu = a * b / p;
                E1T<LI + 1, LJ> E[3];
#pragma unroll
                for (int d = 0; d < 3; ++d) {
                    const double xab = B.pos[3 * ai + d] - B.pos[3 * aj + d];
                    double e00 = exp(-mu * xab * xab);
                    if (d == 0) e00 *= B.coef[ka] * B.coef[kb];
                    herm_e_t<LI + 1, LJ>(a, b, xab, e00, E[d]);
                }
                double dE[NHL], dA[3][NH1];
#pragma unroll
                for (int h = 0; h < NH1; ++h) dA[0][h] = dA[1][h] = dA[2][h] = 0.0;
#pragma unroll
                for (int h = 0; h < NHL; ++h) dE[h] = 0.0;
#pragma unroll
                for (int ia = 0; ia < NI; ++ia) {
                    const int ix = cart_x(LI, ia), iy = cart_y(LI, ia), iz = cart_z(LI, ia);
#pragma unroll
                    for (int ib = 0; ib < NJ; ++ib) {
                        const int jx = cart_x(LJ, ib), jy = cart_y(LJ, ib), jz = cart_z(LJ, ib);
                        const double w = wd[ia * NJ + ib];
#pragma unroll
                        for (int h = 0; h < NH1; ++h) {
                            const int t = HT.t[h], u = HT.u[h], v = HT.v[h];
                            if (t > ix + jx + 1 || u > iy + jy + 1 || v > iz + jz + 1) continue;
                            const double ex = E[0](ix, jx, t), ey = E[1](iy, jy, u), ez = E[2](iz, jz, v);
                            if (h < NHL) dE[h] += w * ex * ey * ez;
                            double dx = 2.0 * a * E[0](ix + 1, jx, t);
                            if (ix > 0) dx -= ix * E[0](ix - 1, jx, t);
                            double dy = 2.0 * a * E[1](iy + 1, jy, u);
                            if (iy > 0) dy -= iy * E[1](iy - 1, jy, u);
                            double dz = 2.0 * a * E[2](iz + 1, jz, v);
                            if (iz > 0) dz -= iz * E[2](iz - 1, jz, v);
                            dA[0][h] += w * dx * ey * ez;
                            dA[1][h] += w * ex * dy * ez;
                            dA[2][h] += w * ex * ey * dz;
                        }
                    }
                }
                const double bp = b / p;
                const double X = (B.pos[3 * ai + 0] - B.pos[3 * c + 0]) - bp * (B.pos[3 * ai + 0] - B.pos[3 * aj + 0]);
                const double Y = (B.pos[3 * ai + 1] - B.pos[3 * c + 1]) - bp * (B.pos[3 * ai + 1] - B.pos[3 * aj + 1]);
                const double Zc = (B.pos[3 * ai + 2] - B.pos[3 * c + 2]) - bp * (B.pos[3 * ai + 2] - B.pos[3 * aj + 2]);
                double R[NH1];
                hermite_r_t<L + 1>(p, X, Y, Zc, tab, R);
                const double pref = -Z[c] * 2.0 * MD_PI / p;
#pragma unroll
                for (int d = 0; d < 3; ++d) {
                    double ga = 0.0, gcd = 0.0;
#pragma unroll
                    for (int h = 0; h < NH1; ++h) ga += dA[d][h] * R[h];
#pragma unroll
                    for (int h = 0; h < NHL; ++h)
                        gcd -= dE[h] * R[hidx(HT.t[h] + (d == 0), HT.u[h] + (d == 1), HT.v[h] + (d == 2))];
                    ga *= pref;
                    gcd *= pref;
                    gA[d] += ga;
                    gc[d] += gcd;
                    gB[d] -= ga + gcd;
                }
            }
        }
#pragma unroll
        for (int d = 0; d < 3; ++d) add(c, d, gc[d]);
    }
}

// ------------------------------------------------------------ runtime l
template <class F>
MD_HD void with_ls(int li, int lj, F&& f) {
    static_for<0, 3>([&](auto ii) {
        static_for<0, 3>([&](auto jj) {
            if (li == decltype(ii)::value && lj == decltype(jj)::value) f(ii, jj);
        });
    });
}

// S, T and dipole block of shells si, sj, written to both triangles
MD_HD void overlap_type_pair(const ShellView& B, int n, int si, int sj, const double* O, double* S, double* T,
                             double* M) {
    with_ls(B.l[si], B.l[sj], [&](auto ii, auto jj) {
        constexpr int LI = decltype(ii)::value, LJ = decltype(jj)::value, NI = NC(LI), NJ = NC(LJ);
        double s[NI * NJ] = {}, t[NI * NJ] = {}, m[3 * NI * NJ] = {};
        overlap_type_acc<LI, LJ>(B, si, sj, O, 0, 1, s, t, m);
        const size_t nn = (size_t)n * n;
        const int oi = B.ao[si], oj = B.ao[sj];
        for (int q = 0; q < NI * NJ; ++q) {
            const size_t u = (size_t)(oi + q / NJ) * n + oj + q % NJ, v = (size_t)(oj + q % NJ) * n + oi + q / NJ;
            S[u] = S[v] = s[q];
            T[u] = T[v] = t[q];
            for (int d = 0; d < 3; ++d) M[d * nn + u] = M[d * nn + v] = m[d * NI * NJ + q];
        }
    });
}

// V block of shells si, sj over point charges (Z_c, C_c), both triangles
MD_HD void nuclear_pair(const ShellView& B, int n, int si, int sj, int nc, const double* Z, const double* C,
                        const double* tab, double* V) {
    with_ls(B.l[si], B.l[sj], [&](auto ii, auto jj) {
        constexpr int LI = decltype(ii)::value, LJ = decltype(jj)::value, NI = NC(LI), NJ = NC(LJ);
        double v[NI * NJ] = {};
        nuclear_acc<LI, LJ>(B, si, sj, 0, 1, nc, Z, C, tab, v);
        const int oi = B.ao[si], oj = B.ao[sj];
        for (int q = 0; q < NI * NJ; ++q) {
            V[(size_t)(oi + q / NJ) * n + oj + q % NJ] = v[q];
            V[(size_t)(oj + q % NJ) * n + oi + q / NJ] = v[q];
        }
    });
}

// d/dR of [sum Wm S + sum D T] of shells si != sj on different atoms: g for atom(si)
MD_HD void ovlkin_grad_pair(const ShellView& B, int n, int si, int sj, const double* Wm, const double* D, double* g) {
    g[0] = g[1] = g[2] = 0.0;
    with_ls(B.l[si], B.l[sj], [&](auto ii, auto jj) {
        ovlkin_grad_acc<decltype(ii)::value, decltype(jj)::value>(B, n, si, sj, Wm, D, 0, 1, g);
    });
}

// d/dR sum D V of shells si, sj: add(atom, dir, value) for every centre
template <class Add>
MD_HD void nuclear_grad_pair(const ShellView& B, int n, int si, int sj, int natm, const double* Z, const double* D,
                             const double* tab, Add add) {
    with_ls(B.l[si], B.l[sj], [&](auto ii, auto jj) {
        double gA[3] = {0.0, 0.0, 0.0}, gB[3] = {0.0, 0.0, 0.0};
        nuclear_grad_acc<decltype(ii)::value, decltype(jj)::value>(B, n, si, sj, 0, 1, natm, Z, D, tab, add, gA, gB);
        for (int d = 0; d < 3; ++d) {
            add(B.atom[si], d, gA[d]);
            add(B.atom[sj], d, gB[d]);
        }
    });
}

}  // namespace md
