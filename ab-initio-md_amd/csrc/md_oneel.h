// One-electron integrals and their nuclear derivatives for one shell pair
// (runtime l <= 2), McMurchie-Davidson as aimd/qc/integrals.py:146-501:
//   S, T, dipole          overlap-type (E^{ij}_0, kinetic via E^{i,j+2}, E^{ij}_1)
//   V                     -sum_C Z_C (2 pi/p) sum_h E_h R_h(p, P - C)
//   d/dR [W S + D T]      shifted expansions 2a E^{i+1,j} - i E^{i-1,j}
//   d/dR [D V]            basis centres + the Hellmann-Feynman term (-R_{h+e_d})
// __host__ __device__ so the check harness runs the same code on the CPU.

#pragma once
#include "md_core.h"
#include "md_pairs.h"

namespace md {

constexpr int OE_IM = 4, OE_JM = 5, OE_TM = 10;   // E buffer: i <= l+1, j <= l+2, t <= i+j+1

struct E1 {
    double e[OE_IM][OE_JM][OE_TM];
    MD_HD double operator()(int i, int j, int t) const { return e[i][j][t]; }
};

// E^{ij}_t for i <= imax, j <= jmax (t <= i+j), zero elsewhere
MD_HD void herm_e_rt(int imax, int jmax, double a, double b, double xab, double e00, E1& E) {
    double p = a + b;
    double xpa = -b * xab / p, xpb = a * xab / p, h = 0.5 / p;
    for (int i = 0; i < OE_IM; ++i)
        for (int j = 0; j < OE_JM; ++j)
            for (int t = 0; t < OE_TM; ++t) E.e[i][j][t] = 0.0;
    E.e[0][0][0] = e00;
    for (int j = 1; j <= jmax; ++j)
        for (int t = 0; t <= j; ++t) {
            double v = xpb * E.e[0][j - 1][t] + (t + 1) * E.e[0][j - 1][t + 1];
            if (t > 0) v += h * E.e[0][j - 1][t - 1];
            E.e[0][j][t] = v;
        }
    for (int i = 1; i <= imax; ++i)
        for (int j = 0; j <= jmax; ++j)
            for (int t = 0; t <= i + j; ++t) {
                double v = xpa * E.e[i - 1][j][t] + (t + 1) * E.e[i - 1][j][t + 1];
                if (t > 0) v += h * E.e[i - 1][j][t - 1];
                E.e[i][j][t] = v;
            }
}

MD_HD int cart_comp(int l, int i, int d) {
    return d == 0 ? cart_x(l, i) : (d == 1 ? cart_y(l, i) : cart_z(l, i));
}

MD_HD double kin1d(const E1& E, int i, int j, double b) {
    double v = -2.0 * b * (2 * j + 1) * E(i, j, 0) + 4.0 * b * b * E(i, j + 2, 0);
    if (j >= 2) v += j * (j - 1) * E(i, j - 2, 0);
    return -0.5 * v;
}

// R_{tuv}(alpha, x, y, z) for t+u+v <= L (runtime L <= 5), in place, into R[NH(5)]
MD_HD void hermite_r_rt(int L, double alpha, double x, double y, double z, const double* tab, double* R) {
    double T = alpha * (x * x + y * y + z * z);
    double F[6];
    F[L] = boys_top(L, T, tab);
    double ex = L > 0 ? exp(-T) : 0.0;
    for (int n = L - 1; n >= 0; --n) F[n] = (2.0 * T * F[n + 1] + ex) * (1.0 / (2 * n + 1));
    double m2a = -2.0 * alpha;
    double pw[6];
    pw[0] = 1.0;
    for (int n = 1; n <= L; ++n) pw[n] = pw[n - 1] * m2a;
    R[0] = pw[L] * F[L];
    for (int n = L - 1; n >= 0; --n) {
        for (int h = NH(L - n) - 1; h >= 1; --h) {
            int t = HT.t[h], u = HT.u[h], v = HT.v[h];
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

// S, T and dipole (origin O) block of shells si >= sj, written to both triangles
MD_HD void overlap_type_pair(const ShellView& B, int n, int si, int sj, const double* O, double* S, double* T,
                             double* M) {
    const int li = B.l[si], lj = B.l[sj], ai = B.atom[si], aj = B.atom[sj];
    const int oi = B.ao[si], oj = B.ao[sj], nci = NC(li), ncj = NC(lj);
    double s[36], t[36], m[3][36];
    for (int k = 0; k < 36; ++k) { s[k] = t[k] = m[0][k] = m[1][k] = m[2][k] = 0.0; }
    const double sq_pi = sqrt(MD_PI);
    E1 E[3];
    for (int ka = B.prim[si]; ka < B.prim[si + 1]; ++ka) {
        double a = B.ex[ka];
        for (int kb = B.prim[sj]; kb < B.prim[sj + 1]; ++kb) {
            double b = B.ex[kb];
            double p = a + b, mu = a * b / p, sq = sq_pi / sqrt(p);
            for (int d = 0; d < 3; ++d) {
                double xab = B.pos[3 * ai + d] - B.pos[3 * aj + d];
                double e00 = exp(-mu * xab * xab) * sq;
                if (d == 0) e00 *= B.coef[ka] * B.coef[kb];
                herm_e_rt(li, lj + 2, a, b, xab, e00, E[d]);
            }
            double P[3];
            for (int d = 0; d < 3; ++d) P[d] = (a * B.pos[3 * ai + d] + b * B.pos[3 * aj + d]) / p - O[d];
            for (int ia = 0; ia < nci; ++ia) {
                int ix = cart_x(li, ia), iy = cart_y(li, ia), iz = cart_z(li, ia);
                double fa = cart_factor(ix, iy, iz);
                for (int ib = 0; ib < ncj; ++ib) {
                    int jx = cart_x(lj, ib), jy = cart_y(lj, ib), jz = cart_z(lj, ib);
                    double f = fa * cart_factor(jx, jy, jz);
                    double sx = E[0](ix, jx, 0), sy = E[1](iy, jy, 0), sz = E[2](iz, jz, 0);
                    double kx = kin1d(E[0], ix, jx, b), ky = kin1d(E[1], iy, jy, b), kz = kin1d(E[2], iz, jz, b);
                    int k = ia * ncj + ib;
                    s[k] += f * sx * sy * sz;
                    t[k] += f * (kx * sy * sz + sx * ky * sz + sx * sy * kz);
                    m[0][k] += f * (E[0](ix, jx, 1) + P[0] * sx) * sy * sz;
                    m[1][k] += f * sx * (E[1](iy, jy, 1) + P[1] * sy) * sz;
                    m[2][k] += f * sx * sy * (E[2](iz, jz, 1) + P[2] * sz);
                }
            }
        }
    }
    const size_t nn = (size_t)n * n;
    for (int ia = 0; ia < nci; ++ia)
        for (int ib = 0; ib < ncj; ++ib) {
            int k = ia * ncj + ib;
            size_t u = (size_t)(oi + ia) * n + oj + ib, v = (size_t)(oj + ib) * n + oi + ia;
            S[u] = S[v] = s[k];
            T[u] = T[v] = t[k];
            for (int d = 0; d < 3; ++d) M[d * nn + u] = M[d * nn + v] = m[d][k];
        }
}

// V block of shells si >= sj over point charges (Z_c, C_c)
MD_HD void nuclear_pair(const ShellView& B, int n, int si, int sj, int nc, const double* Z, const double* C,
                        const double* tab, double* V) {
    const int li = B.l[si], lj = B.l[sj], ai = B.atom[si], aj = B.atom[sj];
    const int oi = B.ao[si], oj = B.ao[sj], nci = NC(li), ncj = NC(lj);
    const int L = li + lj, nh = NH(L);
    double v[36];
    for (int k = 0; k < 36; ++k) v[k] = 0.0;
    E1 E[3];
    double R[NH(5)], Rs[NH(4)];
    for (int ka = B.prim[si]; ka < B.prim[si + 1]; ++ka) {
        double a = B.ex[ka];
        for (int kb = B.prim[sj]; kb < B.prim[sj + 1]; ++kb) {
            double b = B.ex[kb];
            double p = a + b, mu = a * b / p;
            for (int d = 0; d < 3; ++d) {
                double xab = B.pos[3 * ai + d] - B.pos[3 * aj + d];
                double e00 = exp(-mu * xab * xab);
                if (d == 0) e00 *= B.coef[ka] * B.coef[kb];
                herm_e_rt(li, lj, a, b, xab, e00, E[d]);
            }
            double bp = b / p;
            for (int h = 0; h < nh; ++h) Rs[h] = 0.0;
            for (int c = 0; c < nc; ++c) {
                double X = (B.pos[3 * ai + 0] - C[3 * c + 0]) - bp * (B.pos[3 * ai + 0] - B.pos[3 * aj + 0]);
                double Y = (B.pos[3 * ai + 1] - C[3 * c + 1]) - bp * (B.pos[3 * ai + 1] - B.pos[3 * aj + 1]);
                double Zc = (B.pos[3 * ai + 2] - C[3 * c + 2]) - bp * (B.pos[3 * ai + 2] - B.pos[3 * aj + 2]);
                hermite_r_rt(L, p, X, Y, Zc, tab, R);
                double w = -Z[c] * 2.0 * MD_PI / p;
                for (int h = 0; h < nh; ++h) Rs[h] += w * R[h];
            }
            for (int ia = 0; ia < nci; ++ia) {
                int ix = cart_x(li, ia), iy = cart_y(li, ia), iz = cart_z(li, ia);
                double fa = cart_factor(ix, iy, iz);
                for (int ib = 0; ib < ncj; ++ib) {
                    int jx = cart_x(lj, ib), jy = cart_y(lj, ib), jz = cart_z(lj, ib);
                    double acc = 0.0;
                    for (int h = 0; h < nh; ++h)
                        acc += E[0](ix, jx, HT.t[h]) * E[1](iy, jy, HT.u[h]) * E[2](iz, jz, HT.v[h]) * Rs[h];
                    v[ia * ncj + ib] += fa * cart_factor(jx, jy, jz) * acc;
                }
            }
        }
    }
    for (int ia = 0; ia < nci; ++ia)
        for (int ib = 0; ib < ncj; ++ib) {
            V[(size_t)(oi + ia) * n + oj + ib] = v[ia * ncj + ib];
            V[(size_t)(oj + ib) * n + oi + ia] = v[ia * ncj + ib];
        }
}

// d/dR of [sum_mn Wm_mn S_mn + sum_mn D_mn T_mn] from shells si > sj on different
// atoms: g (3 values) for atom ai; atom aj gets -g (integrals.py:332-404).
MD_HD void ovlkin_grad_pair(const ShellView& B, int n, int si, int sj, const double* Wm, const double* D, double* g) {
    const int li = B.l[si], lj = B.l[sj], ai = B.atom[si], aj = B.atom[sj];
    const int oi = B.ao[si], oj = B.ao[sj], nci = NC(li), ncj = NC(lj);
    const double sq_pi = sqrt(MD_PI);
    double g0 = 0.0, g1 = 0.0, g2 = 0.0;
    E1 E[3];
    for (int ka = B.prim[si]; ka < B.prim[si + 1]; ++ka) {
        double a = B.ex[ka];
        for (int kb = B.prim[sj]; kb < B.prim[sj + 1]; ++kb) {
            double b = B.ex[kb];
            double p = a + b, mu = a * b / p, sq = sq_pi / sqrt(p);
            for (int d = 0; d < 3; ++d) {
                double xab = B.pos[3 * ai + d] - B.pos[3 * aj + d];
                double e00 = exp(-mu * xab * xab) * sq;
                if (d == 0) e00 *= B.coef[ka] * B.coef[kb];
                herm_e_rt(li + 1, lj + 2, a, b, xab, e00, E[d]);
            }
            for (int ia = 0; ia < nci; ++ia) {
                double fa = cart_factor(cart_x(li, ia), cart_y(li, ia), cart_z(li, ia));
                for (int ib = 0; ib < ncj; ++ib) {
                    double f = 2.0 * fa * cart_factor(cart_x(lj, ib), cart_y(lj, ib), cart_z(lj, ib));
                    double wmn = f * Wm[(size_t)(oi + ia) * n + oj + ib];
                    double dmn = f * D[(size_t)(oi + ia) * n + oj + ib];
                    double S1[3], K1[3], dS[3], dK[3];
                    for (int d = 0; d < 3; ++d) {
                        int i = cart_comp(li, ia, d), j = cart_comp(lj, ib, d);
                        S1[d] = E[d](i, j, 0);
                        K1[d] = kin1d(E[d], i, j, b);
                        dS[d] = 2.0 * a * E[d](i + 1, j, 0);
                        dK[d] = 2.0 * a * kin1d(E[d], i + 1, j, b);
                        if (i > 0) {
                            dS[d] -= i * E[d](i - 1, j, 0);
                            dK[d] -= i * kin1d(E[d], i - 1, j, b);
                        }
                    }
                    g0 += wmn * dS[0] * S1[1] * S1[2] + dmn * (dK[0] * S1[1] * S1[2] + dS[0] * (K1[1] * S1[2] + S1[1] * K1[2]));
                    g1 += wmn * S1[0] * dS[1] * S1[2] + dmn * (dK[1] * S1[0] * S1[2] + dS[1] * (K1[0] * S1[2] + S1[0] * K1[2]));
                    g2 += wmn * S1[0] * S1[1] * dS[2] + dmn * (dK[2] * S1[0] * S1[1] + dS[2] * (K1[0] * S1[1] + S1[0] * K1[1]));
                }
            }
        }
    }
    g[0] = g0;
    g[1] = g1;
    g[2] = g2;
}

// d/dR sum D V for shells si >= sj: calls add(atom, dir, value) for the basis
// centres and every operator centre c (integrals.py:407-501)
template <class Add>
MD_HD void nuclear_grad_pair(const ShellView& B, int n, int si, int sj, int natm, const double* Z, const double* D,
                             const double* tab, Add add) {
    const int li = B.l[si], lj = B.l[sj], ai = B.atom[si], aj = B.atom[sj];
    const int oi = B.ao[si], oj = B.ao[sj], nci = NC(li), ncj = NC(lj);
    const int L = li + lj, nh = NH(L), nh1 = NH(L + 1);
    const double sym = (si != sj) ? 2.0 : 1.0;
    E1 E[3];
    double R[NH(5)], dE[NH(5)], dA[3][NH(5)];
    for (int ka = B.prim[si]; ka < B.prim[si + 1]; ++ka) {
        double a = B.ex[ka];
        for (int kb = B.prim[sj]; kb < B.prim[sj + 1]; ++kb) {
            double b = B.ex[kb];
            double p = a + b, mu = a * b / p;
            for (int d = 0; d < 3; ++d) {
                double xab = B.pos[3 * ai + d] - B.pos[3 * aj + d];
                double e00 = exp(-mu * xab * xab);
                if (d == 0) e00 *= B.coef[ka] * B.coef[kb];
                herm_e_rt(li + 1, lj, a, b, xab, e00, E[d]);
            }
            for (int h = 0; h < nh1; ++h) dE[h] = dA[0][h] = dA[1][h] = dA[2][h] = 0.0;
            for (int ia = 0; ia < nci; ++ia) {
                int ix = cart_x(li, ia), iy = cart_y(li, ia), iz = cart_z(li, ia);
                double fa = cart_factor(ix, iy, iz);
                for (int ib = 0; ib < ncj; ++ib) {
                    int jx = cart_x(lj, ib), jy = cart_y(lj, ib), jz = cart_z(lj, ib);
                    double w = sym * fa * cart_factor(jx, jy, jz) * D[(size_t)(oi + ia) * n + oj + ib];
                    if (w == 0.0) continue;
                    for (int h = 0; h < nh1; ++h) {
                        int t = HT.t[h], u = HT.u[h], v = HT.v[h];
                        double ex = E[0](ix, jx, t), ey = E[1](iy, jy, u), ez = E[2](iz, jz, v);
                        if (h < nh) dE[h] += w * ex * ey * ez;
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
            double bp = b / p;
            for (int c = 0; c < natm; ++c) {
                double X = (B.pos[3 * ai + 0] - B.pos[3 * c + 0]) - bp * (B.pos[3 * ai + 0] - B.pos[3 * aj + 0]);
                double Y = (B.pos[3 * ai + 1] - B.pos[3 * c + 1]) - bp * (B.pos[3 * ai + 1] - B.pos[3 * aj + 1]);
                double Zc = (B.pos[3 * ai + 2] - B.pos[3 * c + 2]) - bp * (B.pos[3 * ai + 2] - B.pos[3 * aj + 2]);
                hermite_r_rt(L + 1, p, X, Y, Zc, tab, R);
                double pref = -Z[c] * 2.0 * MD_PI / p;
                for (int d = 0; d < 3; ++d) {
                    double gA = 0.0, gC = 0.0;
                    for (int h = 0; h < nh1; ++h) gA += dA[d][h] * R[h];
                    for (int h = 0; h < nh; ++h) {
                        int t = HT.t[h] + (d == 0), u = HT.u[h] + (d == 1), v = HT.v[h] + (d == 2);
                        gC -= dE[h] * R[hidx(t, u, v)];
                    }
                    gA *= pref;
                    gC *= pref;
                    add(ai, d, gA);
                    add(c, d, gC);
                    add(aj, d, -(gA + gC));
                }
            }
        }
    }
}

}  // namespace md
