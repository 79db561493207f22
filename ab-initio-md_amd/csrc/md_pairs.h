// Primitive-pair records (see md_core.h for the layout): Cartesian
// normalisation factors and the 3-D Hermite table of one primitive pair.
#pragma once
#include "md_core.h"

namespace md {

MD_HD long dfact(int n) {
    long r = 1;
    while (n > 1) {
        r *= n;
        n -= 2;
    }
    return r;
}

MD_HD double cart_factor(int lx, int ly, int lz) {
    int l = lx + ly + lz;
    return sqrt((double)dfact(2 * l - 1) / (double)(dfact(2 * lx - 1) * dfact(2 * ly - 1) * dfact(2 * lz - 1)));
}

template <int LA, int LB>
MD_HD void write_record(double a, double b, double ca, double cb, const double* A, const double* B, double* rec) {
    constexpr int TS = LA + LB + 1;
    constexpr int NA = NC(LA), NB = NC(LB), NAB = NA * NB;
    double p = a + b;
    double mu = a * b / p;
    double e[3][(LA + 1) * (LB + 1) * TS];
    for (int d = 0; d < 3; ++d) {
        double xab = A[d] - B[d];
        hermite_e<LA, LB>(a, b, xab, exp(-mu * xab * xab), e[d]);
        rec[R_PX + d] = A[d] - (b / p) * xab;
    }
    rec[R_P] = p;
    rec[R_A] = a;
    rec[R_B] = b;
    rec[R_C] = ca * cb;
    rec[R_IP] = 1.0 / p;
    double* E = rec + REC_HDR;
    for (int i = 0; i < NH(LA + LB) * NAB; ++i) E[i] = 0.0;
    for (int ia = 0; ia < NA; ++ia) {
        int ax = cart_x(LA, ia), ay = cart_y(LA, ia), az = cart_z(LA, ia);
        double fa = cart_factor(ax, ay, az);
        for (int ib = 0; ib < NB; ++ib) {
            int bx = cart_x(LB, ib), by = cart_y(LB, ib), bz = cart_z(LB, ib);
            double f = fa * cart_factor(bx, by, bz) * (ca * cb);
            int ab = ia * NB + ib;
            for (int t = 0; t <= ax + bx; ++t)
                for (int u = 0; u <= ay + by; ++u)
                    for (int v = 0; v <= az + bz; ++v)
                        E[hidx(t, u, v) * NAB + ab] = f * e[0][(ax * (LB + 1) + bx) * TS + t] *
                                                      e[1][(ay * (LB + 1) + by) * TS + u] *
                                                      e[2][(az * (LB + 1) + bz) * TS + v];
        }
    }
}

// record of a primitive pair for a runtime pair class
MD_HD void write_record_any(int la, int lb, double a, double b, double ca, double cb, const double* A,
                            const double* B, double* r) {
    switch (pair_class(la, lb)) {
        case 0: write_record<0, 0>(a, b, ca, cb, A, B, r); break;
        case 1: write_record<1, 0>(a, b, ca, cb, A, B, r); break;
        case 2: write_record<1, 1>(a, b, ca, cb, A, B, r); break;
        case 3: write_record<2, 0>(a, b, ca, cb, A, B, r); break;
        case 4: write_record<2, 1>(a, b, ca, cb, A, B, r); break;
        default: write_record<2, 2>(a, b, ca, cb, A, B, r); break;
    }
}

}  // namespace md
