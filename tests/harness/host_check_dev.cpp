// The device derivative kernel's work items on the host (test infrastructure):
// one shell quartet exactly as grad_class.hip runs it -- NCB bra / NCK ket
// columns per item, the staged Boys slice with its exp(-T_k) column, arithmetic
// in HC_REAL -- over every unique quartet. Compiled once per precision; with
// -DHC_RSQ_MODEL (host_check_common.h) md_rsqrt starts from a model of the
// hardware estimate and the large-T Boys branch uses it too.

#ifndef HC_REAL
#error "compile with -DHC_REAL=double or -DHC_REAL=float"
#endif

#include "host_check_common.h"

// quartet classes instantiated (the sanitizer build keeps the s / p classes, 0..5)
#ifndef HC_NCLASS
#define HC_NCLASS NQCLASS
#endif

#define HC_CAT2(a, b) a##b
#define HC_CAT(a, b) HC_CAT2(a, b)

namespace {
// host twin of boys_stage: BOYS_SLICE orders from Ltop, then exp(-T_k)
std::vector<double> slice_table(int Ltop) {
    std::vector<double> t((size_t)BOYS_NGRID * BOYS_SROW);
    for (int k = 0; k < BOYS_NGRID; ++k)
        for (int c = 0; c < BOYS_SROW; ++c)
            t[(size_t)k * BOYS_SROW + c] = c < BOYS_SLICE ? TAB()[(size_t)k * BOYS_NCOL + Ltop + c] : std::exp(-0.1 * k);
    return t;
}

// One quartet exactly as the gfx950 derivative kernel runs it (grad_class.hip):
// work items of NCB bra / NCK ket columns, staged Boys slice, arithmetic in Real,
// per-item sums added as fixed point
template <int QC, typename Real>
void grad_quartet_device(const State& S, const Basis& Bs, int P, int Q, const double* dc, const double* dx, int nx,
                         double kfac, double thr, const std::vector<double>& qscr, std::vector<long long>& fx) {
    constexpr int hi = [] { int h = 0; while ((h + 1) * (h + 2) / 2 <= QC) ++h; return h; }();
    constexpr int lo = QC - hi * (hi + 1) / 2;
    constexpr int LA = la_of(hi), LB = lb_of(hi), LC = la_of(lo), LD = lb_of(lo);
    using Qt = Quartet<LA, LB, LC, LD>;
    constexpr int NAB = Qt::NAB, NCD = Qt::NCD, NB = Qt::NB, ND = Qt::ND;
    constexpr int NCB = grad_ncol<LA, LB, LC, LD>(), NCK = grad_ncol<LC, LD, LA, LB>();
    constexpr int NIB = NCD / NCB, NIK = NAB / NCK;
    static const std::vector<double> btab = slice_table(LA + LB + LC + LD + 1);
    int sA = S.pa[P], sB = S.pb[P], sC = S.pa[Q], sD = S.pb[Q];
    int aA = (int)Bs.atom[sA], aB = (int)Bs.atom[sB], aC = (int)Bs.atom[sC], aD = (int)Bs.atom[sD];
    if (aA == aB && aB == aC && aC == aD) return;
    double deg = 0.5;
    if (sA != sB) deg *= 2.0;
    if (sC != sD) deg *= 2.0;
    if (P != Q) deg *= 2.0;
    const size_t n = S.nao;
    int oA = S.ao[sA], oB = S.ao[sB], oC = S.ao[sC], oD = S.ao[sD];
    auto gam = [&](int ab, int cd) {
        size_t m = oA + ab / NB, nn = oB + ab % NB, l = oC + cd / ND, s = oD + cd % ND;
        double v = dc[m * n + nn] * dc[l * n + s];
        for (int x = 0; x < nx; ++x) {
            const double* X = dx + x * n * n;
            v -= 0.5 * kfac * (X[m * n + l] * X[nn * n + s] + X[m * n + s] * X[nn * n + l]);
        }
        return v * deg;
    };
    double gmax = 0.0;
    for (int ab = 0; ab < NAB; ++ab)
        for (int cd = 0; cd < NCD; ++cd) gmax = std::max(gmax, std::fabs(gam(ab, cd)));
    if (qscr[P] * qscr[Q] * gmax < thr) return;
    const bool need_bra = aA != aB, need_ket = aC != aD;
    auto add = [&](int a, int d, double x) { long long h, l2; fx_split(x, h, l2); fx[2 * (3 * a + d)] += h; fx[2 * (3 * a + d) + 1] += l2; };
    const double* rb = S.recs.data() + S.off[P];
    const double* rk = S.recs.data() + S.off[Q];
    // the kernel's folded items (grad_class.hip, fold_item)
    constexpr int FOLD = fold_ncol<LA, LB, LC, LD, sizeof(Real) == 8>();
    if constexpr (FOLD > 0) {
        static_for<0, NCD / FOLD>([&](auto jj) {
            constexpr int KC0 = decltype(jj)::value * FOLD;
            Real gm[NAB * FOLD];
            for (int ab = 0; ab < NAB; ++ab)
                for (int c = 0; c < FOLD; ++c) gm[ab * FOLD + c] = Real(gam(ab, KC0 + c));
            double gA[3], gB[3], gC[3];
            grad_pass_c<LA, LB, LC, LD, FOLD, KC0, true, Real>(rb, S.npp[P], rk, S.npp[Q], btab.data(),
                                                               Bs.pos + 3 * aA, Bs.pos + 3 * aB, Bs.pos + 3 * aC,
                                                               aC != aD, gm, gA, gB, gC);
            for (int d = 0; d < 3; ++d) {
                add(aA, d, gA[d]);
                add(aB, d, gB[d]);
                add(aC, d, gC[d]);
                add(aD, d, -(gA[d] + gB[d] + gC[d]));
            }
        });
        return;
    }
    for (int j = 0; j < NIB + NIK; ++j) {
        const bool bra = j < NIB;
        if (bra && !(need_bra || (!need_ket && NAB >= NCD))) continue;
        if (!bra && !(need_ket || (!need_bra && NAB < NCD))) continue;
        double g1[3], g2[3];
        if (bra) {
            const int kc0 = j * NCB;
            Real gm[NAB * NCB];
            for (int ab = 0; ab < NAB; ++ab)
                for (int c = 0; c < NCB; ++c) gm[ab * NCB + c] = Real(gam(ab, kc0 + c));
            grad_pass<LA, LB, LC, LD, NCB, true, Real>(rb, S.npp[P], rk, S.npp[Q], btab.data(), Bs.pos + 3 * aA,
                                                       Bs.pos + 3 * aB, kc0, gm, g1, g2);
            for (int d = 0; d < 3; ++d) {
                if (need_bra) { add(aA, d, g1[d]); add(aB, d, g2[d]); if (!need_ket) add(aC, d, -(g1[d] + g2[d])); }
                else { add(aA, d, g1[d] + g2[d]); add(aC, d, -(g1[d] + g2[d])); }
            }
        } else {
            const int ab0 = (j - NIB) * NCK;
            Real gm[NCD * NCK];
            for (int cd = 0; cd < NCD; ++cd)
                for (int c = 0; c < NCK; ++c) gm[cd * NCK + c] = Real(gam(ab0 + c, cd));
            grad_pass<LC, LD, LA, LB, NCK, true, Real>(rk, S.npp[Q], rb, S.npp[P], btab.data(), Bs.pos + 3 * aC,
                                                       Bs.pos + 3 * aD, ab0, gm, g1, g2);
            for (int d = 0; d < 3; ++d) {
                if (need_ket) { add(aC, d, g1[d]); add(aD, d, g2[d]); if (!need_bra) add(aA, d, -(g1[d] + g2[d])); }
                else { add(aC, d, g1[d] + g2[d]); add(aA, d, -(g1[d] + g2[d])); }
            }
        }
    }
}
}  // namespace

extern "C" void HC_CAT(hc_grad_two_dev_, HC_REAL)(int nshell, int natm, const long long* atom, const long long* l,
                                                  const long long* prim, const long long* block, const double* ex,
                                                  const double* coef, const double* pos, const double* dc, int nx,
                                                  const double* dx, double kfac, double thr, double* g) {
    Basis B{nshell, natm, atom, l, prim, block, ex, coef, pos};
    State S;
    build(B, S);
    std::vector<double> qscr;
    std::vector<long long> fx(6 * natm, 0);
    for_each_quartet(S, B, qscr, [&](int bra, int ket, int qc) {
        static_for<0, HC_NCLASS>([&](auto k) {
            if (decltype(k)::value == qc)
                grad_quartet_device<decltype(k)::value, HC_REAL>(S, B, bra, ket, dc, dx, nx, kfac, thr, qscr, fx);
        });
    });
    for (int i = 0; i < 3 * natm; ++i) g[i] = fx_value(fx[2 * i], fx[2 * i + 1]);
}
