// Test harness (CPU only): runs the engine's own kernel arithmetic -- the
// __host__ __device__ templates of md_core.h / md_pairs.h / md_eri.h /
// md_digest.h / md_grad.h / md_oneel.h -- serially on the host, with the same
// pair orientation, record layout, screening and fixed-point accumulation as
// the gfx950 path. The CPU test suite compares it with the oracle, which
// checks the kernels' mathematics without a GPU. Not part of the product.
// The device kernel's own derivative work items (column groups, staged Boys
// slice, FP64 and FP32) live in host_check_dev.cpp.

#include "host_check_common.h"

namespace {
template <int QC>
void jk_quartet(const State& S, int P, int Q, const double* D, int nset, std::vector<long long>& fx) {
    constexpr int hi = [] { int h = 0; while ((h + 1) * (h + 2) / 2 <= QC) ++h; return h; }();
    constexpr int lo = QC - hi * (hi + 1) / 2;
    constexpr int LA = la_of(hi), LB = lb_of(hi), LC = la_of(lo), LD = lb_of(lo);
    using Qt = Quartet<LA, LB, LC, LD>;
    int sA = S.pa[P], sB = S.pb[P], sC = S.pa[Q], sD = S.pb[Q];
    double w = 1.0;
    if (sA == sB) w *= 0.5;
    if (sC == sD) w *= 0.5;
    if (P == Q) w *= 0.5;
    size_t nn = (size_t)S.nao * S.nao;
    auto sink_of = [&](int s) {
        return HostSink{fx.data() + (4 * s + 0) * nn, fx.data() + (4 * s + 1) * nn, fx.data() + (4 * s + 2) * nn,
                        fx.data() + (4 * s + 3) * nn};
    };
    const double* rb = S.recs.data() + S.off[P];
    const double* rk = S.recs.data() + S.off[Q];
    if constexpr (split_by_cd<LA, LB, LC, LD>()) {
        for (int cd = 0; cd < Qt::NCD; ++cd) {
            double acc[Qt::NAB];
            eri_quartet_cd<LA, LB, LC, LD>(rb, S.npp[P], rk, S.npp[Q], TAB().data(), cd, acc);
            digest_cd<LA, LB, LC, LD>(acc, D, S.nao, nset, sink_of, S.ao[sA], S.ao[sB], S.ao[sC], S.ao[sD], cd, w);
        }
    } else {
        double acc[Qt::NAB * Qt::NCD];
        eri_quartet<LA, LB, LC, LD>(rb, S.npp[P], rk, S.npp[Q], TAB().data(), acc);
        digest_full<LA, LB, LC, LD>(acc, D, S.nao, nset, sink_of, S.ao[sA], S.ao[sB], S.ao[sC], S.ao[sD], w);
    }
}



template <int QC>
void grad_quartet(const State& S, const Basis& Bs, int P, int Q, const double* dc, const double* dx, int nx, double kfac,
                  double thr, const std::vector<double>& qscr, std::vector<long long>& fx) {
    constexpr int hi = [] { int h = 0; while ((h + 1) * (h + 2) / 2 <= QC) ++h; return h; }();
    constexpr int lo = QC - hi * (hi + 1) / 2;
    constexpr int LA = la_of(hi), LB = lb_of(hi), LC = la_of(lo), LD = lb_of(lo);
    using Qt = Quartet<LA, LB, LC, LD>;
    constexpr int NAB = Qt::NAB, NCD = Qt::NCD, NB = Qt::NB, ND = Qt::ND;
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
    bool need_bra = aA != aB, need_ket = aC != aD;
    auto add = [&](int a, int d, double x) { long long h, l2; fx_split(x, h, l2); fx[2 * (3 * a + d)] += h; fx[2 * (3 * a + d) + 1] += l2; };
    const double* rb = S.recs.data() + S.off[P];
    const double* rk = S.recs.data() + S.off[Q];
    if (need_bra || (!need_ket && NAB >= NCD))
        for (int cd = 0; cd < NCD; ++cd) {
            double g[NAB], g1[3], g2[3];
            for (int ab = 0; ab < NAB; ++ab) g[ab] = gam(ab, cd);
            grad_pass_kc<LA, LB, LC, LD>(rb, S.npp[P], rk, S.npp[Q], TAB().data(), Bs.pos + 3 * aA, Bs.pos + 3 * aB, cd, g, g1, g2);
            for (int d = 0; d < 3; ++d) {
                if (need_bra) { add(aA, d, g1[d]); add(aB, d, g2[d]); if (!need_ket) add(aC, d, -(g1[d] + g2[d])); }
                else { add(aA, d, g1[d] + g2[d]); add(aC, d, -(g1[d] + g2[d])); }
            }
        }
    if (need_ket || (!need_bra && NAB < NCD))
        for (int ab = 0; ab < NAB; ++ab) {
            double g[NCD], g1[3], g2[3];
            for (int cd = 0; cd < NCD; ++cd) g[cd] = gam(ab, cd);
            grad_pass_kc<LC, LD, LA, LB>(rk, S.npp[Q], rb, S.npp[P], TAB().data(), Bs.pos + 3 * aC, Bs.pos + 3 * aD, ab, g, g1, g2);
            for (int d = 0; d < 3; ++d) {
                if (need_ket) { add(aC, d, g1[d]); add(aD, d, g2[d]); if (!need_bra) add(aA, d, -(g1[d] + g2[d])); }
                else { add(aC, d, g1[d] + g2[d]); add(aA, d, -(g1[d] + g2[d])); }
            }
        }
}
}  // namespace

extern "C" {

// Schwarz bounds per shell pair and per block pair, the screened quartet
// count, and J/K of nset symmetric densities, all through the kernel code.
long hc_run(int nshell, int natm, const long long* atom, const long long* l, const long long* prim,
            const long long* block, const double* ex, const double* coef, const double* pos, double thr,
            int nset, const double* D, double* J, double* K, double* qshell, double* qblock, long long* quartets,
            long cap) {
    Basis B{nshell, natm, atom, l, prim, block, ex, coef, pos};
    State S;
    build(B, S);
    for (int P = 0; P < S.npair; ++P) {
        double q = 0.0;
        switch (S.cls[P]) {
            case 0: q = schwarz<0, 0>(S, P); break;
            case 1: q = schwarz<1, 0>(S, P); break;
            case 2: q = schwarz<1, 1>(S, P); break;
            case 3: q = schwarz<2, 0>(S, P); break;
            case 4: q = schwarz<2, 1>(S, P); break;
            default: q = schwarz<2, 2>(S, P); break;
        }
        S.qshell[P] = q;
        if (qshell) qshell[P] = q;
    }
    int nblock = (int)block[nshell - 1] + 1;
    int nbp = nblock * (nblock + 1) / 2;
    std::vector<double> qb(nbp, 0.0);
    std::vector<int> bpof(S.npair);
    int P = 0;
    for (int s1 = 0; s1 < nshell; ++s1)
        for (int s2 = 0; s2 <= s1; ++s2, ++P) {
            int b1 = (int)block[s1], b2 = (int)block[s2];
            bpof[P] = b1 * (b1 + 1) / 2 + b2;
            qb[bpof[P]] = std::max(qb[bpof[P]], S.qshell[P]);
        }
    if (qblock) for (int i = 0; i < nbp; ++i) qblock[i] = qb[i];
    size_t nn = (size_t)S.nao * S.nao;
    std::vector<long long> fx(4 * nset * nn, 0);
    std::vector<double> Ds(nset * nn);
    for (int s = 0; s < nset; ++s)
        for (int r = 0; r < S.nao; ++r)
            for (int c = 0; c < S.nao; ++c)
                Ds[s * nn + r * S.nao + c] = 0.5 * (D[s * nn + r * S.nao + c] + D[s * nn + c * S.nao + r]);
    long nq = 0;
    for (P = 0; P < S.npair; ++P)
        for (int Q = 0; Q <= P; ++Q) {
            if (qb[bpof[P]] * qb[bpof[Q]] < thr) continue;
            if (quartets && nq < cap) { quartets[2 * nq] = P; quartets[2 * nq + 1] = Q; }
            ++nq;
            int cp = S.cls[P], cq = S.cls[Q];
            int hi = std::max(cp, cq), lo = std::min(cp, cq);
            int bra = cp >= cq ? P : Q, ket = cp >= cq ? Q : P;
            int qc = hi * (hi + 1) / 2 + lo;
            static_for<0, NQCLASS>([&](auto k) {
                if (decltype(k)::value == qc) jk_quartet<decltype(k)::value>(S, bra, ket, Ds.data(), nset, fx);
            });
        }
    for (int s = 0; s < nset; ++s)
        for (int r = 0; r < S.nao; ++r)
            for (int c = 0; c < S.nao; ++c) {
                size_t ij = r * S.nao + c, ji = c * S.nao + r;
                const long long* f = fx.data() + 4 * s * nn;
                J[s * nn + ij] = fx_value(f[ij] + f[ji], f[nn + ij] + f[nn + ji]);
                K[s * nn + ij] = fx_value(f[2 * nn + ij] + f[2 * nn + ji], f[3 * nn + ij] + f[3 * nn + ji]);
            }
    return nq;
}

// S, T, V, M (3,n,n) and the one-electron gradient (D, W symmetric) via md_oneel.h
void hc_oneel(int nshell, int natm, const long long* atom, const long long* l, const long long* prim, const double* ex,
              const double* coef, const double* pos, const double* Z, const double* D, const double* W, double* S,
              double* T, double* V, double* M, double* g1) {
    std::vector<int> at(nshell), ll(nshell), pr(nshell + 1), ao(nshell + 1, 0);
    for (int s = 0; s < nshell; ++s) { at[s] = (int)atom[s]; ll[s] = (int)l[s]; ao[s + 1] = ao[s] + NC(ll[s]); }
    for (int s = 0; s <= nshell; ++s) pr[s] = (int)prim[s];
    int n = ao[nshell];
    ShellView B{at.data(), ll.data(), pr.data(), ao.data(), ex, coef, pos};
    const double O[3] = {0, 0, 0};
    std::vector<double> Wm((size_t)n * n);
    for (size_t i = 0; i < Wm.size(); ++i) Wm[i] = -W[i];
    std::vector<long long> fx(6 * natm, 0);
    auto add = [&](int a, int d, double x) { long long h, lo; fx_split(x, h, lo); fx[2 * (3 * a + d)] += h; fx[2 * (3 * a + d) + 1] += lo; };
    for (int s1 = 0; s1 < nshell; ++s1)
        for (int s2 = 0; s2 <= s1; ++s2) {
            overlap_type_pair(B, n, s1, s2, O, S, T, M);
            nuclear_pair(B, n, s1, s2, natm, Z, pos, TAB().data(), V);
            if (s1 != s2 && at[s1] != at[s2]) {
                double g[3];
                ovlkin_grad_pair(B, n, s1, s2, Wm.data(), D, g);
                for (int d = 0; d < 3; ++d) { add(at[s1], d, g[d]); add(at[s2], d, -g[d]); }
            }
            nuclear_grad_pair(B, n, s1, s2, natm, Z, D, TAB().data(), add);
        }
    for (int i = 0; i < 3 * natm; ++i) g1[i] = fx_value(fx[2 * i], fx[2 * i + 1]);
}

}  // extern "C"

extern "C" {
void hc_grad_two_dev_double(int, int, const long long*, const long long*, const long long*, const long long*,
                            const double*, const double*, const double*, const double*, int, const double*, double,
                            double, double*);
void hc_grad_two_dev_float(int, int, const long long*, const long long*, const long long*, const long long*,
                           const double*, const double*, const double*, const double*, int, const double*, double,
                           double, double*);

// two-electron gradient through md_grad.h over all unique shell quartets, one ket
// column per pass with the full Boys table in FP64 (the reference's structure)
void hc_grad_two(int nshell, int natm, const long long* atom, const long long* l, const long long* prim,
                 const long long* block, const double* ex, const double* coef, const double* pos, const double* dc,
                 int nx, const double* dx, double kfac, double thr, double* g) {
    Basis B{nshell, natm, atom, l, prim, block, ex, coef, pos};
    State S;
    build(B, S);
    std::vector<double> qscr;
    std::vector<long long> fx(6 * natm, 0);
    for_each_quartet(S, B, qscr, [&](int bra, int ket, int qc) {
        static_for<0, NQCLASS>([&](auto k) {
            if (decltype(k)::value == qc) grad_quartet<decltype(k)::value>(S, B, bra, ket, dc, dx, nx, kfac, thr, qscr, fx);
        });
    });
    for (int i = 0; i < 3 * natm; ++i) g[i] = fx_value(fx[2 * i], fx[2 * i + 1]);
}

// variant 0: hc_grad_two; 1: the device kernel's work items in FP64; 2: in FP32
void hc_grad_two_variant(int variant, int nshell, int natm, const long long* atom, const long long* l,
                         const long long* prim, const long long* block, const double* ex, const double* coef,
                         const double* pos, const double* dc, int nx, const double* dx, double kfac, double thr,
                         double* g) {
    if (variant == 1)
        hc_grad_two_dev_double(nshell, natm, atom, l, prim, block, ex, coef, pos, dc, nx, dx, kfac, thr, g);
    else if (variant == 2)
        hc_grad_two_dev_float(nshell, natm, atom, l, prim, block, ex, coef, pos, dc, nx, dx, kfac, thr, g);
    else
        hc_grad_two(nshell, natm, atom, l, prim, block, ex, coef, pos, dc, nx, dx, kfac, thr, g);
}
}  // extern "C"
