// Shared pieces of the CPU check harness (test infrastructure): the basis and
// pair-record state built with the kernels' own host/device templates, the
// Boys table, Schwarz bounds, and the loop over unique shell quartets that the
// gradient variants run.
#pragma once

#include <cstdint>
#include <cstring>

#ifdef HC_RSQ_MODEL
// Model of the device's reciprocal-square-root estimate: the exact value times
// (1 + 5e-8 s), s a deterministic pseudo-random number in [-1, 1] drawn from
// the bits of x (the FP64 estimate is good to ~5e-8), before md_rsqrt's Newton
// steps; the large-T branch of the staged Boys function then also uses it.
static inline double hc_rsq_model(double x, double y) {
    uint64_t b;
    std::memcpy(&b, &x, sizeof b);
    const double s = (double)((b * 0x9E3779B97F4A7C15ull) >> 11) * 0x1.0p-52 - 1.0;
    return y * (1.0 + 5e-8 * s);
}
#define MD_HOST_RSQ_ESTIMATE hc_rsq_model
#define MD_HOST_LARGE_T_RSQ 1
#endif

#include "md_digest.h"
#include "md_grad.h"
#include "md_oneel.h"
#include "md_pairs.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

using namespace md;

namespace {

struct Basis {
    int nshell, natm;
    const long long *atom, *l, *prim, *block;
    const double *ex, *coef, *pos;
};

struct State {
    int npair, nao;
    std::vector<int> pa, pb, cls, ao, npp;
    std::vector<long long> off;
    std::vector<int> rs;
    std::vector<double> recs, qshell;
};

std::vector<double> boys_tab() {
    const int NG = 402, NR = 43, top = NR - 1;
    std::vector<double> tab((size_t)NG * NR), col(NR);
    for (int k = 0; k < NG; ++k) {
        double x = k * 0.1, term = 1.0 / (2 * top + 1), total = term;
        for (int m = 1; m < 400; ++m) {
            term = term * (2.0 * x) / (double)(2 * top + 2 * m + 1);
            total += term;
        }
        double e = std::exp(-x);
        col[top] = e * total;
        for (int n = top - 1; n >= 0; --n) col[n] = (2.0 * x * col[n + 1] + e) / (double)(2 * n + 1);
        for (int n = 0; n < NR; ++n) tab[(size_t)k * NR + n] = col[n];
    }
    return tab;
}

const std::vector<double>& TAB() {
    static std::vector<double> t = boys_tab();
    return t;
}

void build(const Basis& B, State& S) {
    S.ao.assign(B.nshell + 1, 0);
    for (int s = 0; s < B.nshell; ++s) S.ao[s + 1] = S.ao[s] + NC((int)B.l[s]);
    S.nao = S.ao[B.nshell];
    S.npair = B.nshell * (B.nshell + 1) / 2;
    S.pa.resize(S.npair); S.pb.resize(S.npair); S.cls.resize(S.npair); S.npp.resize(S.npair);
    S.off.resize(S.npair); S.rs.resize(S.npair); S.qshell.resize(S.npair);
    long long off = 0;
    int P = 0;
    for (int s1 = 0; s1 < B.nshell; ++s1)
        for (int s2 = 0; s2 <= s1; ++s2, ++P) {
            int a = s1, b = s2;
            if (B.l[s2] > B.l[s1]) std::swap(a, b);
            S.pa[P] = a; S.pb[P] = b;
            int la = (int)B.l[a], lb = (int)B.l[b];
            S.cls[P] = pair_class(la, lb);
            S.rs[P] = REC_HDR + NH(la + lb) * NC(la) * NC(lb);
            S.off[P] = off;
            off += (long long)S.rs[P] * (B.prim[a + 1] - B.prim[a]) * (B.prim[b + 1] - B.prim[b]);
        }
    S.recs.assign(off, 0.0);
    for (P = 0; P < S.npair; ++P) {
        int a = S.pa[P], b = S.pb[P];
        const double* A = B.pos + 3 * B.atom[a];
        const double* C = B.pos + 3 * B.atom[b];
        double r2 = 0.0;
        for (int d = 0; d < 3; ++d) { double x = A[d] - C[d]; r2 += x * x; }
        int n = 0;
        for (long long ka = B.prim[a]; ka < B.prim[a + 1]; ++ka)
            for (long long kb = B.prim[b]; kb < B.prim[b + 1]; ++kb) {
                double ea = B.ex[ka], eb = B.ex[kb];
                if (!(ea * eb / (ea + eb) * r2 <= 50.0)) continue;
                write_record_any((int)B.l[a], (int)B.l[b], ea, eb, B.coef[ka], B.coef[kb], A, C,
                                 S.recs.data() + S.off[P] + (long long)n * S.rs[P]);
                ++n;
            }
        S.npp[P] = n;
    }
}

template <int LA, int LB>
double schwarz(const State& S, int P) {
    using Qt = Quartet<LA, LB, LA, LB>;
    const double* r = S.recs.data() + S.off[P];
    double m = 0.0;
    if constexpr (split_by_cd<LA, LB, LA, LB>()) {
        for (int cd = 0; cd < Qt::NCD; ++cd) {
            double acc[Qt::NAB];
            eri_quartet_cd<LA, LB, LA, LB>(r, S.npp[P], r, S.npp[P], TAB().data(), cd, acc);
            m = std::max(m, std::fabs(acc[cd]));
        }
    } else {
        double acc[Qt::NAB * Qt::NCD];
        eri_quartet<LA, LB, LA, LB>(r, S.npp[P], r, S.npp[P], TAB().data(), acc);
        for (int ab = 0; ab < Qt::NAB; ++ab) m = std::max(m, std::fabs(acc[ab * Qt::NCD + ab]));
    }
    return std::sqrt(m);
}

constexpr int la_of(int pc) { return pc == 0 ? 0 : pc <= 2 ? 1 : 2; }
constexpr int lb_of(int pc) { return pc - la_of(pc) * (la_of(pc) + 1) / 2; }

struct HostSink {
    long long *Jh, *Jl, *Kh, *Kl;
    void j(size_t i, double x) const { long long h, l; fx_split(x, h, l); Jh[i] += h; Jl[i] += l; }
    void k(size_t i, double x) const { long long h, l; fx_split(x, h, l); Kh[i] += h; Kl[i] += l; }
};


// Schwarz bounds of every shell pair, their block-pair maxima (the screening
// values), then f(bra, ket, class) for every unique shell quartet P >= Q
template <class F>
void for_each_quartet(State& S, const Basis& B, std::vector<double>& qscr, F f) {
    for (int P = 0; P < S.npair; ++P) {
        switch (S.cls[P]) {
            case 0: S.qshell[P] = schwarz<0, 0>(S, P); break;
            case 1: S.qshell[P] = schwarz<1, 0>(S, P); break;
            case 2: S.qshell[P] = schwarz<1, 1>(S, P); break;
            case 3: S.qshell[P] = schwarz<2, 0>(S, P); break;
            case 4: S.qshell[P] = schwarz<2, 1>(S, P); break;
            default: S.qshell[P] = schwarz<2, 2>(S, P); break;
        }
    }
    const int nshell = B.nshell;
    int nblock = (int)B.block[nshell - 1] + 1;
    std::vector<double> qb(nblock * (nblock + 1) / 2, 0.0);
    qscr.assign(S.npair, 0.0);
    std::vector<int> bpof(S.npair);
    int P = 0;
    for (int s1 = 0; s1 < nshell; ++s1)
        for (int s2 = 0; s2 <= s1; ++s2, ++P) {
            int b1 = (int)B.block[s1], b2 = (int)B.block[s2];
            bpof[P] = b1 * (b1 + 1) / 2 + b2;
            qb[bpof[P]] = std::max(qb[bpof[P]], S.qshell[P]);
        }
    for (P = 0; P < S.npair; ++P) qscr[P] = qb[bpof[P]];
    for (P = 0; P < S.npair; ++P)
        for (int Q = 0; Q <= P; ++Q) {
            int cp = S.cls[P], cq = S.cls[Q];
            int hi = std::max(cp, cq), lo = std::min(cp, cq);
            f(cp >= cq ? P : Q, cp >= cq ? Q : P, hi * (hi + 1) / 2 + lo);
        }
}

}  // namespace
