// Derivative-ERI gradient kernel of one quartet class (compiled once per
// class with -DAIMDHF_QCLASS=k, like jk_class.hip).
//
// A pre-pass (one thread per quartet) screens as aimd/qc/integrals.py:1279-1329
// does, with one difference: quartets with all four centres on one atom are
// dropped, and the others are kept while Q_P Q_Q times a bound of max|Gamma|
// reaches the threshold. The bound comes from per-shell-pair density maxima
// and is never below the element-wise max|Gamma| the reference uses, so the
// kept list is a superset of the reference's; every extra quartet has
// Q_P Q_Q max|Gamma| below the threshold, i.e. contributes below it.
// Work item = (quartet, j): j < NIB is the bra pass for ket components
// j*NCB .. j*NCB+NCB-1, NIB <= j the ket pass for NCK bra components, NCB /
// NCK as many as fit the register budget (one R evaluation per primitive
// quartet and item); a pass is replaced by translational invariance when its
// pair sits on one atom. Folded classes (fold_ncol) instead take items of the
// bra pass that also fold in the ket's first-centre derivative (one item per
// ket column, or all columns of a light class in one), so the ket pass and its
// R evaluations disappear. Items are issued wave by wave (64 quartets, one pass
// or column group per wavefront), so the lanes of a wavefront run the same
// code. Contributions are accumulated per workgroup in LDS and flushed once, as
// exact fixed-point integers, so the gradient is bitwise reproducible.

#include "aimdhf_internal.h"
#include "md_grad.h"

#ifndef AIMDHF_QCLASS
#error "compile with -DAIMDHF_QCLASS=<0..20>"
#endif

namespace aimdhf {
using namespace md;

namespace {
constexpr int QC = AIMDHF_QCLASS;
// minimum waves per SIMD asked of the register allocator: two instead of one
// for (pp|ss), (pp|pp) and (ds|ss) in FP64, which then spill little and run
// faster on gfx950; the heavier classes slow down at two waves (more than 80
// spilled registers), (ss|ss) and (ps|ss) at one more wave
template <typename Real>
constexpr int grad_waves() {
    if constexpr (sizeof(Real) == 4) return QC == 10 || (QC >= 14 && QC <= 18) ? 2 : 1;
    return QC == 3 || QC == 5 || QC == 6 ? 2 : 1;
}
constexpr int pc_hi() { int h = 0; while ((h + 1) * (h + 2) / 2 <= QC) ++h; return h; }
constexpr int pc_lo() { return QC - pc_hi() * (pc_hi() + 1) / 2; }
constexpr int la_of(int pc) { return pc == 0 ? 0 : pc <= 2 ? 1 : 2; }
constexpr int lb_of(int pc) { return pc - la_of(pc) * (la_of(pc) + 1) / 2; }
constexpr int LA = la_of(pc_hi()), LB = lb_of(pc_hi()), LC = la_of(pc_lo()), LD = lb_of(pc_lo());
using Qt = Quartet<LA, LB, LC, LD>;
constexpr int NAB = Qt::NAB, NCD = Qt::NCD, NB = Qt::NB, ND = Qt::ND;
constexpr int NCB = grad_ncol<LA, LB, LC, LD>(), NCK = grad_ncol<LC, LD, LA, LB>();
constexpr int NIB = NCD / NCB, NIK = NAB / NCK;
// classes with FOLD<Real> > 0 take all four centres from the bra pass with the
// ket's first-centre derivative folded in (grad_pass_c), FOLD ket columns per
// item; NT<Real> items per quartet
template <typename Real>
constexpr int FOLD = fold_ncol<LA, LB, LC, LD, sizeof(Real) == 8>();
template <typename Real>
constexpr int NT = FOLD<Real> > 0 ? NCD / FOLD<Real> : NIB + NIK;
constexpr int NT_MAX = NT<double> > NT<float> ? NT<double> : NT<float>;
// LDS words for the per-atom sums: with the 35 KB Boys slice, a workgroup stays
// within 64 KB of LDS
constexpr int GRAD_LDS_WORDS = 3600;

__device__ __forceinline__ void lds_add(long long* lds, int atom, int d, double x) {
    long long h, l;
    fx_split(x, h, l);
    if (h) atomicAdd((unsigned long long*)&lds[2 * (3 * atom + d)], (unsigned long long)h);
    if (l) atomicAdd((unsigned long long*)&lds[2 * (3 * atom + d) + 1], (unsigned long long)l);
}

// Gamma element of AOs m, n, l, s (offsets within replica rep)
__device__ __forceinline__ double gamma_el(const GradArgs& a, size_t rep, int m, int n, int l, int s, double deg) {
    const size_t N = a.ld, NN = N * N;
    const double* C = a.dc + rep * NN;
    double v = C[m * N + n] * C[l * N + s];
    for (int x = 0; x < a.nx; ++x) {
        const double* X = a.dx + (rep * a.nx + x) * NN;
        v -= 0.5 * a.kfac * (X[m * N + l] * X[n * N + s] + X[m * N + s] * X[n * N + l]);
    }
    return v * deg;
}

// index of shell pair (s, t) (same replica)
__device__ __forceinline__ int pair_index(const GradArgs& a, int s, int t) {
    int base = 0;
    if (a.shell_first) {
        base = a.shell_pair0[s];
        const int f = a.shell_first[s];
        s -= f;
        t -= f;
    }
    return base + (s > t ? s * (s + 1) / 2 + t : t * (t + 1) / 2 + s);
}

// Screening pre-pass, one thread per quartet (256-quartet blocks): keep[q] = 0
// (dropped), 1 (FP64) or 2 (FP32: Q_P Q_Q max|Gamma| below fp32_switch); each
// block's counts of 1s and 2s go to bcnt[2 * block + {0, 1}]. max|Gamma| is
// bounded from the per-shell-pair density maxima (|Dc_AB| |Dc_CD| +
// |kfac|/2 sum_x (|Dx_AC| |Dx_BD| + |Dx_AD| |Dx_BC|)), an upper bound of the
// element-wise maximum, so no quartet above the threshold is dropped.
template <int Q>
__global__ __launch_bounds__(256) void k_grad_keep(GradArgs a, const int2* __restrict__ list, long long nq) {
    __shared__ int wc[2][4];
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    int keep = 0;
    if (q < nq) {
        const int2 pq = list[q];
        const int sA = a.pair_a[pq.x], sB = a.pair_b[pq.x], sC = a.pair_a[pq.y], sD = a.pair_b[pq.y];
        const int aA = a.shell_atom[sA], aB = a.shell_atom[sB], aC = a.shell_atom[sC], aD = a.shell_atom[sD];
        if (!(aA == aB && aB == aC && aC == aD)) {
            double deg = 0.5;
            if (sA != sB) deg *= 2.0;
            if (sC != sD) deg *= 2.0;
            if (pq.x != pq.y) deg *= 2.0;
            double gb = a.dmc[pair_index(a, sA, sB)] * a.dmc[pair_index(a, sC, sD)];
            const int pAC = pair_index(a, sA, sC), pBD = pair_index(a, sB, sD), pAD = pair_index(a, sA, sD),
                      pBC = pair_index(a, sB, sC);
            for (int x = 0; x < a.nx; ++x) {
                const double* m = a.dmx + (size_t)x * a.npair;
                gb += 0.5 * fabs(a.kfac) * (m[pAC] * m[pBD] + m[pAD] * m[pBC]);
            }
            const double b = a.qscr[pq.x] * a.qscr[pq.y] * (gb * deg);
            keep = (b < a.thr) ? 0 : (b < a.fp32_switch ? 2 : 1);
        }
        a.keep[q] = (unsigned char)keep;
    }
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const unsigned long long m1 = __ballot(keep == 1), m2 = __ballot(keep == 2);
    if (lane == 0) {
        wc[0][w] = (int)__popcll(m1);
        wc[1][w] = (int)__popcll(m2);
    }
    __syncthreads();
    if (threadIdx.x < 2)
        a.bcnt[2 * blockIdx.x + threadIdx.x] =
            wc[threadIdx.x][0] + wc[threadIdx.x][1] + wc[threadIdx.x][2] + wc[threadIdx.x][3];
}

// one 1024-thread block: exclusive scan of the per-block counts (both lists) into
// boff, totals into nsel[0..1]
__global__ __launch_bounds__(1024) void k_grad_scan(const int* __restrict__ bcnt, int nb, int* __restrict__ boff,
                                                    int* __restrict__ nsel) {
    __shared__ int ws[2][16];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    int run0 = 0, run1 = 0;
    for (int base = 0; base < nb; base += 1024) {
        const int i = base + t;
        const int x0 = i < nb ? bcnt[2 * i] : 0, x1 = i < nb ? bcnt[2 * i + 1] : 0;
        int v0 = x0, v1 = x1;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const int y0 = __shfl_up(v0, off, 64), y1 = __shfl_up(v1, off, 64);
            if (lane >= off) {
                v0 += y0;
                v1 += y1;
            }
        }
        if (lane == 63) {
            ws[0][w] = v0;
            ws[1][w] = v1;
        }
        __syncthreads();
        int p0 = 0, p1 = 0, t0 = 0, t1 = 0;
        for (int j = 0; j < 16; ++j) {
            if (j < w) {
                p0 += ws[0][j];
                p1 += ws[1][j];
            }
            t0 += ws[0][j];
            t1 += ws[1][j];
        }
        if (i < nb) {
            boff[2 * i] = run0 + p0 + v0 - x0;
            boff[2 * i + 1] = run1 + p1 + v1 - x1;
        }
        run0 += t0;
        run1 += t1;
        __syncthreads();
    }
    if (t == 0) {
        nsel[0] = run0;
        nsel[1] = run1;
    }
}

// ordered scatter of the selected quartets (same 256-quartet blocks as the pre-pass)
__global__ __launch_bounds__(256) void k_grad_scatter(const unsigned char* __restrict__ keep, long long nq,
                                                      const int* __restrict__ boff, int* __restrict__ sel64,
                                                      int* __restrict__ sel32) {
    __shared__ int wc[2][4];
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const int k = q < nq ? keep[q] : 0;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const unsigned long long m1 = __ballot(k == 1), m2 = __ballot(k == 2);
    if (lane == 0) {
        wc[0][w] = (int)__popcll(m1);
        wc[1][w] = (int)__popcll(m2);
    }
    __syncthreads();
    int o1 = boff[2 * blockIdx.x], o2 = boff[2 * blockIdx.x + 1];
    for (int j = 0; j < w; ++j) {
        o1 += wc[0][j];
        o2 += wc[1][j];
    }
    const unsigned long long below = lane ? (~0ull >> (64 - lane)) : 0ull;
    if (k == 1) sel64[o1 + __popcll(m1 & below)] = (int)q;
    if (k == 2) sel32[o2 + __popcll(m2 & below)] = (int)q;
}

// The folded item for ket columns KC0 .. KC0+FOLD<Real>-1 of quartet pq: A, B and C
// derivatives from one pass, D by translational invariance
template <int KC0, typename Real>
__device__ __forceinline__ void fold_item(const GradArgs& a, int2 pq, const double* btab, long long* lds) {
    const int sA = a.pair_a[pq.x], sB = a.pair_b[pq.x], sC = a.pair_a[pq.y], sD = a.pair_b[pq.y];
    const int aA = a.shell_atom[sA], aB = a.shell_atom[sB], aC = a.shell_atom[sC], aD = a.shell_atom[sD];
    double deg = 0.5;
    if (sA != sB) deg *= 2.0;
    if (sC != sD) deg *= 2.0;
    if (pq.x != pq.y) deg *= 2.0;
    const int oA = a.shell_ao[sA], oB = a.shell_ao[sB], oC = a.shell_ao[sC], oD = a.shell_ao[sD];
    const size_t rep = a.shell_rep ? (size_t)a.shell_rep[sA] : 0;
    constexpr int F = FOLD<Real>;
    Real gam[NAB * F];
    for (int ab = 0; ab < NAB; ++ab)
#pragma unroll
        for (int c = 0; c < F; ++c)
            gam[ab * F + c] =
                Real(gamma_el(a, rep, oA + ab / NB, oB + ab % NB, oC + (KC0 + c) / ND, oD + (KC0 + c) % ND, deg));
    double gA[3], gB[3], gC[3];
    grad_pass_c<LA, LB, LC, LD, F, KC0, true, Real>(a.recs + a.rec_off[pq.x], a.npp[pq.x],
                                                       a.recs + a.rec_off[pq.y], a.npp[pq.y], btab, a.pos + 3 * aA,
                                                       a.pos + 3 * aB, a.pos + 3 * aC, aC != aD, gam, gA, gB, gC);
    for (int d = 0; d < 3; ++d) {
        lds_add(lds, aA, d, gA[d]);
        lds_add(lds, aB, d, gB[d]);
        lds_add(lds, aC, d, gC[d]);
        lds_add(lds, aD, d, -(gA[d] + gB[d] + gC[d]));
    }
}

// Work item = (selected quartet, pass column group), derivative integrals in Real
template <int Q, typename Real>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(grad_waves<Real>()))) void k_grad_class(
    GradArgs a, const int2* __restrict__ list, const int* __restrict__ sel, const int* __restrict__ nsel) {
    extern __shared__ long long lds_acc[];
    __shared__ double btab[BOYS_NGRID * BOYS_SROW];
    // per-atom sums staged in LDS when they fit next to the Boys slice, else added
    // to the output directly (both exact integer adds)
    const bool staged = 6 * a.natm <= GRAD_LDS_WORDS;
    long long* lds = staged ? lds_acc : a.out;
    if (staged)
        for (int i = threadIdx.x; i < 6 * a.natm; i += blockDim.x) lds[i] = 0;
    boys_stage(a.boys, LA + LB + LC + LD + 1, btab);
    // items wave by wave: the 64 lanes of a wave take 64 consecutive selected
    // quartets and one pass / column group, so the pass dispatch is uniform
    constexpr int NTR = NT<Real>;
    const long long ns = *nsel, nitems = (ns + 63) / 64 * 64 * NTR;
    for (long long it = (long long)blockIdx.x * blockDim.x + threadIdx.x; it < nitems;
         it += (long long)gridDim.x * blockDim.x) {
        const long long g = it >> 6;
        const int j = (int)(g % NTR);
        const long long qi = g / NTR * 64 + (it & 63);
        if (qi >= ns) continue;
        const int2 pq = list[sel[qi]];
        if constexpr (FOLD<Real> > 0) {
            static_for<0, NTR>([&](auto jj) {
                constexpr int J = decltype(jj)::value;
                if (j == J) fold_item<J * FOLD<Real>, Real>(a, pq, btab, lds);
            });
            continue;
        }
        const int sA = a.pair_a[pq.x], sB = a.pair_b[pq.x], sC = a.pair_a[pq.y], sD = a.pair_b[pq.y];
        const int aA = a.shell_atom[sA], aB = a.shell_atom[sB], aC = a.shell_atom[sC], aD = a.shell_atom[sD];
        const bool need_bra = aA != aB, need_ket = aC != aD;
        const bool bra = j < NIB;
        if (bra && !(need_bra || (!need_ket && NAB >= NCD))) continue;
        if (!bra && !(need_ket || (!need_bra && NAB < NCD))) continue;
        double deg = 0.5;
        if (sA != sB) deg *= 2.0;
        if (sC != sD) deg *= 2.0;
        if (pq.x != pq.y) deg *= 2.0;
        const int oA = a.shell_ao[sA], oB = a.shell_ao[sB], oC = a.shell_ao[sC], oD = a.shell_ao[sD];
        const size_t rep = a.shell_rep ? (size_t)a.shell_rep[sA] : 0;
        const double* rb = a.recs + a.rec_off[pq.x];
        const double* rk = a.recs + a.rec_off[pq.y];
        double g1[3], g2[3];
        if (bra) {
            const int kc0 = j * NCB;
            Real gam[NAB * NCB];
            for (int ab = 0; ab < NAB; ++ab)
#pragma unroll
                for (int c = 0; c < NCB; ++c)
                    gam[ab * NCB + c] =
                        Real(gamma_el(a, rep, oA + ab / NB, oB + ab % NB, oC + (kc0 + c) / ND, oD + (kc0 + c) % ND, deg));
            grad_pass<LA, LB, LC, LD, NCB, true, Real>(rb, a.npp[pq.x], rk, a.npp[pq.y], btab, a.pos + 3 * aA,
                                                       a.pos + 3 * aB, kc0, gam, g1, g2);
            for (int d = 0; d < 3; ++d) {
                if (need_bra) {
                    lds_add(lds, aA, d, g1[d]);
                    lds_add(lds, aB, d, g2[d]);
                    if (!need_ket) lds_add(lds, aC, d, -(g1[d] + g2[d]));
                } else {
                    lds_add(lds, aA, d, g1[d] + g2[d]);
                    lds_add(lds, aC, d, -(g1[d] + g2[d]));
                }
            }
        } else {
            const int ab0 = (j - NIB) * NCK;
            Real gam[NCD * NCK];
            for (int cd = 0; cd < NCD; ++cd)
#pragma unroll
                for (int c = 0; c < NCK; ++c)
                    gam[cd * NCK + c] =
                        Real(gamma_el(a, rep, oA + (ab0 + c) / NB, oB + (ab0 + c) % NB, oC + cd / ND, oD + cd % ND, deg));
            grad_pass<LC, LD, LA, LB, NCK, true, Real>(rk, a.npp[pq.y], rb, a.npp[pq.x], btab, a.pos + 3 * aC,
                                                       a.pos + 3 * aD, ab0, gam, g1, g2);
            for (int d = 0; d < 3; ++d) {
                if (need_ket) {
                    lds_add(lds, aC, d, g1[d]);
                    lds_add(lds, aD, d, g2[d]);
                    if (!need_bra) lds_add(lds, aA, d, -(g1[d] + g2[d]));
                } else {
                    lds_add(lds, aC, d, g1[d] + g2[d]);
                    lds_add(lds, aA, d, -(g1[d] + g2[d]));
                }
            }
        }
    }
    if (staged) {
        __syncthreads();
        for (int i = threadIdx.x; i < 6 * a.natm; i += blockDim.x)
            if (lds[i]) atomicAdd((unsigned long long*)&a.out[i], (unsigned long long)lds[i]);
    }
}

}  // namespace

#define CAT2(a, b) a##b
#define CAT(a, b) CAT2(a, b)

// pre-pass + scan + ordered scatter (a.bcnt / a.boff hold 2 ints per 256 quartets)
void CAT(launch_grad_keep_, AIMDHF_QCLASS)(const GradArgs& a, const int2* list, long long nq, hipStream_t st) {
    if (nq <= 0) return;
    const int nb = (int)((nq + 255) / 256);
    k_grad_keep<QC><<<dim3(nb), dim3(256), 0, st>>>(a, list, nq);
    k_grad_scan<<<dim3(1), dim3(1024), 0, st>>>(a.bcnt, nb, a.boff, a.nsel);
    k_grad_scatter<<<dim3(nb), dim3(256), 0, st>>>(a.keep, nq, a.boff, a.sel64, a.sel32);
}

// FP64 items, then FP32 items (the grid is sized for the whole list; each kernel
// reads its selection count on the device)
void CAT(launch_grad_class_, AIMDHF_QCLASS)(const GradArgs& a, const int2* list, long long nq, hipStream_t st) {
    long long items = nq * NT_MAX;
    if (items <= 0) return;
    long long blocks = (items + 255) / 256;
    if (blocks > 1024) blocks = 1024;
    const size_t shm = 6 * a.natm <= GRAD_LDS_WORDS ? sizeof(long long) * 6 * a.natm : 0;
    k_grad_class<QC, double><<<dim3((unsigned)blocks), dim3(256), shm, st>>>(a, list, a.sel64, a.nsel);
    if (a.fp32_switch > 0.0)
        k_grad_class<QC, float><<<dim3((unsigned)blocks), dim3(256), shm, st>>>(a, list, a.sel32, a.nsel + 1);
}

}  // namespace aimdhf
