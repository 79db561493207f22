// gfx950 kernels of the Hartree-Fock force engine: primitive-pair build,
// Schwarz bounds, screened quartet lists, and the integral-direct J/K build.
//
// One thread evaluates one contracted shell quartet (class-templated, l <= 2)
// and digests it straight from registers into J and K, accumulated as exact
// fixed-point integers (md_core.h: fx_split), so the Fock build is bitwise
// reproducible from run to run whatever order the waves execute in.

#include "aimdhf_internal.h"
#include "md_eri.h"
#include "md_pairs.h"

namespace aimdhf {
using namespace md;

// ------------------------------------------------------------- pair build
// Primitive pairs kept iff mu |A-B|^2 <= 50 (aimd/qc/integrals.py:504-599).
// The test is evaluated without contractions so it is bit-identical to the
// CPU restatement.
__device__ __noinline__ bool keep_prim_pair(double a, double b, double r2) {
#pragma clang fp contract(off)
    return a * b / (a + b) * r2 <= 50.0;
}

__device__ __noinline__ double dist2(const double* A, const double* B) {
#pragma clang fp contract(off)
    double r2 = 0.0;
    for (int d = 0; d < 3; ++d) {
        double x = A[d] - B[d];
        r2 += x * x;
    }
    return r2;
}

// thread per shell pair: npp[P] and the records of the kept primitive pairs
__global__ void k_build_pairs(PairArgs g) {
    int P = blockIdx.x * blockDim.x + threadIdx.x;
    if (P >= g.npair) return;
    int sA = g.pair_a[P], sB = g.pair_b[P];
    int la = g.shell_l[sA], lb = g.shell_l[sB];
    const double* A = g.pos + 3 * g.shell_atom[sA];
    const double* B = g.pos + 3 * g.shell_atom[sB];
    double r2 = dist2(A, B);
    double* rec = g.recs + g.rec_off[P];
    int rs = g.rec_size[P];
    int n = 0;
    for (int ka = g.shell_prim[sA]; ka < g.shell_prim[sA + 1]; ++ka) {
        double a = g.prim_exp[ka];
        for (int kb = g.shell_prim[sB]; kb < g.shell_prim[sB + 1]; ++kb) {
            double b = g.prim_exp[kb];
            if (!keep_prim_pair(a, b, r2)) continue;
            double ca = g.prim_coef[ka], cb = g.prim_coef[kb];
            double* r = rec + n * rs;
            write_record_any(la, lb, a, b, ca, cb, A, B, r);
            ++n;
        }
    }
    g.npp[P] = n;
}

// ---------------------------------------------------------------- Schwarz
// One thread per (shell pair, component ab): the diagonal integral (ab|ab) from
// the ab column of (AB|AB); qshell[P] receives max |(ab|ab)| through an integer
// max on the bit pattern (non-negative doubles order like their bits, so the
// result does not depend on the order), and k_block_q takes the square root.
template <int LA, int LB>
__global__ __launch_bounds__(64) void k_schwarz(PairArgs g) {
    using Qt = Quartet<LA, LB, LA, LB>;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int np = g.ncls_pairs[pair_class(LA, LB)];
    if (t >= np * Qt::NCD) return;
    const int i = t / Qt::NCD, cd = t % Qt::NCD;
    const int P = g.cls_pairs[g.cls_pair_off[pair_class(LA, LB)] + i];
    const double* rec = g.recs + g.rec_off[P];
    double acc[Qt::NAB];
    eri_quartet_cd<LA, LB, LA, LB>(rec, g.npp[P], rec, g.npp[P], g.boys, cd, acc);
    atomicMax((unsigned long long*)&g.qshell[P], (unsigned long long)__double_as_longlong(fabs(acc[cd])));
}

// block-pair bound = max over its shell pairs; every shell pair screens with it
__global__ void k_block_q(PairArgs g) {
    int bp = blockIdx.x * blockDim.x + threadIdx.x;
    if (bp >= g.nbpair) return;
    double m = 0.0;
    for (int i = g.bp_off[bp]; i < g.bp_off[bp + 1]; ++i) m = fmax(m, g.qshell[g.bp_list[i]]);
    m = sqrt(m);
    g.qblock[bp] = m;
    for (int i = g.bp_off[bp]; i < g.bp_off[bp + 1]; ++i) g.qscr[g.bp_list[i]] = m;
}

// ------------------------------------------------------------ quartet lists
// Unique shell quartets (P >= Q) kept iff not (q_P q_Q < thr), the test of
// aimd/qc/integrals.py:902; binned by class with the heavier pair as bra.
__device__ __forceinline__ int qclass(int cp, int cq) {
    int hi = cp > cq ? cp : cq, lo = cp > cq ? cq : cp;
    return hi * (hi + 1) / 2 + lo;
}

// Within a class, quartets are further binned by precision (q_P q_Q >= sw
// first: FP64; then below sw: FP32 in mixed precision) and cost npp_P * npp_Q
// (kept primitive pairs), heaviest bin first, so the lanes of a wavefront run
// loops of similar length; within a bin the order is P-major, Q ascending.
__device__ __forceinline__ int cost_bin(int c) {
    return c > 256 ? 0 : c > 64 ? 1 : c > 16 ? 2 : c > 4 ? 3 : 4;
}

// list bin of a kept quartet: precision part, then cost bin
__device__ __forceinline__ int list_bin(double qq, double sw, int cost) {
    return (qq < sw ? NCOSTBIN : 0) + cost_bin(cost);
}

// One wavefront per pair P, lanes over Q <= P in chunks of 64: counts per
// (class, list bin) into counts[vc * npair + P], vc = class * NB + bin, with
// NB = NLBIN (precision split at sw) or NCOSTBIN (no split). Per (partner
// class, bin) histogram in LDS (integer adds: order-independent).
template <int NB>
__global__ __launch_bounds__(64) void k_count_quartets(PairArgs g, double thr, double sw, int* counts) {
    __shared__ int hist[6 * NB];
    const int P = blockIdx.x;
    const int lane = threadIdx.x;
    const double qp = g.qscr[P];
    const int cp = g.pair_cls[P];
    const int np = g.npp[P];
    const int lo = g.pair_lo ? g.pair_lo[P] : 0;
    for (int i = lane; i < 6 * NB; i += 64) hist[i] = 0;
    __syncthreads();
    for (int Q = lo + lane; Q <= P; Q += 64) {
        const double qq = qp * g.qscr[Q];
        if (qq < thr) continue;
        atomicAdd(&hist[g.pair_cls[Q] * NB + (NB == NLBIN ? list_bin(qq, sw, np * g.npp[Q]) : cost_bin(np * g.npp[Q]))],
                  1);
    }
    __syncthreads();
    for (int vc = lane; vc < NQCLASS * NB; vc += 64) {
        const int k = vc / NB, b = vc % NB;
        int v = 0;
        for (int cq = 0; cq < 6; ++cq)
            if (qclass(cp, cq) == k) v += hist[cq * NB + b];
        counts[(size_t)vc * g.npair + P] = v;
    }
}

// Exclusive scan of each (class, bin) column over P: one 1024-thread block per
// column; totals[vc] receives the column sum.
__global__ __launch_bounds__(1024) void k_scan_columns(const int* counts, int npair, long long* offs,
                                                       long long* totals) {
    __shared__ long long wsum[16];
    const int vc = blockIdx.x;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int* col = counts + (size_t)vc * npair;
    long long* out = offs + (size_t)vc * npair;
    long long run = 0;
    for (int base = 0; base < npair; base += 1024) {
        const int i = base + t;
        const long long x = i < npair ? col[i] : 0;
        long long v = x;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const long long y = __shfl_up(v, off, 64);
            if (lane >= off) v += y;
        }
        if (lane == 63) wsum[w] = v;
        __syncthreads();
        long long pre = 0;
        for (int j = 0; j < w; ++j) pre += wsum[j];
        long long tot = 0;
        for (int j = 0; j < 16; ++j) tot += wsum[j];
        if (i < npair) out[i] = run + pre + v - x;
        run += tot;
        __syncthreads();
    }
    if (t == 0) totals[vc] = run;
}

// base[vc] = exclusive scan of the column totals (class-major, bin-minor)
__global__ void k_scan_totals(const long long* totals, int n, long long* base) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        long long r = 0;
        for (int i = 0; i < n; ++i) {
            base[i] = r;
            r += totals[i];
        }
    }
}

// Ordered fill: per 64-partner chunk, one ballot per key present (the chunk's
// lanes grouped by (partner class, bin)), running offsets in LDS.
template <int NB>
__global__ __launch_bounds__(64) void k_fill_quartets(PairArgs g, double thr, double sw, const long long* offs,
                                                      const long long* base, int2* list) {
    __shared__ long long o[6 * NB];
    const int P = blockIdx.x;
    const int lane = threadIdx.x;
    const double qp = g.qscr[P];
    const int cp = g.pair_cls[P];
    const int np = g.npp[P];
    for (int i = lane; i < 6 * NB; i += 64) {
        const int cq = i / NB, b = i % NB;
        const int vc = qclass(cp, cq) * NB + b;
        o[i] = base[vc] + offs[(size_t)vc * g.npair + P];
    }
    __syncthreads();
    const unsigned long long below = (lane ? (~0ull >> (64 - lane)) : 0ull);
    for (int Q0 = g.pair_lo ? g.pair_lo[P] : 0; Q0 <= P; Q0 += 64) {
        const int Q = Q0 + lane;
        const double qq = Q <= P ? qp * g.qscr[Q] : 0.0;
        const bool in = Q <= P && !(qq < thr);
        const int key = in ? g.pair_cls[Q] * NB +
                                 (NB == NLBIN ? list_bin(qq, sw, np * g.npp[Q]) : cost_bin(np * g.npp[Q]))
                           : -1;
        unsigned long long active = __ballot(in);
        while (active) {
            const int kk = __shfl(key, __ffsll((long long)active) - 1, 64);
            const unsigned long long m = __ballot(key == kk);
            const long long at = o[kk];
            if (key == kk) {
                const int cq = kk / NB;
                list[at + __popcll(m & below)] = cp >= cq ? make_int2(P, Q) : make_int2(Q, P);
            }
            __syncthreads();
            if (lane == 0) o[kk] = at + __popcll(m);
            __syncthreads();
            active &= ~m;
        }
    }
}

// J = J' + J'^T, K = K' + K'^T from the fixed-point accumulators
__global__ void k_jk_finalize(const long long* fx, int nset, int n, double* J, double* K) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t nn = (size_t)n * n;
    if (i >= (size_t)nset * nn) return;
    size_t s = i / nn, r = (i % nn) / n, c = i % n;
    size_t ij = r * n + c, ji = c * n + r;
    const long long* Jh = fx + (4 * s + 0) * nn;
    const long long* Jl = fx + (4 * s + 1) * nn;
    const long long* Kh = fx + (4 * s + 2) * nn;
    const long long* Kl = fx + (4 * s + 3) * nn;
    J[i] = fx_value(Jh[ij] + Jh[ji], Jl[ij] + Jl[ji]);
    K[i] = fx_value(Kh[ij] + Kh[ji], Kl[ij] + Kl[ji]);
}

__global__ void k_symmetrize(const double* D, int nset, int n, double* out) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t nn = (size_t)n * n;
    if (i >= (size_t)nset * nn) return;
    size_t s = i / nn, r = (i % nn) / n, c = i % n;
    out[i] = 0.5 * (D[s * nn + r * n + c] + D[s * nn + c * n + r]);
}

// max |x| over n doubles: per-block maxima, then one block reduces them
__global__ void k_absmax_part(const double* x, size_t n, double* part) {
    __shared__ double sm[256];
    double m = 0.0;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        m = fmax(m, fabs(x[i]));
    sm[threadIdx.x] = m;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) sm[threadIdx.x] = fmax(sm[threadIdx.x], sm[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sm[0];
}

double device_absmax(const double* x, size_t n, double* scratch, hipStream_t st) {
    const int nb = 256;
    k_absmax_part<<<dim3(nb), dim3(256), 0, st>>>(x, n, scratch);
    k_absmax_part<<<dim3(1), dim3(256), 0, st>>>(scratch, nb, scratch + nb);
    double h = 0.0;
    if (hipMemcpyAsync(&h, scratch + nb, sizeof(double), hipMemcpyDeviceToHost, st) != hipSuccess) return -1.0;
    if (hipStreamSynchronize(st) != hipSuccess) return -1.0;
    return h;
}

__global__ void k_zero64(long long* p, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

// ------------------------------------------------------------- launchers
#define GRID(n, b) dim3((unsigned)(((n) + (b) - 1) / (b))), dim3(b)

void launch_build_pairs(const PairArgs& g, hipStream_t st) {
    if (g.npair) k_build_pairs<<<GRID(g.npair, 64), 0, st>>>(g);
}

void launch_schwarz(const PairArgs& g, const int* ncls_host, hipStream_t st) {
    if (g.npair) k_zero64<<<GRID(g.npair, 256), 0, st>>>((long long*)g.qshell, g.npair);
    if (ncls_host[0]) k_schwarz<0, 0><<<GRID(ncls_host[0] * 1, 64), 0, st>>>(g);
    if (ncls_host[1]) k_schwarz<1, 0><<<GRID(ncls_host[1] * 3, 64), 0, st>>>(g);
    if (ncls_host[2]) k_schwarz<1, 1><<<GRID(ncls_host[2] * 9, 64), 0, st>>>(g);
    if (ncls_host[3]) k_schwarz<2, 0><<<GRID(ncls_host[3] * 6, 64), 0, st>>>(g);
    if (ncls_host[4]) k_schwarz<2, 1><<<GRID(ncls_host[4] * 18, 64), 0, st>>>(g);
    if (ncls_host[5]) k_schwarz<2, 2><<<GRID(ncls_host[5] * 36, 64), 0, st>>>(g);
    if (g.nbpair) k_block_q<<<GRID(g.nbpair, 64), 0, st>>>(g);
}

// columns per class: NLBIN with a precision split (sw > 0), else NCOSTBIN
int list_bins(double sw) { return sw > 0.0 ? NLBIN : NCOSTBIN; }

void launch_count_quartets(const PairArgs& g, double thr, double sw, int* counts, long long* offs, long long* totals,
                           long long* base, hipStream_t st) {
    if (!g.npair) return;
    const int nb = list_bins(sw);
    if (nb == NLBIN) k_count_quartets<NLBIN><<<dim3(g.npair), dim3(64), 0, st>>>(g, thr, sw, counts);
    else k_count_quartets<NCOSTBIN><<<dim3(g.npair), dim3(64), 0, st>>>(g, thr, sw, counts);
    k_scan_columns<<<dim3(NQCLASS * nb), dim3(1024), 0, st>>>(counts, g.npair, offs, totals);
    k_scan_totals<<<dim3(1), dim3(1), 0, st>>>(totals, NQCLASS * nb, base);
}

void launch_fill_quartets(const PairArgs& g, double thr, double sw, const long long* offs, const long long* base,
                          int2* list, hipStream_t st) {
    if (!g.npair) return;
    if (list_bins(sw) == NLBIN) k_fill_quartets<NLBIN><<<dim3(g.npair), dim3(64), 0, st>>>(g, thr, sw, offs, base, list);
    else k_fill_quartets<NCOSTBIN><<<dim3(g.npair), dim3(64), 0, st>>>(g, thr, sw, offs, base, list);
}

// per shell pair (thread per pair): largest |D| over its block, for Dc and each Dx
__global__ void k_pair_dmax(GradArgs a, const int* shell_l, double* dmc, double* dmx) {
    const int P = blockIdx.x * blockDim.x + threadIdx.x;
    if (P >= a.npair) return;
    const int X = a.pair_a[P], Y = a.pair_b[P];
    const int lX = shell_l[X], lY = shell_l[Y];
    const int oX = a.shell_ao[X], nX = (lX + 1) * (lX + 2) / 2, oY = a.shell_ao[Y], nY = (lY + 1) * (lY + 2) / 2;
    const size_t n = a.ld, nn = n * n;
    const size_t r = a.shell_rep ? (size_t)a.shell_rep[X] : 0;
    const double* dc = a.dc + r * nn;
    const double* dx = a.dx + r * a.nx * nn;
    double m = 0.0, mx[2] = {0.0, 0.0};
    for (int i = 0; i < nX; ++i)
        for (int j = 0; j < nY; ++j) {
            const size_t e = (size_t)(oX + i) * n + oY + j;
            m = fmax(m, fabs(dc[e]));
            for (int x = 0; x < a.nx; ++x) mx[x] = fmax(mx[x], fabs(dx[x * nn + e]));
        }
    dmc[P] = m;
    for (int x = 0; x < a.nx; ++x) dmx[(size_t)x * a.npair + P] = mx[x];
}

void launch_pair_dmax(const GradArgs& a, const int* shell_l, double* dmc, double* dmx, hipStream_t st) {
    if (a.npair > 0) k_pair_dmax<<<GRID(a.npair, 256), 0, st>>>(a, shell_l, dmc, dmx);
}

void launch_zero64(long long* p, size_t n, hipStream_t st) {
    if (n) k_zero64<<<GRID(n, 256), 0, st>>>(p, n);
}

void launch_symmetrize(const double* D, int nset, int n, double* out, hipStream_t st) {
    size_t m = (size_t)nset * n * n;
    if (m) k_symmetrize<<<GRID(m, 256), 0, st>>>(D, nset, n, out);
}

void launch_jk_finalize(const long long* fx, int nset, int n, double* J, double* K, hipStream_t st) {
    size_t m = (size_t)nset * n * n;
    if (m) k_jk_finalize<<<GRID(m, 256), 0, st>>>(fx, nset, n, J, K);
}

}  // namespace aimdhf
