// Per-geometry gather lists of the store-based J and K kernels. Both follow
// fixed enumerations, so the per-iteration sums have a fixed order:
//   K: for output block (X, Y) and each (l_J, l_L), the partners J of X in
//      descending-q order and, for each, the prefix of the partners L of Y
//      (descending q) with q_XJ q_YL >= threshold;
//   J: for pair P and each partner class, the prefix of the q-sorted pairs Q
//      of that class with q_P q_Q >= threshold.
// An entry packs the quartet's index in its class store (bits 0-31), two AO
// offsets (14 bits each at 32 and 46) and orientation flags (bits 60-62).
// Each segment is ordered by Schwarz-product band (q q in [band_lo[b],
// band_lo[b - 1])), each band in the order above; in mixed precision the
// leading bands are held in FP64 and the rest in FP32.

#include "aimdhf_internal.h"

namespace aimdhf {

namespace {

__device__ __forceinline__ int pcls(int a, int b) { return a > b ? a * (a + 1) / 2 + b : b * (b + 1) / 2 + a; }

__device__ __forceinline__ int wave_isum(int v) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
    return v;
}

// number of partners L (in [r0, r1), q descending) with qx * q_L >= thr
__device__ __forceinline__ int valid_prefix(const double* __restrict__ q, int r0, int r1, double qx, double thr) {
    int lo = r0, hi = r1;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (qx * q[mid] < thr) hi = mid;
        else lo = mid + 1;
    }
    return lo - r0;
}

__device__ __forceinline__ unsigned long long checked(const ListArgs& a, int k, int v) {
    if (v < 0 || v >= a.qcount[k]) {
        a.err[0] = 1;
        return 0;
    }
    return (unsigned long long)v;
}

__device__ __forceinline__ int qcls(int c1, int c2) { return c1 > c2 ? c1 * (c1 + 1) / 2 + c2 : c2 * (c2 + 1) / 2 + c1; }

__device__ __forceinline__ long long pair_key(int P, int Q) {
    const long long hi = P > Q ? P : Q, lo = P > Q ? Q : P;
    return hi * (hi + 1) / 2 + lo;
}

// per segment: counts[seg] entries with q q >= thr, band[seg * NBAND + b] of them in band b
__global__ __launch_bounds__(64) void k_klist_count(ListArgs a, long long* counts, long long* band) {
    const int o = blockIdx.x;
    const int X = a.out_x[o], Y = a.out_y[o];
    const int lane = threadIdx.x;
    for (int sub = 0; sub < 9; ++sub) {
        const int LJ = sub / 3, LL = sub % 3;
        const int p0 = a.part_off[X * 3 + LJ], p1 = a.part_off[X * 3 + LJ + 1];
        const int r0 = a.part_off[Y * 3 + LL], r1 = a.part_off[Y * 3 + LL + 1];
        int cb[NBAND];
        for (int b = 0; b < NBAND; ++b) cb[b] = 0;
        for (int jj = p0 + lane; jj < p1; jj += 64) {
            int prev = 0;
            for (int b = 0; b < NBAND; ++b) {
                const int n = valid_prefix(a.part_q, r0, r1, a.part_q[jj], a.band_lo[b]);
                cb[b] += n - prev;
                prev = n;
            }
        }
        int cnt = 0;
        for (int b = 0; b < NBAND; ++b) {
            cb[b] = wave_isum(cb[b]);
            cnt += cb[b];
        }
        if (lane == 0) {
            counts[(size_t)o * 9 + sub] = cnt;
            for (int b = 0; b < NBAND; ++b) band[((size_t)o * 9 + sub) * NBAND + b] = cb[b];
        }
    }
}

// Entries of a segment, band by band; within a band every (J, L) with q_XJ
// q_YL in the band, J by descending q and, for each J, its L in descending q
// (a contiguous run of Y's sorted partners).
__global__ __launch_bounds__(64) void k_klist_fill(ListArgs a) {
    __shared__ int incl[NBAND][64], lbeg[NBAND][64];
    const int o = blockIdx.x;
    const int X = a.out_x[o], Y = a.out_y[o];
    const int lX = a.shell_l[X], lY = a.shell_l[Y];
    const int lane = threadIdx.x;
    for (int sub = 0; sub < 9; ++sub) {
        const int LJ = sub / 3, LL = sub % 3;
        const int p0 = a.part_off[X * 3 + LJ], p1 = a.part_off[X * 3 + LJ + 1];
        const int r0 = a.part_off[Y * 3 + LL], r1 = a.part_off[Y * 3 + LL + 1];
        const int c1 = pcls(lX, LJ), c2 = pcls(lY, LL);
        const size_t seg = (size_t)o * 9 + sub;
        long long base[NBAND];
        base[0] = a.kseg[seg];
        for (int b = 1; b < NBAND; ++b) base[b] = base[b - 1] + a.kband[seg * NBAND + b - 1];
        auto put = [&](long long pos, int jidx, int lidx) {
            const int J = a.part_shell[jidx], PXJ = a.part_pair[jidx];
            const int L = a.part_shell[lidx], PYL = a.part_pair[lidx];
            const unsigned long long qi = checked(a, qcls(c1, c2), a.qmap[pair_key(PXJ, PYL)]);
            const bool fb = (lX == LJ) ? (a.pair_a[PXJ] != X) : (lX < LJ);
            const bool fk = (lY == LL) ? (a.pair_a[PYL] != Y) : (lY < LL);
            const bool sw = (c1 == c2) ? (PXJ < PYL) : (c1 < c2);
            a.klist[pos] = (long long)(qi | ((unsigned long long)a.shell_ao[J] << 32) |
                                       ((unsigned long long)a.shell_ao[L] << 46) | ((unsigned long long)sw << 60) |
                                       ((unsigned long long)fb << 61) | ((unsigned long long)fk << 62));
        };
        for (int jb = p0; jb < p1; jb += 64) {
            const int jj = jb + lane;
            // this J's run of L in each band: [prev prefix, prefix at the band's lower bound)
            int prev = 0;
            for (int b = 0; b < NBAND; ++b) {
                const int n = jj < p1 ? valid_prefix(a.part_q, r0, r1, a.part_q[jj], a.band_lo[b]) : 0;
                int x = n - prev;
#pragma unroll
                for (int off = 1; off < 64; off <<= 1) {
                    const int t = __shfl_up(x, off, 64);
                    if (lane >= off) x += t;
                }
                incl[b][lane] = x;
                lbeg[b][lane] = prev;
                prev = n;
            }
            __syncthreads();
            for (int b = 0; b < NBAND; ++b) {
                const int tot = incl[b][63];
                for (int t = lane; t < tot; t += 64) {
                    int lo = 0, hi = 63;
                    while (lo < hi) {
                        const int mid = (lo + hi) >> 1;
                        if (incl[b][mid] > t) hi = mid;
                        else lo = mid + 1;
                    }
                    const int before = lo ? incl[b][lo - 1] : 0;
                    put(base[b] + t, jb + lo, r0 + lbeg[b][lo] + (t - before));
                }
                base[b] += tot;
            }
            __syncthreads();
        }
    }
}

// per (P, partner class): counts = partners with q_P q_Q >= thr, band = those
// in each band (the partners are in descending q, so each band is a run)
__global__ void k_jlist_count(ListArgs a, long long* counts, long long* band) {
    const int P = blockIdx.x * blockDim.x + threadIdx.x;
    if (P >= a.npair) return;
    const double qp = a.qscr[P];
    for (int c = 0; c < 6; ++c) {
        int prev = 0;
        for (int b = 0; b < NBAND; ++b) {
            const int n = valid_prefix(a.ket_q, a.ket_off[c], a.ket_off[c + 1], qp, a.band_lo[b]);
            band[((size_t)P * 6 + c) * NBAND + b] = n - prev;
            prev = n;
        }
        counts[(size_t)P * 6 + c] = prev;
    }
}

__global__ __launch_bounds__(64) void k_jlist_fill(ListArgs a) {
    const int P = blockIdx.x;
    const int cp = a.pair_cls[P];
    for (int c = 0; c < 6; ++c) {
        const long long e0 = a.jseg[(size_t)P * 6 + c], e1 = a.jseg[(size_t)P * 6 + c + 1];
        for (long long i = e0 + threadIdx.x; i < e1; i += 64) {
            const int Q = a.ket_pair[a.ket_off[c] + (i - e0)];
            const unsigned long long qi = checked(a, qcls(cp, c), a.qmap[pair_key(P, Q)]);
            const int C = a.pair_a[Q], D = a.pair_b[Q];
            const bool sw = (cp == c) ? (P < Q) : (cp < c);
            const unsigned long long ent = qi | ((unsigned long long)a.shell_ao[C] << 32) |
                                           ((unsigned long long)a.shell_ao[D] << 46) |
                                           ((unsigned long long)sw << 60) | ((unsigned long long)(C != D) << 61);
            a.jlist[i] = (long long)ent;
        }
    }
}

__device__ __forceinline__ double wave_dmax(double v) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v = fmax(v, __shfl_xor(v, m, 64));
    return v;
}

// largest q_P q_Q over the entries of each K unit (one wavefront per unit)
__global__ __launch_bounds__(64) void k_kunit_q(UnitQArgs a) {
    const int u = blockIdx.x;
    const int code = a.kunit_os[u];
    const int o = code >> 5, sub = code & 15;
    const int k = qcls(pcls(a.shell_l[a.out_x[o]], sub / 3), pcls(a.shell_l[a.out_y[o]], sub % 3));
    const int2* list = a.list + a.qoff[k];
    double m = 0.0;
    for (long long e = a.kunit_e[2 * u] + threadIdx.x; e < a.kunit_e[2 * u + 1]; e += 64) {
        const int2 q = list[(unsigned long long)a.klist[e] & 0xffffffffull];
        m = fmax(m, a.qscr[q.x] * a.qscr[q.y]);
    }
    m = wave_dmax(m);
    if (threadIdx.x == 0) a.kunit_q[u] = m;
}

// the same for the J units
__global__ __launch_bounds__(64) void k_junit_q(UnitQArgs a) {
    const int u = blockIdx.x;
    const int code = a.junit_pc[u];
    const int k = qcls(a.pair_cls[code >> 4], code & 7);
    const int2* list = a.list + a.qoff[k];
    double m = 0.0;
    for (long long e = a.junit_e[2 * u] + threadIdx.x; e < a.junit_e[2 * u + 1]; e += 64) {
        const int2 q = list[(unsigned long long)a.jlist[e] & 0xffffffffull];
        m = fmax(m, a.qscr[q.x] * a.qscr[q.y]);
    }
    m = wave_dmax(m);
    if (threadIdx.x == 0) a.junit_q[u] = m;
}

// qmap entry of every listed quartet (thread per quartet; class by its offset)
__global__ void k_qmap(QmapArgs a) {
    const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= a.qoff[21]) return;
    int k = 0;
    while (k < 20 && g >= a.qoff[k + 1]) ++k;
    const int2 pq = a.list[g];
    a.qmap[pair_key(pq.x, pq.y)] = (int)(g - a.qoff[k]);
}

// out[0] = max |x| (one block; the max does not depend on the order)
__global__ __launch_bounds__(1024) void k_absmax(const double* __restrict__ x, size_t n, double* out) {
    __shared__ double part[16];
    double m = 0.0;
    for (size_t i = threadIdx.x; i < n; i += 1024) m = fmax(m, fabs(x[i]));
    m = wave_dmax(m);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x < 64) {
        m = threadIdx.x < 16 ? part[threadIdx.x] : 0.0;
        m = wave_dmax(m);
        if (threadIdx.x == 0) out[0] = m;
    }
}

__global__ void k_fill_int(int* p, size_t n, int v) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}

}  // namespace

void launch_fill_int(int* p, size_t n, int v, hipStream_t st) {
    if (n) k_fill_int<<<dim3(2048), dim3(256), 0, st>>>(p, n, v);
}

void launch_unit_q(const UnitQArgs& a, hipStream_t st) {
    if (a.nkunit > 0) k_kunit_q<<<dim3(a.nkunit), dim3(64), 0, st>>>(a);
    if (a.njunit > 0) k_junit_q<<<dim3(a.njunit), dim3(64), 0, st>>>(a);
}

void launch_qmap(const QmapArgs& a, hipStream_t st) {
    const long long n = a.qoff[21];
    if (n > 0) k_qmap<<<dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st>>>(a);
}

void launch_absmax(const double* x, size_t n, double* out, hipStream_t st) {
    k_absmax<<<dim3(1), dim3(1024), 0, st>>>(x, n, out);
}

void launch_klist_count(const ListArgs& a, int nout, long long* counts, long long* band, hipStream_t st) {
    if (nout > 0) k_klist_count<<<dim3(nout), dim3(64), 0, st>>>(a, counts, band);
}

void launch_klist_fill(const ListArgs& a, int nout, hipStream_t st) {
    if (nout > 0) k_klist_fill<<<dim3(nout), dim3(64), 0, st>>>(a);
}

void launch_jlist_count(const ListArgs& a, long long* counts, long long* band, hipStream_t st) {
    if (a.npair > 0) k_jlist_count<<<dim3((a.npair + 63) / 64), dim3(64), 0, st>>>(a, counts, band);
}

void launch_jlist_fill(const ListArgs& a, hipStream_t st) {
    if (a.npair > 0) k_jlist_fill<<<dim3(a.npair), dim3(64), 0, st>>>(a);
}

}  // namespace aimdhf
