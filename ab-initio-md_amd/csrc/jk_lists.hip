// Per-geometry gather lists of the store-based J and K kernels. Both follow
// fixed enumerations, so the per-iteration sums have a fixed order:
//   K: for output block (X, Y) and each (l_J, l_L), the partners J of X in
//      descending-q order and, for each, the prefix of the partners L of Y
//      (descending q) with q_XJ q_YL >= threshold;
//   J: for pair P and each partner class, the prefix of the q-sorted pairs Q
//      of that class with q_P q_Q >= threshold.
// An entry packs the quartet's index in its class store (bits 0-31), two AO
// offsets (14 bits each at 32 and 46) and orientation flags (bits 60-62).
// In mixed precision the entries with q q >= sw (held in FP64) come first in
// each segment, those below (held in FP32) after them.

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

// per segment: counts[seg] entries with q q >= thr, of which big[seg] have q q >= sw
__global__ __launch_bounds__(64) void k_klist_count(ListArgs a, long long* counts, long long* big) {
    const int o = blockIdx.x;
    const int X = a.out_x[o], Y = a.out_y[o];
    const int lane = threadIdx.x;
    for (int sub = 0; sub < 9; ++sub) {
        const int LJ = sub / 3, LL = sub % 3;
        const int p0 = a.part_off[X * 3 + LJ], p1 = a.part_off[X * 3 + LJ + 1];
        const int r0 = a.part_off[Y * 3 + LL], r1 = a.part_off[Y * 3 + LL + 1];
        int cnt = 0, cb = 0;
        for (int jj = p0 + lane; jj < p1; jj += 64) {
            cnt += valid_prefix(a.part_q, r0, r1, a.part_q[jj], a.thr);
            cb += valid_prefix(a.part_q, r0, r1, a.part_q[jj], a.sw);
        }
        cnt = wave_isum(cnt);
        cb = wave_isum(cb);
        if (lane == 0) {
            counts[(size_t)o * 9 + sub] = cnt;
            big[(size_t)o * 9 + sub] = cb;
        }
    }
}

// Entries of a segment: first every (J, L) with q_XJ q_YL >= sw, then the rest
// (q q in [thr, sw)), each part in partner order (J by descending q, L prefix).
__global__ __launch_bounds__(64) void k_klist_fill(ListArgs a) {
    __shared__ int inb[64], ins[64], nbj[64];
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
        long long base_b = a.kseg[seg], base_s = a.kseg[seg] + a.kbig[seg];
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
            int vt = 0, vb = 0;
            if (jj < p1) {
                vt = valid_prefix(a.part_q, r0, r1, a.part_q[jj], a.thr);
                vb = valid_prefix(a.part_q, r0, r1, a.part_q[jj], a.sw);
            }
            int xb = vb, xs = vt - vb;
#pragma unroll
            for (int off = 1; off < 64; off <<= 1) {
                const int tb = __shfl_up(xb, off, 64), ts = __shfl_up(xs, off, 64);
                if (lane >= off) {
                    xb += tb;
                    xs += ts;
                }
            }
            inb[lane] = xb;
            ins[lane] = xs;
            nbj[lane] = vb;
            __syncthreads();
            const int tot_b = inb[63], tot_s = ins[63];
            for (int t = lane; t < tot_b; t += 64) {
                int lo = 0, hi = 63;
                while (lo < hi) {
                    const int mid = (lo + hi) >> 1;
                    if (inb[mid] > t) hi = mid;
                    else lo = mid + 1;
                }
                const int prev = lo ? inb[lo - 1] : 0;
                put(base_b + t, jb + lo, r0 + (t - prev));
            }
            for (int t = lane; t < tot_s; t += 64) {
                int lo = 0, hi = 63;
                while (lo < hi) {
                    const int mid = (lo + hi) >> 1;
                    if (ins[mid] > t) hi = mid;
                    else lo = mid + 1;
                }
                const int prev = lo ? ins[lo - 1] : 0;
                put(base_s + t, jb + lo, r0 + nbj[lo] + (t - prev));
            }
            base_b += tot_b;
            base_s += tot_s;
            __syncthreads();
        }
    }
}

// per (P, partner class): counts = partners with q_P q_Q >= thr, big = those >= sw (a prefix)
__global__ void k_jlist_count(ListArgs a, long long* counts, long long* big) {
    const int P = blockIdx.x * blockDim.x + threadIdx.x;
    if (P >= a.npair) return;
    const double qp = a.qscr[P];
    for (int c = 0; c < 6; ++c) {
        counts[(size_t)P * 6 + c] = valid_prefix(a.ket_q, a.ket_off[c], a.ket_off[c + 1], qp, a.thr);
        big[(size_t)P * 6 + c] = valid_prefix(a.ket_q, a.ket_off[c], a.ket_off[c + 1], qp, a.sw);
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

__global__ void k_fill_int(int* p, size_t n, int v) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}

}  // namespace

void launch_fill_int(int* p, size_t n, int v, hipStream_t st) {
    if (n) k_fill_int<<<dim3(2048), dim3(256), 0, st>>>(p, n, v);
}

void launch_klist_count(const ListArgs& a, int nout, long long* counts, long long* big, hipStream_t st) {
    if (nout > 0) k_klist_count<<<dim3(nout), dim3(64), 0, st>>>(a, counts, big);
}

void launch_klist_fill(const ListArgs& a, int nout, hipStream_t st) {
    if (nout > 0) k_klist_fill<<<dim3(nout), dim3(64), 0, st>>>(a);
}

void launch_jlist_count(const ListArgs& a, long long* counts, long long* big, hipStream_t st) {
    if (a.npair > 0) k_jlist_count<<<dim3((a.npair + 63) / 64), dim3(64), 0, st>>>(a, counts, big);
}

void launch_jlist_fill(const ListArgs& a, hipStream_t st) {
    if (a.npair > 0) k_jlist_fill<<<dim3(a.npair), dim3(64), 0, st>>>(a);
}

}  // namespace aimdhf
