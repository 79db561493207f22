// Per-geometry tables of the integral store's gather segments, built on the
// device from the per-band entry counts of the list builders: entry offsets,
// FP64 entry counts, tiled data offsets of both precisions and the work units
// (each segment band cut into pieces of at most unit_items lane items). Only
// a handful of totals go back to the host (for allocation sizes and launch
// grids). The scans are three-pass (block scans, a scan of the block totals,
// the add) with a fixed association, so the tables never depend on timing.

#include <algorithm>

#include "aimdhf_internal.h"

namespace aimdhf {

namespace {

constexpr int SCAN_T = 1024, SCAN_V = 4, SCAN_B = SCAN_T * SCAN_V;   // items per scan block

__device__ __forceinline__ int ncart(int l) { return (l + 1) * (l + 2) / 2; }
__device__ __forceinline__ int ncp(int c) {   // Cartesian components of a pair class
    const int la = c == 0 ? 0 : c <= 2 ? 1 : 2, lb = c - la * (la + 1) / 2;
    return ncart(la) * ncart(lb);
}
// a part of n entries: n * ns lane items of blk / ns values, in tiles of 64 items
__device__ __forceinline__ long long tiled(long long n, long long blk, long long ns) {
    return (n * ns + 63) / 64 * 64 * (blk / ns);
}

// exclusive scan of v over the block's threads (thread order); total = block sum
__device__ long long block_scan(long long v, long long* wsum, long long& total) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    long long x = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const long long t = __shfl_up(x, off, 64);
        if (lane >= off) x += t;
    }
    if (lane == 63) wsum[w] = x;
    __syncthreads();
    if (w == 0) {
        long long s = lane < SCAN_T / 64 ? wsum[lane] : 0;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const long long t = __shfl_up(s, off, 64);
            if (lane >= off) s += t;
        }
        if (lane < SCAN_T / 64) wsum[lane] = s;   // inclusive wave prefix
    }
    __syncthreads();
    total = wsum[SCAN_T / 64 - 1];
    const long long before = w ? wsum[w - 1] : 0;
    __syncthreads();
    return before + x - v;
}

// out[i] = sum of in[j < i] within each block of SCAN_B items; bsum[block] = block total
__global__ __launch_bounds__(SCAN_T) void k_scan_blocks(const long long* __restrict__ in, long long* out, long long n,
                                                       long long* bsum) {
    __shared__ long long wsum[SCAN_T / 64];
    const long long i0 = (long long)blockIdx.x * SCAN_B + threadIdx.x * SCAN_V;
    long long v[SCAN_V], s = 0;
#pragma unroll
    for (int k = 0; k < SCAN_V; ++k) {
        v[k] = i0 + k < n ? in[i0 + k] : 0;
        s += v[k];
    }
    long long total;
    long long p = block_scan(s, wsum, total);
#pragma unroll
    for (int k = 0; k < SCAN_V; ++k) {
        if (i0 + k < n) out[i0 + k] = p;
        p += v[k];
    }
    if (threadIdx.x == 0) bsum[blockIdx.x] = total;
}

// exclusive scan of the nb <= SCAN_B block totals in place; out[n] = grand total
__global__ __launch_bounds__(SCAN_T) void k_scan_totals(long long* bsum, int nb, long long* out, long long n) {
    __shared__ long long wsum[SCAN_T / 64];
    const int i0 = threadIdx.x * SCAN_V;
    long long v[SCAN_V], s = 0;
#pragma unroll
    for (int k = 0; k < SCAN_V; ++k) {
        v[k] = i0 + k < nb ? bsum[i0 + k] : 0;
        s += v[k];
    }
    long long total;
    long long p = block_scan(s, wsum, total);
#pragma unroll
    for (int k = 0; k < SCAN_V; ++k) {
        if (i0 + k < nb) bsum[i0 + k] = p;
        p += v[k];
    }
    if (threadIdx.x == 0) out[n] = total;
}

__global__ __launch_bounds__(256) void k_scan_add(long long* out, long long n, const long long* bsum) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] += bsum[i / SCAN_B];
}

// K segment i = o * 9 + sub: entry count, FP64 count, tiled sizes, units per band
__global__ void k_kseg_sizes(TabArgs a) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long long)a.nout * 9) return;
    const int o = (int)(i / 9), sub = (int)(i % 9);
    const int nx = ncart(a.shell_l[a.out_x[o]]), ny = ncart(a.shell_l[a.out_y[o]]);
    const int nj = ncart(sub / 3), nl = ncart(sub % 3);
    const int blk = nx * nj * ny * nl, ns = blk > 36 ? ny * nl : 1;
    const long long per = a.unit_items / ns > 0 ? a.unit_items / ns : 1;
    long long cnt = 0, big = 0;
    for (int b = 0; b < NBAND; ++b) {
        const long long nb = a.kband[i * NBAND + b];
        cnt += nb;
        if (b < a.nbig) big += nb;
        a.knu[i * NBAND + b] = (nb + per - 1) / per;
    }
    a.kcount[i] = cnt;
    a.kbig[i] = big;
    a.ksz64[i] = tiled(big, blk, ns);
    a.ksz32[i] = tiled(cnt - big, blk, ns);
}

// J segment of the pair at class-order position idx and partner class cq; the
// sizes are indexed by pair (P * 6 + cq), the units by class order
__global__ void k_jseg_sizes(TabArgs a) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)a.npair * 6) return;
    const int idx = (int)(t / 6), cq = (int)(t % 6);
    const int P = a.cls_pairs[idx];
    const long long i = (long long)P * 6 + cq;
    const int blk = ncp(a.pair_cls[P]) * ncp(cq), ns = blk > 36 ? ncp(cq) : 1;
    const long long per = a.unit_items / ns > 0 ? a.unit_items / ns : 1;
    long long cnt = 0, big = 0;
    for (int b = 0; b < NBAND; ++b) {
        const long long nb = a.jband[i * NBAND + b];
        cnt += nb;
        if (b < a.nbig) big += nb;
        a.jnu[t * NBAND + b] = (nb + per - 1) / per;
    }
    a.jcount[i] = cnt;
    a.jbig[i] = big;
    a.jsz64[i] = tiled(big, blk, ns);
    a.jsz32[i] = tiled(cnt - big, blk, ns);
}

// the units of K segment band (i, b); out_unit[o] = first unit of output o
__global__ void k_kunits(TabArgs a) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long nks = (long long)a.nout * 9;
    if (t > nks * NBAND) return;
    if (t == nks * NBAND) {
        a.out_unit[a.nout] = (int)a.kustart[t];
        return;
    }
    const long long i = t / NBAND;
    const int b = (int)(t % NBAND);
    const int o = (int)(i / 9), sub = (int)(i % 9);
    if (sub == 0 && b == 0) a.out_unit[o] = (int)a.kustart[t];
    const long long u0 = a.kustart[t], nu = a.kustart[t + 1] - u0;
    if (nu == 0) return;
    const int ny = ncart(a.shell_l[a.out_y[o]]), nx = ncart(a.shell_l[a.out_x[o]]);
    const int nj = ncart(sub / 3), nl = ncart(sub % 3);
    const int blk = nx * nj * ny * nl, ns = blk > 36 ? ny * nl : 1;
    const long long per = a.unit_items / ns > 0 ? a.unit_items / ns : 1;
    long long e0 = a.kseg[i];
    for (int bb = 0; bb < b; ++bb) e0 += a.kband[i * NBAND + bb];
    const long long e1 = e0 + a.kband[i * NBAND + b];
    const int code = o << 5 | (b >= a.nbig) << 4 | sub;
    for (long long u = 0; u < nu; ++u) {
        a.kunit_os[u0 + u] = code;
        a.kunit_e[2 * (u0 + u)] = e0 + u * per;
        a.kunit_e[2 * (u0 + u) + 1] = e0 + (u + 1) * per < e1 ? e0 + (u + 1) * per : e1;
    }
}

// the units of J segment band (idx, cq, b) in class order; pair_unit[idx] = first unit of position idx
__global__ void k_junits(TabArgs a) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long nseg = (long long)a.npair * 6;
    if (t > nseg * NBAND) return;
    if (t == nseg * NBAND) {
        a.pair_unit[a.npair] = (int)a.justart[t];
        return;
    }
    const long long s = t / NBAND;
    const int b = (int)(t % NBAND);
    const int idx = (int)(s / 6), cq = (int)(s % 6);
    if (cq == 0 && b == 0) a.pair_unit[idx] = (int)a.justart[t];
    const long long u0 = a.justart[t], nu = a.justart[t + 1] - u0;
    if (nu == 0) return;
    const int P = a.cls_pairs[idx];
    const long long i = (long long)P * 6 + cq;
    const int blk = ncp(a.pair_cls[P]) * ncp(cq), ns = blk > 36 ? ncp(cq) : 1;
    const long long per = a.unit_items / ns > 0 ? a.unit_items / ns : 1;
    long long e0 = a.jseg[i];
    for (int bb = 0; bb < b; ++bb) e0 += a.jband[i * NBAND + bb];
    const long long e1 = e0 + a.jband[i * NBAND + b];
    const int code = P << 4 | (b >= a.nbig) << 3 | cq;
    for (long long u = 0; u < nu; ++u) {
        a.junit_pc[u0 + u] = code;
        a.junit_e[2 * (u0 + u)] = e0 + u * per;
        a.junit_e[2 * (u0 + u) + 1] = e0 + (u + 1) * per < e1 ? e0 + (u + 1) * per : e1;
    }
}

// consistency of the scanned tables with their sizes: any mismatch sets *bad
__global__ void k_tab_check(TabArgs a, long long* bad) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long nks = (long long)a.nout * 9, njs = (long long)a.npair * 6;
    bool ok = true;
    if (t < nks)
        ok = a.kseg[t + 1] - a.kseg[t] == a.kcount[t] && a.kdoff[t + 1] - a.kdoff[t] == a.ksz64[t] &&
             a.kdoff32[t + 1] - a.kdoff32[t] == a.ksz32[t] && a.kbig[t] <= a.kcount[t] && a.kcount[t] >= 0;
    if (t < njs)
        ok = ok && a.jseg[t + 1] - a.jseg[t] == a.jcount[t] && a.jdoff[t + 1] - a.jdoff[t] == a.jsz64[t] &&
             a.jdoff32[t + 1] - a.jdoff32[t] == a.jsz32[t] && a.jbig[t] <= a.jcount[t] && a.jcount[t] >= 0;
    if (t < nks * NBAND) ok = ok && a.kustart[t + 1] - a.kustart[t] == a.knu[t] && a.knu[t] >= 0;
    if (t < njs * NBAND) ok = ok && a.justart[t + 1] - a.justart[t] == a.jnu[t] && a.jnu[t] >= 0;
    if (t == 0) ok = ok && a.kseg[0] == 0 && a.jseg[0] == 0 && a.kustart[0] == 0 && a.justart[0] == 0;
    if (!ok) *bad = 1;
}

// the totals the host needs, gathered into one small array
__global__ void k_tab_totals(TabArgs a, long long* out) {
    const long long nks = (long long)a.nout * 9, njs = (long long)a.npair * 6;
    out[0] = a.kseg[nks];
    out[1] = a.jseg[njs];
    out[2] = a.kdoff[nks];
    out[3] = a.kdoff32[nks];
    out[4] = a.jdoff[njs];
    out[5] = a.jdoff32[njs];
    out[6] = a.kustart[nks * NBAND];
    out[7] = a.justart[njs * NBAND];
    for (int t = 0; t < 10; ++t) out[8 + t] = a.kustart[(long long)a.out_off[t] * 9 * NBAND];
    for (int c = 0; c < 7; ++c) out[18 + c] = a.justart[(long long)a.cls_pair_off[c] * 6 * NBAND];
}

__global__ void k_zero_flag(long long* p) { p[0] = 0; }

int grid(long long n, int b) { return (int)((n + b - 1) / b); }

}  // namespace

void launch_scan(const long long* in, long long* out, long long n, long long* bsum, hipStream_t st) {
    const int nb = grid(n, SCAN_B);
    if (n > 0) k_scan_blocks<<<dim3(nb), dim3(SCAN_T), 0, st>>>(in, out, n, bsum);
    k_scan_totals<<<dim3(1), dim3(SCAN_T), 0, st>>>(bsum, nb, out, n);
    if (n > 0) k_scan_add<<<dim3(grid(n, 256)), dim3(256), 0, st>>>(out, n, bsum);
}

long long scan_max_items() { return (long long)SCAN_B * SCAN_B; }

void launch_seg_sizes(const TabArgs& a, hipStream_t st) {
    const long long nks = (long long)a.nout * 9, njs = (long long)a.npair * 6;
    if (nks) k_kseg_sizes<<<dim3(grid(nks, 256)), dim3(256), 0, st>>>(a);
    if (njs) k_jseg_sizes<<<dim3(grid(njs, 256)), dim3(256), 0, st>>>(a);
}

void launch_units(const TabArgs& a, hipStream_t st) {
    const long long nk = (long long)a.nout * 9 * NBAND + 1, nj = (long long)a.npair * 6 * NBAND + 1;
    k_kunits<<<dim3(grid(nk, 256)), dim3(256), 0, st>>>(a);
    k_junits<<<dim3(grid(nj, 256)), dim3(256), 0, st>>>(a);
}

// out[0..24] totals, out[25] = 1 if a table failed its check (else 0)
void launch_tab_totals(const TabArgs& a, long long* out, hipStream_t st) {
    const long long n = std::max((long long)a.nout * 9, (long long)a.npair * 6) * NBAND;
    k_tab_totals<<<dim3(1), dim3(1), 0, st>>>(a, out);
    k_zero_flag<<<dim3(1), dim3(1), 0, st>>>(out + 25);
    k_tab_check<<<dim3(grid(n, 256)), dim3(256), 0, st>>>(a, out + 25);
}

}  // namespace aimdhf
