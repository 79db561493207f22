// Exchange kernels for output blocks (X, Y) of angular momenta
// (AIMDHF_LX, AIMDHF_LY) (compiled once per type): one wave per output block,
// lanes over the screened (J, L) partner pairs of each (l_J, l_L) sub-class,
// enumerated from per-shell partner lists sorted by Schwarz bound (for each J
// the valid L form a prefix). Per-lane blocks are summed by a fixed butterfly
// and written to the (l_J, l_L) partial slice.

#include "aimdhf_internal.h"
#include "md_jk2.h"

#if !defined(AIMDHF_LX) || !defined(AIMDHF_LY)
#error "compile with -DAIMDHF_LX=<0..2> -DAIMDHF_LY=<0..2>"
#endif

namespace aimdhf {
using namespace md;

namespace {
constexpr int LX = AIMDHF_LX, LY = AIMDHF_LY;
constexpr int NX = NC(LX), NY = NC(LY);
constexpr int OT = LX * 3 + LY;

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
    return v;
}

template <int LJ, int LL, bool FB, bool FK, int NSET>
__device__ __forceinline__ void k_item(const KOArgs& a, int PXJ, int PYL, int J, int L, int cd, double* Kacc) {
    constexpr int LA = FB ? LJ : LX, LB = FB ? LX : LJ, LC = FK ? LL : LY, LD = FK ? LY : LL;
    using Qt = Quartet<LA, LB, LC, LD>;
    const double* rb = a.recs + a.rec_off[PXJ];
    const double* rk = a.recs + a.rec_off[PYL];
    const size_t n = a.nao, nn = n * n;
    if constexpr (split_by_cd<LA, LB, LC, LD>()) {
        double acc[Qt::NAB];
        eri_quartet_cd<LA, LB, LC, LD>(rb, a.npp[PXJ], rk, a.npp[PYL], a.boys, cd, acc);
        k_digest_cd<LX, LJ, LY, LL, FB, FK>(acc, cd, a.dens, n, nn, NSET, a.shell_ao[J], a.shell_ao[L], Kacc);
    } else {
        double acc[Qt::NAB * Qt::NCD];
        eri_quartet<LA, LB, LC, LD>(rb, a.npp[PXJ], rk, a.npp[PYL], a.boys, acc);
        k_digest_full<LX, LJ, LY, LL, FB, FK>(acc, a.dens, n, nn, NSET, a.shell_ao[J], a.shell_ao[L], Kacc);
    }
}

template <int LJ, int LL>
constexpr int nsplit() {
    constexpr int LA = LX > LJ ? LX : LJ, LB = LX > LJ ? LJ : LX;
    constexpr int LC = LY > LL ? LY : LL, LD = LY > LL ? LL : LY;
    return split_by_cd<LA, LB, LC, LD>() ? NC(LC) * NC(LD) : 1;
}

template <int LX_, int LY_, int LJ, int LL, int NSET>
__global__ __launch_bounds__(64) void k_kown(KOArgs a) {
    constexpr int NS = nsplit<LJ, LL>();
    __shared__ int incl[64];
    const int o = a.out_off[OT] + blockIdx.x;
    const int X = a.out_x[o], Y = a.out_y[o];
    const int lane = threadIdx.x;
    const int p0 = a.part_off[X * 3 + LJ], p1 = a.part_off[X * 3 + LJ + 1];
    const int r0 = a.part_off[Y * 3 + LL], r1 = a.part_off[Y * 3 + LL + 1];
    double Kacc[NSET * NX * NY];
#pragma unroll
    for (int i = 0; i < NSET * NX * NY; ++i) Kacc[i] = 0.0;
    for (int jb = p0; jb < p1; jb += 64) {
        const int jj = jb + lane;
        int nl = 0;
        if (jj < p1) {
            const double qxj = a.part_q[jj];
            int lo = r0, hi = r1;
            while (lo < hi) {
                int mid = (lo + hi) >> 1;
                if (qxj * a.part_q[mid] < a.thr) hi = mid;
                else lo = mid + 1;
            }
            nl = lo - r0;
        }
        int v = nl;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            int t = __shfl_up(v, off, 64);
            if (lane >= off) v += t;
        }
        incl[lane] = v;
        __syncthreads();
        const int total = incl[63];
        for (int t = lane; t < total * NS; t += 64) {
            const int item = t / NS, cd = t % NS;
            int lo = 0, hi = 63;
            while (lo < hi) {
                int mid = (lo + hi) >> 1;
                if (incl[mid] > item) hi = mid;
                else lo = mid + 1;
            }
            const int prev = lo ? incl[lo - 1] : 0;
            const int jidx = jb + lo, lidx = r0 + (item - prev);
            const int J = a.part_shell[jidx], PXJ = a.part_pair[jidx];
            const int L = a.part_shell[lidx], PYL = a.part_pair[lidx];
            const bool fb = (LX == LJ) ? (a.pair_a[PXJ] != X) : (LX < LJ);
            const bool fk = (LY == LL) ? (a.pair_a[PYL] != Y) : (LY < LL);
            if constexpr (LX == LJ && LY == LL) {
                if (!fb && !fk) k_item<LJ, LL, false, false, NSET>(a, PXJ, PYL, J, L, cd, Kacc);
                else if (fb && !fk) k_item<LJ, LL, true, false, NSET>(a, PXJ, PYL, J, L, cd, Kacc);
                else if (!fb && fk) k_item<LJ, LL, false, true, NSET>(a, PXJ, PYL, J, L, cd, Kacc);
                else k_item<LJ, LL, true, true, NSET>(a, PXJ, PYL, J, L, cd, Kacc);
            } else if constexpr (LX == LJ) {
                constexpr bool FK = LY < LL;
                if (!fb) k_item<LJ, LL, false, FK, NSET>(a, PXJ, PYL, J, L, cd, Kacc);
                else k_item<LJ, LL, true, FK, NSET>(a, PXJ, PYL, J, L, cd, Kacc);
            } else if constexpr (LY == LL) {
                constexpr bool FB = LX < LJ;
                if (!fk) k_item<LJ, LL, FB, false, NSET>(a, PXJ, PYL, J, L, cd, Kacc);
                else k_item<LJ, LL, FB, true, NSET>(a, PXJ, PYL, J, L, cd, Kacc);
            } else {
                k_item<LJ, LL, (LX < LJ), (LY < LL), NSET>(a, PXJ, PYL, J, L, cd, Kacc);
            }
        }
        __syncthreads();
    }
    const size_t n = a.nao, nn = n * n;
    double* out = a.kpart + (size_t)(LJ * 3 + LL) * a.nset * nn;
    const int oX = a.shell_ao[X], oY = a.shell_ao[Y];
#pragma unroll
    for (int i = 0; i < NSET * NX * NY; ++i) {
        double v = wave_sum(Kacc[i]);
        const int s = i / (NX * NY), xy = i % (NX * NY);
        if (lane == 0) out[s * nn + (size_t)(oX + xy / NY) * n + oY + xy % NY] = v;
    }
}

}  // namespace

#define CAT2(a, b) a##b
#define CAT3(a, b, c) a##b##_##c
#define CAT3X(a, b, c) CAT3(a, b, c)

void CAT3X(launch_kown_, AIMDHF_LX, AIMDHF_LY)(const KOArgs& a, int nout, const int* has_l, hipStream_t st) {
    if (nout <= 0) return;
    static_for<0, 3>([&](auto jj) {
        constexpr int LJ = decltype(jj)::value;
        if (!has_l[LJ]) return;
        static_for<0, 3>([&](auto ll) {
            constexpr int LL = decltype(ll)::value;
            if (!has_l[LL]) return;
            if (a.nset == 1) k_kown<LX, LY, LJ, LL, 1><<<dim3(nout), dim3(64), 0, st>>>(a);
            else k_kown<LX, LY, LJ, LL, 2><<<dim3(nout), dim3(64), 0, st>>>(a);
        });
    });
}

}  // namespace aimdhf
