// Exchange from the integral store for output blocks (X, Y) of angular
// momenta (AIMDHF_LX, AIMDHF_LY) (compiled once per type): one wave per
// output block, lanes over its gather list (built once per geometry, in the
// fixed order of the per-shell partner lists), each entry one stored quartet
// (XJ|YL) or one stored-ket column of it. Per-lane partial blocks are summed
// by a fixed butterfly and written to K[X][Y] and K[Y][X]; no atomics, so
// the result is bitwise reproducible run to run.

#include <type_traits>

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
constexpr int pcls(int a, int b) { return a > b ? a * (a + 1) / 2 + b : b * (b + 1) / 2 + a; }
constexpr int qcls(int c1, int c2) { return c1 > c2 ? c1 * (c1 + 1) / 2 + c2 : c2 * (c2 + 1) / 2 + c1; }

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
    return v;
}

template <bool DYN, bool VAL, class F>
__device__ __forceinline__ void with_flag(bool v, F&& f) {
    if constexpr (DYN) {
        if (v) f(std::true_type{});
        else f(std::false_type{});
    } else {
        f(std::integral_constant<bool, VAL>{});
    }
}

template <int LJ, int LL, bool FB, bool FK, bool SW, int NSET>
__device__ __forceinline__ void k_item(const KSArgs& a, const double* __restrict__ S, int cs, int oJ, int oL,
                                       double* Kacc) {
    constexpr int NJ = NC(LJ), NL = NC(LL);
    // stored bra / ket records
    constexpr int BA = SW ? (LY > LL ? LY : LL) : (LX > LJ ? LX : LJ);
    constexpr int BB = SW ? (LY > LL ? LL : LY) : (LX > LJ ? LJ : LX);
    constexpr int KA = SW ? (LX > LJ ? LX : LJ) : (LY > LL ? LY : LL);
    constexpr int KB = SW ? (LX > LJ ? LJ : LX) : (LY > LL ? LL : LY);
    constexpr int NAB = NC(BA) * NC(BB), NCD = NC(KA) * NC(KB);
    const size_t n = a.nao, nn = n * n;
    if constexpr (store_split<BA, BB, KA, KB>()) {
        double v[NAB];
#pragma unroll
        for (int i = 0; i < NAB; ++i) v[i] = S[i * NCD + cs];
        k_store_col<LX, LJ, LY, LL, FB, FK, SW, NSET>(v, cs, a.dens, n, nn, oJ, oL, Kacc);
    } else {
        double blk[NAB * NCD];
#pragma unroll
        for (int i = 0; i < NAB * NCD; ++i) blk[i] = S[i];
        double djl[NSET * NJ * NL];
#pragma unroll
        for (int s = 0; s < NSET; ++s)
#pragma unroll
            for (int j = 0; j < NJ; ++j)
#pragma unroll
                for (int l = 0; l < NL; ++l) djl[(s * NJ + j) * NL + l] = a.dens[s * nn + (size_t)(oJ + j) * n + oL + l];
        k_store_full<LX, LJ, LY, LL, FB, FK, SW, NSET>(blk, djl, Kacc);
    }
}

template <int LX_, int LY_, int NSET>
__global__ __launch_bounds__(64) void k_kstore(KSArgs a) {
    const int o = a.out_off[OT] + blockIdx.x;
    const int X = a.out_x[o], Y = a.out_y[o];
    const int lane = threadIdx.x;
    double Kacc[NSET * NX * NY];
#pragma unroll
    for (int i = 0; i < NSET * NX * NY; ++i) Kacc[i] = 0.0;
    static_for<0, 3>([&](auto jj) {
        constexpr int LJ = decltype(jj)::value;
        static_for<0, 3>([&](auto ll) {
            constexpr int LL = decltype(ll)::value;
            constexpr int sub = LJ * 3 + LL;
            const long long e0 = a.kseg[(size_t)o * 9 + sub], e1 = a.kseg[(size_t)o * 9 + sub + 1];
            if (e0 == e1) return;
            constexpr int c1 = pcls(LX, LJ), c2 = pcls(LY, LL), k = qcls(c1, c2);
            // stored ket size (the same for both orientations when c1 == c2)
            constexpr int BA = c1 >= c2 ? (LX > LJ ? LX : LJ) : (LY > LL ? LY : LL);
            constexpr int BB = c1 >= c2 ? (LX > LJ ? LJ : LX) : (LY > LL ? LL : LY);
            constexpr int KA = c1 >= c2 ? (LY > LL ? LY : LL) : (LX > LJ ? LX : LJ);
            constexpr int KB = c1 >= c2 ? (LY > LL ? LL : LY) : (LX > LJ ? LJ : LX);
            constexpr int BLK = NC(BA) * NC(BB) * NC(KA) * NC(KB);
            constexpr int NS = store_split<BA, BB, KA, KB>() ? NC(KA) * NC(KB) : 1;
            const double* base = a.store + a.store_base[k];
            const long long nt = (e1 - e0) * NS;
            for (long long t = lane; t < nt; t += 64) {
                const unsigned long long ent = (unsigned long long)a.klist[e0 + t / NS];
                const int cs = (int)(t % NS);
                const long long qi = (long long)(ent & 0xffffffffull);
                const int oJ = (int)((ent >> 32) & 0x3fff), oL = (int)((ent >> 46) & 0x3fff);
                const bool sw = (ent >> 60) & 1, fb = (ent >> 61) & 1, fk = (ent >> 62) & 1;
                const double* S = base + qi * BLK;
                with_flag<(c1 == c2), (c1 < c2)>(sw, [&](auto swc) {
                    with_flag<(LX == LJ), (LX < LJ)>(fb, [&](auto fbc) {
                        with_flag<(LY == LL), (LY < LL)>(fk, [&](auto fkc) {
                            k_item<LJ, LL, decltype(fbc)::value, decltype(fkc)::value, decltype(swc)::value, NSET>(
                                a, S, cs, oJ, oL, Kacc);
                        });
                    });
                });
            }
        });
    });
    const size_t n = a.nao, nn = n * n;
    const int oX = a.shell_ao[X], oY = a.shell_ao[Y];
#pragma unroll
    for (int i = 0; i < NSET * NX * NY; ++i) {
        const double v = wave_sum(Kacc[i]);
        const int s = i / (NX * NY), x = (i % (NX * NY)) / NY, y = i % NY;
        if (lane == 0) {
            a.K[s * nn + (size_t)(oX + x) * n + oY + y] = v;
            if (X != Y) a.K[s * nn + (size_t)(oY + y) * n + oX + x] = v;
        }
    }
}

}  // namespace

#define CAT3(a, b, c) a##b##_##c
#define CAT3X(a, b, c) CAT3(a, b, c)

void CAT3X(launch_kstore_, AIMDHF_LX, AIMDHF_LY)(const KSArgs& a, int nout, hipStream_t st) {
    if (nout <= 0) return;
    if (a.nset == 1) k_kstore<LX, LY, 1><<<dim3(nout), dim3(64), 0, st>>>(a);
    else k_kstore<LX, LY, 2><<<dim3(nout), dim3(64), 0, st>>>(a);
}

}  // namespace aimdhf
