// Exchange from the integral store for output blocks (X, Y) of angular
// momenta (AIMDHF_LX, AIMDHF_LY) (compiled once per type).
//
// The gather list of every output block is cut into work units of at most a
// fixed number of lane items (entries, or (entry, yl) columns for blocks
// above 36 doubles), one wavefront each, so heavy and light blocks balance.
// Per geometry, k_kcopy lays each unit's stored quartets out in gather
// order, reoriented to (XJ|YL) with the (y, l) index major. A lane item
// (an entry, or an (entry, yl) column when split) holds W consecutive values
// of that order; items are stored in lane-interleaved tiles of 64, value
// major: item T, value v at seg + (T / 64) 64 W + v 64 + T % 64, so every
// load of the digestion and every store of the copy covers 64 consecutive
// words. Per SCF
// iteration, k_kstore streams them and writes one partial block per unit
// (lanes summed by a fixed butterfly; nset NX NY doubles per unit, each
// type in its own region); k_kreduce adds each block's unit
// partials in unit order and writes K[X][Y] and K[Y][X]. For a density
// difference (incremental Fock) units whose largest Schwarz product times
// max |dD| is below a threshold are skipped, and in mixed precision, once
// max |dD| is small, every unit is digested from FP32 copies. No atomics, so the
// result is bitwise reproducible run to run.

#include "aimdhf_internal.h"
#include "md_core.h"

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

template <int LJ, int LL>
struct Sub {
    static constexpr int NJ = NC(LJ), NL = NC(LL), NXJ = NX * NJ, NYL = NY * NL, BLK = NXJ * NYL;
    static constexpr bool SPLIT = BLK > 36;
    static constexpr int NS = SPLIT ? NYL : 1;     // lane items per entry
    static constexpr int W = BLK / NS;               // values per lane item
    static constexpr int K = qcls(pcls(LX, LJ), pcls(LY, LL));
};

// run f(Sub<LJ, LL>) for the runtime subclass index sub = LJ * 3 + LL
template <class F>
__device__ __forceinline__ void with_sub(int sub, F&& f) {
    static_for<0, 9>([&](auto ss) {
        constexpr int S = decltype(ss)::value;
        if (sub == S) f(Sub<S / 3, S % 3>{});
    });
}

template <int LX_, int LY_>
__global__ __launch_bounds__(64) void k_kcopy(KSArgs a) {
    const int u = a.unit_base[OT] + blockIdx.x;
    const int code = a.unit_os[u];
    const int o = code >> 5, f32 = (code >> 4) & 1, sub = code & 15;
    const long long u0 = a.unit_e[2 * u], u1 = a.unit_e[2 * u + 1];
    const int lane = threadIdx.x;
    with_sub(sub, [&](auto sb) {
        using SB = decltype(sb);
        constexpr int NJ = SB::NJ, NL = SB::NL, NXJ = SB::NXJ, NYL = SB::NYL, BLK = SB::BLK;
        constexpr int NS = SB::NS, W = SB::W;
        const size_t seg = (size_t)o * 9 + sub;
        const double* base = a.store + a.store_base[SB::K];
        const long long t0 = (u0 - a.kseg[seg] - (f32 ? a.kbig[seg] : 0)) * NS;
        const long long nt = (u1 - u0) * NS;
        // lanes over consecutive items, so each value row is written by one
        // coalesced (non-temporal: the copies are read only by later builds) store;
        // the W source loads of an item are independent
        for (long long i = lane; i < nt; i += 64) {
            const unsigned long long ent = (unsigned long long)a.klist[u0 + i / NS];
            const int s = (int)(i % NS);
            const double* blk = base + (long long)(ent & 0xffffffffull) * BLK;
            const bool sw = (ent >> 60) & 1, fb = (ent >> 61) & 1, fk = (ent >> 62) & 1;
            double val[W];
#pragma unroll
            for (int v = 0; v < W; ++v) {
                // element (xj, yl) of the (XJ|YL) order
                const int yl = SB::SPLIT ? s : v / NXJ, xj = SB::SPLIT ? v : v % NXJ;
                const int x = xj / NJ, j = xj % NJ, y = yl / NL, l = yl % NL;
                const int i1 = fb ? j * NX + x : x * NJ + j;
                const int i2 = fk ? l * NY + y : y * NL + l;
                val[v] = blk[sw ? i2 * NXJ + i1 : i1 * NYL + i2];
            }
            const long long T = t0 + i;
            const long long dst = (T >> 6) * (64 * W) + (T & 63);
            if (f32) {
                float* out = a.kdata32 + a.kdoff32[seg] + dst;
#pragma unroll
                for (int v = 0; v < W; ++v) __builtin_nontemporal_store((float)val[v], out + v * 64);
            } else {
                double* out = a.kdata + a.kdoff[seg] + dst;
#pragma unroll
                for (int v = 0; v < W; ++v) __builtin_nontemporal_store(val[v], out + v * 64);
                if (a.kdata32b) {
                    float* out32 = a.kdata32b + a.kdoff[seg] + dst;
#pragma unroll
                    for (int v = 0; v < W; ++v) __builtin_nontemporal_store((float)val[v], out32 + v * 64);
                }
            }
        }
    });
}

// items per lane and pass of the digestion: small items are batched so each
// lane has several independent loads in flight
template <int W>
constexpr int batch_items() { return W <= 2 ? 4 : W <= 4 ? 2 : 1; }

// one work unit of subclass SB in storage / arithmetic precision Real: the
// part's tiles start at tile0, the unit's first lane item is t0 within the
// part; its per-lane sums are added to acc
template <typename Real, class SB, int NSET>
__device__ __forceinline__ void k_unit(const KSArgs& a, const Real* __restrict__ tile0, long long t0, long long u0,
                                       long long u1, int lane, Real* __restrict__ acc) {
    constexpr int NJ = SB::NJ, NL = SB::NL, NXJ = SB::NXJ, NYL = SB::NYL, NS = SB::NS, W = SB::W;
    constexpr int U = batch_items<W>();
    const size_t n = a.nao, nn = n * n;
    const long long nt = (u1 - u0) * NS;
    // the list entries run one pass ahead: a pass's density gathers (addressed by
    // its entries) issue with its integral loads instead of after an extra round trip.
    // Entries and integral copies are read once per build and loaded non-temporally,
    // so they do not push the density (read by every unit) out of the caches
    unsigned long long ent_next[U];
#pragma unroll
    for (int k = 0; k < U; ++k) {
        const long long tk = lane + 64 * k;
        ent_next[k] = lane < nt ? (unsigned long long)__builtin_nontemporal_load(a.klist + u0 + (tk < nt ? tk : lane) / NS) : 0ull;
    }
    for (long long t = lane; t < nt; t += 64 * U) {
        // loads of all U items first; an item past the end repeats item t
        // with its density weight zeroed
        unsigned long long ent[U];
        Real v[U][W], wk[U];
        int sk[U];
#pragma unroll
        for (int k = 0; k < U; ++k) {
            const long long tk = t + 64 * k;
            const long long tc = tk < nt ? tk : t;
            ent[k] = ent_next[k];
            wk[k] = tk < nt ? Real(1) : Real(0);
            sk[k] = (int)(tc % NS);
            const long long T = t0 + tc;
            const Real* p = tile0 + (T >> 6) * (64 * W) + (T & 63);
#pragma unroll
            for (int i = 0; i < W; ++i) v[k][i] = __builtin_nontemporal_load(p + i * 64);
        }
        const long long tn = t + 64 * U;
#pragma unroll
        for (int k = 0; k < U; ++k) {
            const long long tk = tn + 64 * k;
            ent_next[k] = tn < nt ? (unsigned long long)__builtin_nontemporal_load(a.klist + u0 + (tk < nt ? tk : tn) / NS) : 0ull;
        }
#pragma unroll
        for (int k = 0; k < U; ++k) {
            const int oJ = (int)((ent[k] >> 32) & 0x3fff), oL = (int)((ent[k] >> 46) & 0x3fff);
            if constexpr (SB::SPLIT) {
                // item = (entry, yl) column of NXJ values
                const int y = sk[k] / NL, l = sk[k] % NL;
#pragma unroll
                for (int s = 0; s < NSET; ++s) {
                    Real dj[NJ];
#pragma unroll
                    for (int j = 0; j < NJ; ++j) dj[j] = wk[k] * Real(a.dens[s * nn + (size_t)(oJ + j) * n + oL + l]);
#pragma unroll
                    for (int x = 0; x < NX; ++x) {
                        Real t2 = 0;
#pragma unroll
                        for (int j = 0; j < NJ; ++j) t2 += v[k][x * NJ + j] * dj[j];
#pragma unroll
                        for (int yy = 0; yy < NY; ++yy)
                            if (yy == y) acc[(s * NX + x) * NY + yy] += t2;
                    }
                }
            } else {
#pragma unroll
                for (int s = 0; s < NSET; ++s) {
                    Real djl[NJ * NL];
#pragma unroll
                    for (int j = 0; j < NJ; ++j)
#pragma unroll
                        for (int l = 0; l < NL; ++l)
                            djl[j * NL + l] = wk[k] * Real(a.dens[s * nn + (size_t)(oJ + j) * n + oL + l]);
#pragma unroll
                    for (int x = 0; x < NX; ++x)
#pragma unroll
                        for (int y = 0; y < NY; ++y) {
                            Real t2 = 0;
#pragma unroll
                            for (int j = 0; j < NJ; ++j)
#pragma unroll
                                for (int l = 0; l < NL; ++l)
                                    t2 += v[k][(y * NL + l) * NXJ + x * NJ + j] * djl[j * NL + l];
                            acc[(s * NX + x) * NY + y] += t2;
                        }
                }
            }
        }
    }
}

template <int LX_, int LY_, int NSET>
__global__ __launch_bounds__(64) void k_kstore(KSArgs a) {
    const int u = a.unit_base[OT] + blockIdx.x;
    const int code = a.unit_os[u];
    const int o = code >> 5, f32 = (code >> 4) & 1, sub = code & 15;
    const long long u0 = a.unit_e[2 * u], u1 = a.unit_e[2 * u + 1];
    const int lane = threadIdx.x;
    double* part = a.kpart + a.part_base[OT] + (size_t)(u - a.unit_base[OT]) * (NSET * NX * NY);
    if (a.delta && a.unit_q[u] * a.dmax[0] < a.dthr) {
        // screened out for this density difference
        for (int i = lane; i < NSET * NX * NY; i += 64) part[i] = 0.0;
        return;
    }
    double Kacc[NSET * NX * NY];
#pragma unroll
    for (int i = 0; i < NSET * NX * NY; ++i) Kacc[i] = 0.0;
    with_sub(sub, [&](auto sb) {
        using SB = decltype(sb);
        const size_t seg = (size_t)o * 9 + sub;
        if (f32 || (a.delta && a.kdata32b && a.dmax[0] <= a.d32)) {
            // FP32 sums per unit, folded into the FP64 block
            float acc[NSET * NX * NY];
#pragma unroll
            for (int i = 0; i < NSET * NX * NY; ++i) acc[i] = 0.0f;
            if (f32)
                k_unit<float, SB, NSET>(a, a.kdata32 + a.kdoff32[seg], (u0 - a.kseg[seg] - a.kbig[seg]) * SB::NS, u0,
                                        u1, lane, acc);
            else
                k_unit<float, SB, NSET>(a, a.kdata32b + a.kdoff[seg], (u0 - a.kseg[seg]) * SB::NS, u0, u1, lane, acc);
#pragma unroll
            for (int i = 0; i < NSET * NX * NY; ++i) Kacc[i] = double(acc[i]);
        } else
            k_unit<double, SB, NSET>(a, a.kdata + a.kdoff[seg], (u0 - a.kseg[seg]) * SB::NS, u0, u1, lane, Kacc);
    });
#pragma unroll
    for (int i = 0; i < NSET * NX * NY; ++i) {
        const double v = wave_sum(Kacc[i]);
        if (lane == 0) part[i] = v;
    }
}

// one thread per (output block of this type, set, element): unit partials in unit order
template <int LX_, int LY_, int NSET>
__global__ void k_kreduce(KSArgs a, int nout) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nout * NSET * NX * NY) return;
    const int o = a.out_off[OT] + t / (NSET * NX * NY), i = t % (NSET * NX * NY);
    double v = 0.0;
    for (int u = a.out_unit[o]; u < a.out_unit[o + 1]; ++u)
        v += a.kpart[a.part_base[OT] + (size_t)(u - a.unit_base[OT]) * (NSET * NX * NY) + i];
    const int X = a.out_x[o], Y = a.out_y[o];
    const size_t n = a.nao, nn = n * n;
    const int s = i / (NX * NY), x = (i % (NX * NY)) / NY, y = i % NY;
    const int oX = a.shell_ao[X], oY = a.shell_ao[Y];
    a.K[s * nn + (size_t)(oX + x) * n + oY + y] = v;
    if (X != Y) a.K[s * nn + (size_t)(oY + y) * n + oX + x] = v;
}

}  // namespace

#define CAT3(a, b, c) a##b##_##c
#define CAT3X(a, b, c) CAT3(a, b, c)

// units of this type: a.unit_base[OT] .. a.unit_base[OT + 1]; nout output blocks
void CAT3X(launch_kstore_, AIMDHF_LX, AIMDHF_LY)(const KSArgs& a, int nout, hipStream_t st) {
    if (nout <= 0) return;
    const int nu = a.unit_base[OT + 1] - a.unit_base[OT];
    if (nu <= 0) return;
    if (a.nset == 1)
        k_kstore<LX, LY, 1><<<dim3(nu), dim3(64), 0, st>>>(a);
    else
        k_kstore<LX, LY, 2><<<dim3(nu), dim3(64), 0, st>>>(a);
}

// the unit partials of launch_kstore into K (every output block is written)
void CAT3X(launch_kreduce_, AIMDHF_LX, AIMDHF_LY)(const KSArgs& a, int nout, hipStream_t st) {
    if (nout <= 0) return;
    const int nr = nout * a.nset * NX * NY;
    if (a.nset == 1)
        k_kreduce<LX, LY, 1><<<dim3((nr + 127) / 128), dim3(128), 0, st>>>(a, nout);
    else
        k_kreduce<LX, LY, 2><<<dim3((nr + 127) / 128), dim3(128), 0, st>>>(a, nout);
}

void CAT3X(launch_kcopy_, AIMDHF_LX, AIMDHF_LY)(const KSArgs& a, int nout, hipStream_t st) {
    const int nu = a.unit_base[OT + 1] - a.unit_base[OT];
    if (nout > 0 && nu > 0) k_kcopy<LX, LY><<<dim3(nu), dim3(64), 0, st>>>(a);
}

}  // namespace aimdhf
