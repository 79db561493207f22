// Coulomb from the integral store for bra pairs of class AIMDHF_PCLASS
// (compiled once per class).
//
// The gather list of every pair P (each partner Q with q_P q_Q >= threshold,
// by partner class) is cut into work units of at most a fixed number of lane
// items, one wavefront each. Per geometry, k_jcopy lays each unit's stored
// quartets out in gather order, oriented (P|Q) with the partner's component
// index major. A lane item (an entry, or an (entry, cd) column when split)
// holds W consecutive values of that order; items are stored in
// lane-interleaved tiles of 64, value major: item T, value v at
// seg + (T / 64) 64 W + v 64 + T % 64, so every load of the digestion and
// every store of the copy covers 64 consecutive words. Per SCF
// iteration, k_jstore forms sum_cd (ab|cd) w_Q D_cd (w_Q = 2 for C != D) per
// unit (lanes summed by a fixed butterfly) and k_jreduce adds each pair's
// unit partials in unit order into J_P, so J is bitwise reproducible. For a
// density difference (incremental Fock) units whose largest Schwarz product
// times max |dD| is below a threshold are skipped, and in mixed precision,
// once max |dD| is small, every unit is digested from FP32 copies.

#include "aimdhf_internal.h"
#include "md_core.h"

#ifndef AIMDHF_PCLASS
#error "compile with -DAIMDHF_PCLASS=<0..5>"
#endif

namespace aimdhf {
using namespace md;

namespace {
constexpr int CP = AIMDHF_PCLASS;
constexpr int la_of(int pc) { return pc == 0 ? 0 : pc <= 2 ? 1 : 2; }
constexpr int lb_of(int pc) { return pc - la_of(pc) * (la_of(pc) + 1) / 2; }
constexpr int LA = la_of(CP), LB = lb_of(CP);
constexpr int NAB = NC(LA) * NC(LB);
constexpr int qcls(int c1, int c2) { return c1 > c2 ? c1 * (c1 + 1) / 2 + c2 : c2 * (c2 + 1) / 2 + c1; }

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
    return v;
}

template <int CQ>
struct Part {
    static constexpr int NDc = NC(lb_of(CQ)), NCD = NC(la_of(CQ)) * NDc, BLK = NAB * NCD;
    static constexpr bool SPLIT = BLK > 36;
    static constexpr int NS = SPLIT ? NCD : 1;       // lane items per entry
    static constexpr int W = BLK / NS;               // values per lane item
    static constexpr int K = qcls(CP, CQ);
};

template <class F>
__device__ __forceinline__ void with_part(int cq, F&& f) {
    static_for<0, 6>([&](auto qq) {
        constexpr int C = decltype(qq)::value;
        if (cq == C) f(Part<C>{});
    });
}

template <int CP_>
__global__ __launch_bounds__(64) void k_jcopy(JSArgs a) {
    const int u = a.unit_base[CP] + blockIdx.x;
    const int code = a.unit_pc[u];
    const int P = code >> 4, f32 = (code >> 3) & 1, cq = code & 7;
    const long long u0 = a.unit_e[2 * u], u1 = a.unit_e[2 * u + 1];
    const int lane = threadIdx.x;
    with_part(cq, [&](auto pt) {
        using PT = decltype(pt);
        constexpr int NCD = PT::NCD, BLK = PT::BLK, NS = PT::NS, W = PT::W;
        const size_t seg = (size_t)P * 6 + cq;
        const double* base = a.store + a.store_base[PT::K];
        const long long t0 = (u0 - a.jseg[seg] - (f32 ? a.jbig[seg] : 0)) * NS;
        const long long nt = (u1 - u0) * NS;
        // lanes over consecutive items, so each value row is written by one
        // coalesced (non-temporal: the copies are read only by later builds) store;
        // the W source loads of an item are independent
        for (long long i = lane; i < nt; i += 64) {
            const unsigned long long ent = (unsigned long long)a.jlist[u0 + i / NS];
            const int s = (int)(i % NS);
            const double* blk = base + (long long)(ent & 0xffffffffull) * BLK;
            const bool sw = (ent >> 60) & 1;
            double val[W];
#pragma unroll
            for (int v = 0; v < W; ++v) {
                // element (cd, ab) of the (P|Q) order
                const int cd = PT::SPLIT ? s : v / NAB, ab = PT::SPLIT ? v : v % NAB;
                val[v] = blk[sw ? cd * NAB + ab : ab * NCD + cd];
            }
            const long long T = t0 + i;
            const long long dst = (T >> 6) * (64 * W) + (T & 63);
            if (f32) {
                float* out = a.jdata32 + a.jdoff32[seg] + dst;
#pragma unroll
                for (int v = 0; v < W; ++v) __builtin_nontemporal_store((float)val[v], out + v * 64);
            } else {
                double* out = a.jdata + a.jdoff[seg] + dst;
#pragma unroll
                for (int v = 0; v < W; ++v) __builtin_nontemporal_store(val[v], out + v * 64);
                if (a.jdata32b) {
                    float* out32 = a.jdata32b + a.jdoff[seg] + dst;
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

// one work unit of partner class PT in storage / arithmetic precision Real:
// the part's tiles start at tile0, the unit's first lane item is t0 within
// the part; its per-lane sums are added to acc
template <typename Real, class PT, int NSET>
__device__ __forceinline__ void j_unit(const JSArgs& a, const Real* __restrict__ tile0, long long t0, long long u0,
                                       long long u1, int lane, Real* __restrict__ acc) {
    constexpr int NDc = PT::NDc, NCD = PT::NCD, NS = PT::NS, W = PT::W;
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
        ent_next[k] = lane < nt ? (unsigned long long)__builtin_nontemporal_load(a.jlist + u0 + (tk < nt ? tk : lane) / NS) : 0ull;
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
            wk[k] = tk >= nt ? Real(0) : ((ent[k] >> 61) & 1) ? Real(2) : Real(1);
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
            ent_next[k] = tn < nt ? (unsigned long long)__builtin_nontemporal_load(a.jlist + u0 + (tk < nt ? tk : tn) / NS) : 0ull;
        }
#pragma unroll
        for (int k = 0; k < U; ++k) {
            const int oC = (int)((ent[k] >> 32) & 0x3fff), oD = (int)((ent[k] >> 46) & 0x3fff);
            if constexpr (PT::SPLIT) {
                // item = (entry, cd) column of NAB values
                const int cd = sk[k];
#pragma unroll
                for (int s = 0; s < NSET; ++s) {
                    const Real dv = wk[k] * Real(a.dens[s * nn + (size_t)(oC + cd / NDc) * n + oD + cd % NDc]);
#pragma unroll
                    for (int ab = 0; ab < NAB; ++ab) acc[s * NAB + ab] += v[k][ab] * dv;
                }
            } else {
#pragma unroll
                for (int s = 0; s < NSET; ++s) {
                    Real dq[NCD];
#pragma unroll
                    for (int cd = 0; cd < NCD; ++cd)
                        dq[cd] = wk[k] * Real(a.dens[s * nn + (size_t)(oC + cd / NDc) * n + oD + cd % NDc]);
#pragma unroll
                    for (int ab = 0; ab < NAB; ++ab) {
                        Real t2 = 0;
#pragma unroll
                        for (int cd = 0; cd < NCD; ++cd) t2 += v[k][cd * NAB + ab] * dq[cd];
                        acc[s * NAB + ab] += t2;
                    }
                }
            }
        }
    }
}

template <int CP_, int NSET>
__global__ __launch_bounds__(64) void k_jstore(JSArgs a) {
    const int u = a.unit_base[CP] + blockIdx.x;
    const int code = a.unit_pc[u];
    const int P = code >> 4, f32 = (code >> 3) & 1, cq = code & 7;
    const long long u0 = a.unit_e[2 * u], u1 = a.unit_e[2 * u + 1];
    const int lane = threadIdx.x;
    double* part = a.jpart + a.part_base[CP] + (size_t)(u - a.unit_base[CP]) * (NSET * NAB);
    if (a.delta && a.unit_q[u] * a.dmax[0] < a.dthr) {
        // screened out for this density difference
        for (int i = lane; i < NSET * NAB; i += 64) part[i] = 0.0;
        return;
    }
    double Jacc[NSET * NAB];
#pragma unroll
    for (int i = 0; i < NSET * NAB; ++i) Jacc[i] = 0.0;
    with_part(cq, [&](auto pt) {
        using PT = decltype(pt);
        const size_t seg = (size_t)P * 6 + cq;
        if (f32 || (a.delta && a.jdata32b && a.dmax[0] <= a.d32)) {
            // FP32 sums per unit, folded into the FP64 block
            float acc[NSET * NAB];
#pragma unroll
            for (int i = 0; i < NSET * NAB; ++i) acc[i] = 0.0f;
            if (f32)
                j_unit<float, PT, NSET>(a, a.jdata32 + a.jdoff32[seg], (u0 - a.jseg[seg] - a.jbig[seg]) * PT::NS, u0,
                                        u1, lane, acc);
            else
                j_unit<float, PT, NSET>(a, a.jdata32b + a.jdoff[seg], (u0 - a.jseg[seg]) * PT::NS, u0, u1, lane, acc);
#pragma unroll
            for (int i = 0; i < NSET * NAB; ++i) Jacc[i] = double(acc[i]);
        } else
            j_unit<double, PT, NSET>(a, a.jdata + a.jdoff[seg], (u0 - a.jseg[seg]) * PT::NS, u0, u1, lane, Jacc);
    });
#pragma unroll
    for (int i = 0; i < NSET * NAB; ++i) {
        const double v = wave_sum(Jacc[i]);
        if (lane == 0) part[i] = v;
    }
}

// one thread per (pair of this class, set, element): unit partials in unit order
template <int CP_, int NSET>
__global__ void k_jreduce(JSArgs a, int npairs) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= npairs * NSET * NAB) return;
    const int idx = a.cls_off + t / (NSET * NAB), i = t % (NSET * NAB);
    double v = 0.0;
    for (int u = a.pair_unit[idx]; u < a.pair_unit[idx + 1]; ++u)
        v += a.jpart[a.part_base[CP] + (size_t)(u - a.unit_base[CP]) * (NSET * NAB) + i];
    const int P = a.cls_pairs[idx];
    const int A = a.pair_a[P], B = a.pair_b[P];
    const int oA = a.shell_ao[A], oB = a.shell_ao[B];
    const size_t n = a.nao, nn = n * n;
    constexpr int NBc = NC(LB);
    const int s = i / NAB, x = (i % NAB) / NBc, y = i % NBc;
    a.J[s * nn + (size_t)(oA + x) * n + oB + y] = v;
    if (A != B) a.J[s * nn + (size_t)(oB + y) * n + oA + x] = v;
}

}  // namespace

#define CAT2(a, b) a##b
#define CAT(a, b) CAT2(a, b)

// units of this class: a.unit_base[CP] .. a.unit_base[CP + 1]; npairs pairs from a.cls_off
void CAT(launch_jstore_, AIMDHF_PCLASS)(const JSArgs& a, int npairs, hipStream_t st) {
    if (npairs <= 0) return;
    const int nu = a.unit_base[CP + 1] - a.unit_base[CP];
    if (nu <= 0) return;
    if (a.nset == 1)
        k_jstore<CP, 1><<<dim3(nu), dim3(64), 0, st>>>(a);
    else
        k_jstore<CP, 2><<<dim3(nu), dim3(64), 0, st>>>(a);
}

// the unit partials of launch_jstore into the J blocks of this pair class
void CAT(launch_jreduce_, AIMDHF_PCLASS)(const JSArgs& a, int npairs, hipStream_t st) {
    if (npairs <= 0) return;
    const int nr = npairs * a.nset * NAB;
    if (a.nset == 1)
        k_jreduce<CP, 1><<<dim3((nr + 127) / 128), dim3(128), 0, st>>>(a, npairs);
    else
        k_jreduce<CP, 2><<<dim3((nr + 127) / 128), dim3(128), 0, st>>>(a, npairs);
}

void CAT(launch_jcopy_, AIMDHF_PCLASS)(const JSArgs& a, int npairs, hipStream_t st) {
    const int nu = a.unit_base[CP + 1] - a.unit_base[CP];
    if (npairs > 0 && nu > 0) k_jcopy<CP><<<dim3(nu), dim3(64), 0, st>>>(a);
}

}  // namespace aimdhf
