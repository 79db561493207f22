// Per-geometry evaluation of the screened unique quartets of one class
// (compiled once per class with -DAIMDHF_QCLASS=k) into the HBM integral
// store. Quartet t of the class list (bra = heavier pair, ties: higher pair
// index) keeps its contracted block at store[t * NAB * NCD + ab * NCD + cd],
// ab / cd over the bra / ket pair records. The list holds the class's FP64
// quartets first; in mixed precision the rest (q_P q_Q below the switch, held
// and digested in FP32 afterwards) are also evaluated in FP32.

#include "aimdhf_internal.h"
#include "md_eri.h"

#ifndef AIMDHF_QCLASS
#error "compile with -DAIMDHF_QCLASS=<0..20>"
#endif

namespace aimdhf {
using namespace md;

namespace {
constexpr int QC = AIMDHF_QCLASS;
constexpr int pc_hi() { int h = 0; while ((h + 1) * (h + 2) / 2 <= QC) ++h; return h; }
constexpr int pc_lo() { return QC - pc_hi() * (pc_hi() + 1) / 2; }
constexpr int la_of(int pc) { return pc == 0 ? 0 : pc <= 2 ? 1 : 2; }
constexpr int lb_of(int pc) { return pc - la_of(pc) * (la_of(pc) + 1) / 2; }
constexpr int LA = la_of(pc_hi()), LB = lb_of(pc_hi()), LC = la_of(pc_lo()), LD = lb_of(pc_lo());
using Qt = Quartet<LA, LB, LC, LD>;
constexpr bool SPLIT = split_by_cd<LA, LB, LC, LD>();
constexpr int BLK = Qt::NAB * Qt::NCD;
constexpr int NSPLIT = SPLIT ? Qt::NCD : 1;

template <int LA_, int LB_, int LC_, int LD_, typename Real>
__global__ __launch_bounds__(256) void k_eri_store(StoreArgs a, const int2* __restrict__ list, long long nq) {
    __shared__ double btab[BOYS_NGRID * BOYS_SROW];
    boys_stage(a.boys, LA_ + LB_ + LC_ + LD_, btab);
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nq * NSPLIT) return;
    const long long q = t / NSPLIT;
    const int2 pq = list[q];
    const double* rb = a.recs + a.rec_off[pq.x];
    const double* rk = a.recs + a.rec_off[pq.y];
    double* out = a.store + q * BLK;
    if constexpr (SPLIT) {
        const int cd = (int)(t % NSPLIT);
        Real acc[Qt::NAB];
        eri_quartet_cd<LA_, LB_, LC_, LD_, true, Real>(rb, a.npp[pq.x], rk, a.npp[pq.y], btab, cd, acc);
#pragma unroll
        for (int ab = 0; ab < Qt::NAB; ++ab) out[ab * Qt::NCD + cd] = double(acc[ab]);
    } else {
        Real acc[BLK];
        eri_quartet<LA_, LB_, LC_, LD_, true, Real>(rb, a.npp[pq.x], rk, a.npp[pq.y], btab, acc);
#pragma unroll
        for (int i = 0; i < BLK; ++i) out[i] = double(acc[i]);
    }
}
}  // namespace

#define CAT2(a, b) a##b
#define CAT(a, b) CAT2(a, b)

// quartets [0, n64) of the class list in FP64, [n64, nq) in FP32
void CAT(launch_eri_store_, AIMDHF_QCLASS)(const StoreArgs& a, const int2* list, long long nq, long long n64,
                                           hipStream_t st) {
    const long long nt = n64 * NSPLIT;
    if (nt > 0)
        k_eri_store<LA, LB, LC, LD, double><<<dim3((unsigned)((nt + 255) / 256)), dim3(256), 0, st>>>(a, list, n64);
    const long long nt32 = (nq - n64) * NSPLIT;
    if (nt32 > 0) {
        StoreArgs a32 = a;
        a32.store = a.store + n64 * BLK;
        k_eri_store<LA, LB, LC, LD, float><<<dim3((unsigned)((nt32 + 255) / 256)), dim3(256), 0, st>>>(
            a32, list + n64, nq - n64);
    }
}

}  // namespace aimdhf
