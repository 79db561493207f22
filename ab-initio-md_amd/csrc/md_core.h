// McMurchie-Davidson building blocks shared by the gfx950 kernels.
//
// Everything here is __host__ __device__ so that the host-side check harness
// (host_check.cpp, test-only) evaluates exactly the arithmetic the kernels do.
//
//   Boys function      F_n(T): 402 x 43 grid table, 10-term Taylor expansion
//                      about the nearest grid point for T < 40, asymptotic
//                      F_0 and upward recursion above (the scheme of
//                      aimd/qc/boys.py:49-140); lower orders by the downward
//                      recursion F_n = (2T F_{n+1} + e^{-T}) / (2n+1).
//   Hermite R          R^n_{tuv}(alpha, PQ) levels n = L..0 updated in place,
//                      highest degree first (aimd/qc/hermite.py:144-178).
//   Hermite E          1-D E^{ij}_t recursion (aimd/qc/hermite.py:109-141).
//
// Hermite indices are linearised by degree exactly as the reference does
// (aimd/qc/hermite.py:73-105): within degree d, t from d down to 0, then u
// from d-t down to 0. hidx() computes that position arithmetically so that
// fully unrolled loops index registers with compile-time constants.

#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <type_traits>

#define MD_HD __host__ __device__ __forceinline__

namespace md {

constexpr int NH(int L) { return L < 0 ? 0 : (L + 1) * (L + 2) * (L + 3) / 6; }
constexpr int NC(int l) { return (l + 1) * (l + 2) / 2; }

// linear Hermite index of (t, u, v)
constexpr int hidx(int t, int u, int v) {
    return NH(t + u + v - 1) + (u + v) * (u + v + 1) / 2 + v;
}
// (d-t)(d-t+1)/2 + (d-t-u) with d = t+u+v equals (u+v)(u+v+1)/2 + v

// Cartesian component i of angular momentum l in PySCF order (xx, xy, xz, yy, yz, zz)
constexpr int cart_x(int l, int i) {
    int n = 0;
    for (int lx = l; lx >= 0; --lx)
        for (int ly = l - lx; ly >= 0; --ly) {
            if (n == i) return lx;
            ++n;
        }
    return 0;
}
constexpr int cart_y(int l, int i) {
    int n = 0;
    for (int lx = l; lx >= 0; --lx)
        for (int ly = l - lx; ly >= 0; --ly) {
            if (n == i) return ly;
            ++n;
        }
    return 0;
}
constexpr int cart_z(int l, int i) { return l - cart_x(l, i) - cart_y(l, i); }

constexpr int BOYS_NGRID = 402;
constexpr int BOYS_NCOL = 43;
constexpr double BOYS_XSWITCH = 40.0;
constexpr double MD_PI = 3.141592653589793;
constexpr double TWO_PI_2_5 = 34.986836655249725;  // 2 pi^(5/2)

// Boys table layouts: the full table (rows of BOYS_NCOL orders from 0), or a
// per-class slice of BOYS_SLICE orders starting at the class's top order,
// staged in LDS by the kernels (boys_stage) in rows of BOYS_SROW with
// exp(-T_k) of the grid point last. BS = row stride, C0 = first order.
constexpr int BOYS_SLICE = 10;
constexpr int BOYS_SROW = BOYS_SLICE + 1;

// exp / sqrt of the working precision (float arithmetic stays float)
MD_HD inline double md_exp(double x) { return exp(x); }
MD_HD inline float md_exp(float x) { return expf(x); }
MD_HD inline double md_sqrt(double x) { return sqrt(x); }
MD_HD inline float md_sqrt(float x) { return sqrtf(x); }

// 1/sqrt(x), x > 0: the hardware estimate refined by Newton steps
// y <- y (3 - x y^2) / 2, each taking a relative error e to ~1.5 e^2. The
// FP64 estimate is good to ~5e-8 (measured on gfx950), so FP64 takes two
// steps (~1e-16; one step leaves ~4e-15, which the cancellation between
// core-orbital derivative integrals turns into visible force errors); the
// FP32 estimate is already near 1 ulp and takes one.
MD_HD inline double md_rsqrt(double x) {
#ifdef __HIP_DEVICE_COMPILE__
    double y = __builtin_amdgcn_rsq(x);
#else
    double y = 1.0 / sqrt(x);
#ifdef MD_HOST_RSQ_ESTIMATE
    y = MD_HOST_RSQ_ESTIMATE(x, y);     // host checks: model of the hardware estimate's error
#endif
#endif
    const double hx = 0.5 * x;
    y = fma(y, fma(-hx * y, y, 0.5), y);
    return fma(y, fma(-hx * y, y, 0.5), y);
}
MD_HD inline float md_rsqrt(float x) {
#ifdef __HIP_DEVICE_COMPILE__
    const float y = __builtin_amdgcn_rsqf(x);
#else
    const float y = 1.0f / sqrtf(x);
#endif
    const float h = 0.5f * x * y;
    return fmaf(y, fmaf(-h, y, 0.5f), y);
}

// alpha = p q / (p + q) and the ERI prefactor 2 pi^(5/2) / (p q sqrt(p + q))
// from the pair exponents, their reciprocals (stored in the pair records)
// and one refined 1/sqrt(p + q): no division
template <typename Real>
MD_HD void quartet_factors(Real p, Real q, Real ip, Real iq, Real& alpha, Real& pref) {
    const Real r = md_rsqrt(p + q);
    alpha = p * q * (r * r);
    pref = Real(TWO_PI_2_5) * (ip * iq) * r;
}

// F_L(T) (single top order), as aimd/qc/boys.py:87-114, in precision Real
template <int BS = BOYS_NCOL, int C0 = 0, typename Real = double>
MD_HD Real boys_top(int L, Real T, const double* __restrict__ tab) {
    if (T < Real(BOYS_XSWITCH)) {
        int k = (int)(T * Real(10.0) + Real(0.5));
        Real dx = Real(k) * Real(0.1) - T;
        const double* row = tab + k * BS + (L - C0);
        Real acc = 0, fac = 1;
        acc += Real(row[0]) * fac; fac *= dx * Real(1.0);
        acc += Real(row[1]) * fac; fac *= dx * Real(0.5);
        acc += Real(row[2]) * fac; fac *= dx * Real(1.0 / 3.0);
        acc += Real(row[3]) * fac; fac *= dx * Real(0.25);
        acc += Real(row[4]) * fac; fac *= dx * Real(0.2);
        acc += Real(row[5]) * fac; fac *= dx * Real(1.0 / 6.0);
        acc += Real(row[6]) * fac; fac *= dx * Real(1.0 / 7.0);
        acc += Real(row[7]) * fac; fac *= dx * Real(0.125);
        acc += Real(row[8]) * fac; fac *= dx * Real(1.0 / 9.0);
        acc += Real(row[9]) * fac;
        return acc;
    }
    Real f = Real(0.5) * md_sqrt(Real(MD_PI) / T);
    if (L > 0) {
        Real ex = md_exp(-T);
        Real inv2x = Real(0.5) / T;
        for (int m = 0; m < L; ++m) f = (Real(2 * m + 1) * f - ex) * inv2x;
    }
    return f;
}

// F_0..F_L into F (downward recursion from F_L, as the reference's R code does)
template <int L, int BS = BOYS_NCOL, int C0 = 0, typename Real = double>
MD_HD void boys_all(Real T, const double* __restrict__ tab, Real* F) {
    F[L] = boys_top<BS, C0, Real>(L, T, tab);
    if (L > 0) {
        Real ex = md_exp(-T);
        Real t2 = Real(2.0) * T;
#pragma unroll
        for (int n = L - 1; n >= 0; --n) F[n] = (t2 * F[n + 1] + ex) * (Real(1.0) / Real(2 * n + 1));
    }
}

// The same from a staged slice (rows of BOYS_SROW: orders L .. L+9, then
// exp(-T_k)). Below the switch, F_L is the 10-term Taylor expansion and
// exp(-T) = exp(-T_k) exp(dx), dx = T_k - T in [-0.05, 0.05], with exp(dx)
// its degree-9 Taylor polynomial (remainder < 0.05^10 / 10! ~ 3e-20).
// Above it, F_L by upward recursion from F_0 = sqrt(pi / T) / 2, with one
// exp(-T) shared by both recursions.
template <int L, typename Real = double>
MD_HD void boys_all_slice(Real T, const double* __restrict__ tab, Real* F) {
    Real ex;
    if (T < Real(BOYS_XSWITCH)) {
        const int k = (int)(T * Real(10.0) + Real(0.5));
        const Real dx = Real(k) * Real(0.1) - T;
        const double* row = tab + k * BOYS_SROW;
        Real acc = 0, fac = 1;
        acc += Real(row[0]) * fac; fac *= dx * Real(1.0);
        acc += Real(row[1]) * fac; fac *= dx * Real(0.5);
        acc += Real(row[2]) * fac; fac *= dx * Real(1.0 / 3.0);
        acc += Real(row[3]) * fac; fac *= dx * Real(0.25);
        acc += Real(row[4]) * fac; fac *= dx * Real(0.2);
        acc += Real(row[5]) * fac; fac *= dx * Real(1.0 / 6.0);
        acc += Real(row[6]) * fac; fac *= dx * Real(1.0 / 7.0);
        acc += Real(row[7]) * fac; fac *= dx * Real(0.125);
        acc += Real(row[8]) * fac; fac *= dx * Real(1.0 / 9.0);
        acc += Real(row[9]) * fac;
        F[L] = acc;
        if constexpr (L > 0) {
            Real e = Real(1.0 / 362880.0);
            e = e * dx + Real(1.0 / 40320.0);
            e = e * dx + Real(1.0 / 5040.0);
            e = e * dx + Real(1.0 / 720.0);
            e = e * dx + Real(1.0 / 120.0);
            e = e * dx + Real(1.0 / 24.0);
            e = e * dx + Real(1.0 / 6.0);
            e = e * dx + Real(0.5);
            e = e * dx + Real(1.0);
            e = e * dx + Real(1.0);
            ex = Real(row[BOYS_SLICE]) * e;
        }
    } else {
        // sqrt and division here, not md_rsqrt: with the hardware reciprocal
        // square root in both this branch and quartet_factors, gfx950 builds
        // of the derivative kernels gave force errors ~5e-5 that neither use
        // alone shows (and the host build of the same source does not)
#ifdef MD_HOST_LARGE_T_RSQ
        const Real rt = md_rsqrt(T);                // host checks of that variant's arithmetic
        Real f = Real(0.5 * 1.7724538509055159) * rt;
#else
        Real f = Real(0.5) * md_sqrt(Real(MD_PI) / T);
#endif
        if constexpr (L > 0) {
            ex = md_exp(-T);
#ifdef MD_HOST_LARGE_T_RSQ
            const Real inv2x = Real(0.5) * rt * rt;
#else
            const Real inv2x = Real(0.5) / T;
#endif
            for (int m = 0; m < L; ++m) f = (Real(2 * m + 1) * f - ex) * inv2x;
        }
        F[L] = f;
    }
    if constexpr (L > 0) {
        const Real t2 = Real(2.0) * T;
#pragma unroll
        for (int n = L - 1; n >= 0; --n) F[n] = (t2 * F[n + 1] + ex) * (Real(1.0) / Real(2 * n + 1));
    }
}

// compile-time loop: f(std::integral_constant<int, i>) for i in [B, E)
template <int B, int E, class F>
MD_HD void static_for(F&& f) {
    if constexpr (B < E) {
        f(std::integral_constant<int, B>{});
        static_for<B + 1, E>(f);
    }
}

// (t, u, v) of every linear Hermite index up to degree HDEG_MAX
constexpr int HDEG_MAX = 10;
struct HermTable {
    int t[NH(HDEG_MAX)], u[NH(HDEG_MAX)], v[NH(HDEG_MAX)];
};
constexpr HermTable make_herm_table() {
    HermTable h{};
    int n = 0;
    for (int d = 0; d <= HDEG_MAX; ++d)
        for (int t = d; t >= 0; --t)
            for (int u = d - t; u >= 0; --u) {
                h.t[n] = t; h.u[n] = u; h.v[n] = d - t - u;
                ++n;
            }
    return h;
}
inline constexpr HermTable HT = make_herm_table();

// R[hidx(t,u,v)] = scale * R_{tuv}(alpha, (x,y,z)) for t+u+v <= L
// tab: the full table (SLICE false) or the slice starting at order L (SLICE true)
template <int L, bool SLICE = false, typename Real = double>
MD_HD void hermite_r(Real alpha, Real x, Real y, Real z, Real scale,
                     const double* __restrict__ tab, Real* R) {
    Real T = alpha * (x * x + y * y + z * z);
    Real F[L + 1];
    if constexpr (SLICE) boys_all_slice<L, Real>(T, tab, F);
    else boys_all<L, BOYS_NCOL, 0, Real>(T, tab, F);
    Real m2a = Real(-2.0) * alpha;
    Real pw[L + 1];
    pw[0] = scale;
#pragma unroll
    for (int n = 1; n <= L; ++n) pw[n] = pw[n - 1] * m2a;
    R[0] = pw[L] * F[L];
    static_for<0, L>([&](auto nn) {
        constexpr int n = L - 1 - decltype(nn)::value;
        constexpr int nhn = NH(L - n);
        // highest degree first, so the level-(n+1) values read below are not yet overwritten
        static_for<1, nhn>([&](auto hh) {
            constexpr int h = nhn - decltype(hh)::value;
            constexpr int t = HT.t[h], u = HT.u[h], v = HT.v[h];
            Real r;
            if constexpr (t > 0) {
                r = x * R[hidx(t - 1, u, v)];
                if constexpr (t > 1) r += Real(t - 1) * R[hidx(t - 2, u, v)];
            } else if constexpr (u > 0) {
                r = y * R[hidx(0, u - 1, v)];
                if constexpr (u > 1) r += Real(u - 1) * R[hidx(0, u - 2, v)];
            } else {
                r = z * R[hidx(0, 0, v - 1)];
                if constexpr (v > 1) r += Real(v - 1) * R[hidx(0, 0, v - 2)];
            }
            R[h] = r;
        });
        R[0] = pw[n] * F[n];
    });
}

// 1-D Hermite expansion E^{ij}_t, i <= IM, j <= JM, t <= i+j; e[(i*(JM+1)+j)*(IM+JM+1)+t]
// (entries with t > i+j are zero)
template <int IM, int JM, typename Real = double>
MD_HD void hermite_e(Real a, Real b, Real xab, Real e00, Real* e) {
    constexpr int TS = IM + JM + 1;
    constexpr int JS = JM + 1;
    const Real p = a + b;
    const Real xpa = -b * xab / p;
    const Real xpb = a * xab / p;
    const Real h = Real(0.5) / p;
#pragma unroll
    for (int i = 0; i < (IM + 1) * JS * TS; ++i) e[i] = Real(0);
    e[0] = e00;
    static_for<1, JM + 1>([&](auto jj) {
        constexpr int j = decltype(jj)::value;
        static_for<0, j + 1>([&](auto tt) {
            constexpr int t = decltype(tt)::value;
            Real v = xpb * e[(j - 1) * TS + t];
            if constexpr (t + 1 <= j - 1) v += Real(t + 1) * e[(j - 1) * TS + t + 1];
            if constexpr (t > 0) v += h * e[(j - 1) * TS + t - 1];
            e[j * TS + t] = v;
        });
    });
    static_for<1, IM + 1>([&](auto ii) {
        constexpr int i = decltype(ii)::value;
        static_for<0, JM + 1>([&](auto jj) {
            constexpr int j = decltype(jj)::value;
            static_for<0, i + j + 1>([&](auto tt) {
                constexpr int t = decltype(tt)::value;
                Real v = xpa * e[((i - 1) * JS + j) * TS + t];
                if constexpr (t + 1 <= i - 1 + j) v += Real(t + 1) * e[((i - 1) * JS + j) * TS + t + 1];
                if constexpr (t > 0) v += h * e[((i - 1) * JS + j) * TS + t - 1];
                e[(i * JS + j) * TS + t] = v;
            });
        });
    });
}

// ---------------------------------------------------------------- pair records
// A primitive pair of shells (A, B), l_A >= l_B, is stored as a record of
// REC_HDR header doubles followed by the 3-D Hermite table
//   E[h][ab] = f_a f_b c_a c_b E^{a_x b_x}_t E^{a_y b_y}_u E^{a_z b_z}_v,  h = hidx(t,u,v)
// (NH(l_A+l_B) x NC(l_A) NC(l_B), row-major), i.e. the contraction
// coefficients and Cartesian normalisation factors are folded in.
enum { R_P = 0, R_PX = 1, R_PY = 2, R_PZ = 3, R_A = 4, R_B = 5, R_C = 6, R_IP = 7, REC_HDR = 8 };

constexpr int rec_size(int la, int lb) { return REC_HDR + NH(la + lb) * NC(la) * NC(lb); }

// pair class index of (la >= lb): ss 0, ps 1, pp 2, ds 3, dp 4, dd 5
constexpr int pair_class(int la, int lb) { return la * (la + 1) / 2 + lb; }
constexpr int NPAIRCLASS = 6;
constexpr int NQCLASS = NPAIRCLASS * (NPAIRCLASS + 1) / 2;   // 21

// ---------------------------------------------------------------- fixed point
// Order-independent accumulation: x is split into hi = round(x 2^32) and the
// remainder lo = round((x 2^32 - hi) 2^42); hi and lo are summed in separate
// int64 words. Integer addition is associative, so the sums (and the double
// they are converted to) do not depend on the order in which contributions
// arrive. Resolution 2^-74 per contribution, range |sum| < 2^31.
MD_HD void fx_split(double x, long long& hi, long long& lo) {
    double xs = x * 4294967296.0;          // 2^32, exact
    double h = rint(xs);
    double r = xs - h;                     // exact
    hi = (long long)h;
    lo = (long long)rint(r * 4398046511104.0);   // 2^42
}

MD_HD double fx_value(long long hi, long long lo) {
    return (double)hi * (1.0 / 4294967296.0) + (double)lo * (1.0 / 18889465931478580854784.0);  // 2^-32, 2^-74
}

#ifdef __HIPCC__
// Stage the BOYS_SLICE orders from Ltop of the full table (global) and
// exp(-T_k) of every grid point into an LDS slice (BOYS_NGRID * BOYS_SROW
// doubles) with the whole workgroup; ends with a barrier.
__device__ inline void boys_stage(const double* __restrict__ tab, int Ltop, double* lds) {
    for (int i = threadIdx.x; i < BOYS_NGRID * BOYS_SROW; i += blockDim.x) {
        const int k = i / BOYS_SROW, c = i % BOYS_SROW;
        lds[i] = c < BOYS_SLICE ? tab[k * BOYS_NCOL + Ltop + c] : exp(-0.1 * k);
    }
    __syncthreads();
}
#endif

}  // namespace md
