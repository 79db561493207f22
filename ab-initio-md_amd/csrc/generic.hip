// Runtime-l McMurchie-Davidson kernels for bases with f and g shells (l <= 4).
//
// The class-specialised kernels cover l <= 2 with compile-time shapes; a plan
// whose basis holds a shell of l = 3 or 4 runs every integral through the
// kernels below instead (aimd/qc/hermite.py:51 sets LMAX = 4 for the
// reference's kernels). Shapes are runtime values, so one kernel serves every
// class:
//
//   one workgroup of one wavefront per shell pair (Schwarz bounds, one-electron
//   matrices and gradient) or per unique shell quartet (J/K, derivative ERIs,
//   grid-stride over the list); lanes over the Cartesian elements of the pair
//   or quartet, in chunks of 64; primitive pairs kept by the reference's
//   cutoff mu |AB|^2 <= 50 (aimd/qc/integrals.py:511-560) for the
//   two-electron work, all primitive pairs for the one-electron work
//   (aimd/qc/integrals.py:146-501).
//
//   Per primitive pair, lanes 0-2 build the 1-D Hermite tables E^{ij}_t of
//   x, y, z in LDS (aimd/qc/hermite.py:109-141, with rows past l for the
//   kinetic and derivative shifts); per primitive quartet (or pair and
//   nucleus) the wavefront builds R^0_{tuv} level by level in two LDS buffers
//   (aimd/qc/hermite.py:144-178), and every lane contracts its element.
//
//   Derivatives by the shift rule d/dA_x phi_a = 2 alpha phi_{a+1x} - a_x
//   phi_{a-1x}; the fourth centre of a quartet and the second of an overlap
//   or kinetic pair by translational invariance; the Hellmann-Feynman term
//   of the nuclear attraction from R_{t+1,u,v}.
//
// J, K and the gradients are summed in the exact two-word fixed point of
// md_core.h (fx_split), so every result is independent of the order in which
// workgroups finish: bitwise reproducible run to run.

#include "aimdhf_internal.h"
#include "md_core.h"
#include "md_pairs.h"

namespace aimdhf {
using namespace md;

namespace {
constexpr int GL = GEN_LMAX;                    // highest shell l
constexpr int GI = GL + 2;                      // highest 1-D index of an E table (l_B + 2 for kinetic)
constexpr int GT = 2 * GI + 1;                  // t <= i + j
constexpr int GE = (GI + 1) * (GI + 1) * GT;    // one 1-D table
constexpr int GLR = 4 * GL + 1;                 // highest R degree (derivative ERIs)
constexpr int GNH = NH(GLR);
constexpr int WG = 64;

struct Lds {
    double eb[3 * GE];                          // bra (or one-electron pair) tables x, y, z
    double ek[3 * GE];                          // ket tables
    double r[2 * GNH];                          // R levels, ping-pong
    double f[GLR + 2];                          // scaled Boys values of the current R
};

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
    return v;
}

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v = fmax(v, __shfl_xor(v, m, 64));
    return v;
}

__device__ __forceinline__ void fx_add(long long* acc, size_t i, double x) {
    long long h, l;
    fx_split(x, h, l);
    if (h) atomicAdd((unsigned long long*)&acc[2 * i], (unsigned long long)h);
    if (l) atomicAdd((unsigned long long*)&acc[2 * i + 1], (unsigned long long)l);
}

__device__ __forceinline__ int ncart_of(int l) { return (l + 1) * (l + 2) / 2; }

// Cartesian exponents of component c of angular momentum l (PySCF order)
__device__ __forceinline__ void comp(int l, int c, int& x, int& y, int& z) {
    int n = 0;
    for (int lx = l; lx >= 0; --lx)
        for (int ly = l - lx; ly >= 0; --ly) {
            if (n == c) {
                x = lx;
                y = ly;
                z = l - lx - ly;
                return;
            }
            ++n;
        }
    x = y = z = 0;
}

// 1-D Hermite table e[(i (GI+1) + j) GT + t] = E^{ij}_t for i <= imax, j <= jmax,
// E^{00}_0 = exp(-mu xab^2), the recursion of md::hermite_e with runtime bounds
__device__ void e1d(double a, double b, double xab, int imax, int jmax, double* e) {
    const double p = a + b;
    const double xpa = -b * xab / p, xpb = a * xab / p, h = 0.5 / p;
    for (int i = 0; i <= imax; ++i)
        for (int j = 0; j <= jmax; ++j)
            for (int t = 0; t < GT; ++t) e[(i * (GI + 1) + j) * GT + t] = 0.0;
    e[0] = exp(-a * b / p * xab * xab);
    for (int j = 1; j <= jmax; ++j) {
        const double* q = e + (j - 1) * GT;
        for (int t = 0; t <= j; ++t) {
            double v = xpb * q[t];
            if (t + 1 <= j - 1) v += (t + 1) * q[t + 1];
            if (t > 0) v += h * q[t - 1];
            e[j * GT + t] = v;
        }
    }
    for (int i = 1; i <= imax; ++i)
        for (int j = 0; j <= jmax; ++j) {
            const double* q = e + ((i - 1) * (GI + 1) + j) * GT;
            for (int t = 0; t <= i + j; ++t) {
                double v = xpa * q[t];
                if (t + 1 <= i - 1 + j) v += (t + 1) * q[t + 1];
                if (t > 0) v += h * q[t - 1];
                e[(i * (GI + 1) + j) * GT + t] = v;
            }
        }
}

__device__ __forceinline__ const double* erow(const double* e, int d, int i, int j) {
    return e + d * GE + (i * (GI + 1) + j) * GT;
}

// (t, u, v) of linear Hermite index h (degree-major, t descending, then u descending)
__device__ __forceinline__ void hdecode(int h, int& t, int& u, int& v) {
    int d = 0;
    while (NH(d) <= h) ++d;
    const int r = h - NH(d - 1);
    int m = 0;
    while ((m + 1) * (m + 2) / 2 <= r) ++m;
    u = m - (r - m * (m + 1) / 2);
    t = d - m;
    v = m - u;
}

// R^0_{tuv}(alpha, (x, y, z)) * scale for t+u+v <= L, built by the whole
// wavefront; returns the LDS buffer holding it. Ends with a barrier.
__device__ const double* r_table(Lds& s, int L, double alpha, double x, double y, double z, double scale,
                                 const double* __restrict__ boys) {
    if (threadIdx.x == 0) {
        const double T = alpha * (x * x + y * y + z * z);
        double* F = s.f;
        F[L] = boys_top<BOYS_NCOL, 0, double>(L, T, boys);
        const double ex = exp(-T);
        for (int n = L - 1; n >= 0; --n) F[n] = (2.0 * T * F[n + 1] + ex) / (2 * n + 1);
        double pw = scale;
        for (int n = 0; n <= L; ++n) {
            F[n] *= pw;
            pw *= -2.0 * alpha;
        }
        s.r[0] = F[L];
    }
    __syncthreads();
    double* in = s.r;
    double* out = s.r + GNH;
    for (int n = L - 1; n >= 0; --n) {
        const int nh = NH(L - n);
        for (int h = threadIdx.x; h < nh; h += blockDim.x) {
            if (h == 0) {
                out[0] = s.f[n];
                continue;
            }
            int t, u, v;
            hdecode(h, t, u, v);
            double r;
            if (t > 0) {
                r = x * in[hidx(t - 1, u, v)];
                if (t > 1) r += (t - 1) * in[hidx(t - 2, u, v)];
            } else if (u > 0) {
                r = y * in[hidx(0, u - 1, v)];
                if (u > 1) r += (u - 1) * in[hidx(0, u - 2, v)];
            } else {
                r = z * in[hidx(0, 0, v - 1)];
                if (v > 1) r += (v - 1) * in[hidx(0, 0, v - 2)];
            }
            out[h] = r;
        }
        __syncthreads();
        double* tmp = in;
        in = out;
        out = tmp;
    }
    return in;
}

// sum_{tuv} E^{ab}_{tuv} sum_{tau nu phi} (-1)^{tau+nu+phi} E^{cd}_{tau nu phi} R_{t+tau, u+nu, v+phi}
__device__ double eri_elem(const double* eb, const double* ek, const double* R, int ax, int ay, int az, int bx,
                           int by, int bz, int cx, int cy, int cz, int dx, int dy, int dz) {
    const double* ex = erow(eb, 0, ax, bx);
    const double* ey = erow(eb, 1, ay, by);
    const double* ez = erow(eb, 2, az, bz);
    const double* kx = erow(ek, 0, cx, dx);
    const double* ky = erow(ek, 1, cy, dy);
    const double* kz = erow(ek, 2, cz, dz);
    const int sx = cx + dx, sy = cy + dy, sz = cz + dz;
    double s = 0.0;
    for (int t = 0; t <= ax + bx; ++t)
        for (int u = 0; u <= ay + by; ++u) {
            const double e2 = ex[t] * ey[u];
            for (int v = 0; v <= az + bz; ++v) {
                double in = 0.0;
                for (int tau = 0; tau <= sx; ++tau) {
                    const double f1 = (tau & 1) ? -kx[tau] : kx[tau];
                    for (int nu = 0; nu <= sy; ++nu) {
                        const double f2 = f1 * ((nu & 1) ? -ky[nu] : ky[nu]);
                        for (int ph = 0; ph <= sz; ++ph) {
                            const double f3 = (ph & 1) ? -kz[ph] : kz[ph];
                            in += f2 * f3 * R[hidx(t + tau, u + nu, v + ph)];
                        }
                    }
                }
                s += e2 * ez[v] * in;
            }
        }
    return s;
}

// sum_{tuv} E^{ab}_{tuv} R_{t+ox, u+oy, v+oz} (one-electron nuclear attraction; o: derivative offset)
__device__ double nuc_elem(const double* eb, const double* R, int ax, int ay, int az, int bx, int by, int bz, int ox,
                           int oy, int oz) {
    const double* ex = erow(eb, 0, ax, bx);
    const double* ey = erow(eb, 1, ay, by);
    const double* ez = erow(eb, 2, az, bz);
    double s = 0.0;
    for (int t = 0; t <= ax + bx; ++t)
        for (int u = 0; u <= ay + by; ++u)
            for (int v = 0; v <= az + bz; ++v) s += ex[t] * ey[u] * ez[v] * R[hidx(t + ox, u + oy, v + oz)];
    return s;
}

struct Shell {
    int atom, l, p0, p1, ao, nc;
};

__device__ __forceinline__ Shell shell(const GenArgs& g, int s) {
    Shell r;
    r.atom = g.shell_atom[s];
    r.l = g.shell_l[s];
    r.p0 = g.shell_prim[s];
    r.p1 = g.shell_prim[s + 1];
    r.ao = g.shell_ao[s];
    r.nc = ncart_of(r.l);
    return r;
}

// Walk the primitive quartets of (I J | K L): per kept bra pair, lanes 0-2 build
// the bra tables (rows up to l + sh), per kept ket pair the ket tables, then R
// of degree L; f(R, a, b, c, d) runs on every lane with the four exponents.
template <class Fn>
__device__ void quartet_walk(const GenArgs& g, Lds& s, const Shell& I, const Shell& J, const Shell& K, const Shell& Lsh,
                             int shi, int shj, int shk, int shl, int L, Fn&& f) {
    const int lane = threadIdx.x;
    const double* A = g.pos + 3 * I.atom;
    const double* B = g.pos + 3 * J.atom;
    const double* C = g.pos + 3 * K.atom;
    const double* D = g.pos + 3 * Lsh.atom;
    double rab = 0.0, rcd = 0.0;
    for (int d = 0; d < 3; ++d) {
        rab += (A[d] - B[d]) * (A[d] - B[d]);
        rcd += (C[d] - D[d]) * (C[d] - D[d]);
    }
    for (int ia = I.p0; ia < I.p1; ++ia)
        for (int ib = J.p0; ib < J.p1; ++ib) {
            const double a = g.prim_exp[ia], b = g.prim_exp[ib];
            if (!(a * b / (a + b) * rab <= 50.0)) continue;
            const double p = a + b;
            __syncthreads();                    // the previous pair's readers are done
            if (lane < 3) e1d(a, b, A[lane] - B[lane], I.l + shi, J.l + shj, s.eb + lane * GE);
            for (int ic = K.p0; ic < K.p1; ++ic)
                for (int id = Lsh.p0; id < Lsh.p1; ++id) {
                    const double c = g.prim_exp[ic], dd = g.prim_exp[id];
                    if (!(c * dd / (c + dd) * rcd <= 50.0)) continue;
                    const double q = c + dd;
                    __syncthreads();
                    if (lane < 3) e1d(c, dd, C[lane] - D[lane], K.l + shk, Lsh.l + shl, s.ek + lane * GE);
                    double P[3], Q[3];
                    for (int k = 0; k < 3; ++k) {
                        P[k] = (a * A[k] + b * B[k]) / p;
                        Q[k] = (c * C[k] + dd * D[k]) / q;
                    }
                    const double alpha = p * q / (p + q);
                    const double pref = TWO_PI_2_5 / (p * q * sqrt(p + q)) * g.prim_coef[ia] * g.prim_coef[ib] *
                                        g.prim_coef[ic] * g.prim_coef[id];
                    __syncthreads();                // tables written
                    const double* R = r_table(s, L, alpha, P[0] - Q[0], P[1] - Q[1], P[2] - Q[2], pref, g.boys);
                    f(R, a, b, c, dd);
                }
        }
    __syncthreads();
}

// ----------------------------------------------------------------- Schwarz
// q[P] = sqrt(max_ab |(ab|ab)|) of shell pair P (i >= j)
__global__ __launch_bounds__(WG) void k_gen_schwarz(GenArgs g) {
    __shared__ Lds s;
    const int P = blockIdx.x;
    const Shell I = shell(g, g.pair_i[P]), J = shell(g, g.pair_j[P]);
    const int ne = I.nc * J.nc;
    double best = 0.0;
    for (int e0 = 0; e0 < ne; e0 += WG) {
        const int e = e0 + threadIdx.x;
        const bool ok = e < ne;
        int ax = 0, ay = 0, az = 0, bx = 0, by = 0, bz = 0;
        if (ok) {
            comp(I.l, e / J.nc, ax, ay, az);
            comp(J.l, e % J.nc, bx, by, bz);
        }
        double acc = 0.0;
        quartet_walk(g, s, I, J, I, J, 0, 0, 0, 0, 2 * (I.l + J.l), [&](const double* R, double, double, double, double) {
            if (ok) acc += eri_elem(s.eb, s.ek, R, ax, ay, az, bx, by, bz, ax, ay, az, bx, by, bz);
        });
        if (ok) {
            const double f = cart_factor(ax, ay, az) * cart_factor(bx, by, bz);
            best = fmax(best, fabs(acc * f * f));
        }
    }
    best = wave_max(best);
    if (threadIdx.x == 0) g.q[P] = sqrt(best);
}

// ------------------------------------------------------------------ J / K
// unique quartet (IJ|KL), P = (i >= j) >= Q = (k >= l): every element adds its
// 8-fold images to J' and K' (J = J' + J'^T, K = K' + K'^T, k_jk_finalize)
__global__ __launch_bounds__(WG) void k_gen_jk(GenArgs g) {
    __shared__ Lds s;
    const size_t n = g.nao, nn = n * n;
    for (long long qi = blockIdx.x; qi < g.nq; qi += gridDim.x) {
        const int2 pq = g.quartets[qi];
        const int si = g.pair_i[pq.x], sj = g.pair_j[pq.x], sk = g.pair_i[pq.y], sl = g.pair_j[pq.y];
        const Shell I = shell(g, si), J = shell(g, sj), K = shell(g, sk), Lsh = shell(g, sl);
        double deg = 1.0;
        if (si == sj) deg *= 0.5;
        if (sk == sl) deg *= 0.5;
        if (pq.x == pq.y) deg *= 0.5;
        const int nb = J.nc, nc = K.nc, nd = Lsh.nc;
        const int ne = I.nc * nb * nc * nd;
        const int L = I.l + J.l + K.l + Lsh.l;
        for (int e0 = 0; e0 < ne; e0 += WG) {
            const int e = e0 + threadIdx.x;
            const bool ok = e < ne;
            int ca = 0, cb = 0, cc = 0, cd = 0;
            int ax = 0, ay = 0, az = 0, bx = 0, by = 0, bz = 0, cx = 0, cy = 0, cz = 0, dx = 0, dy = 0, dz = 0;
            if (ok) {
                cd = e % nd;
                cc = (e / nd) % nc;
                cb = (e / (nd * nc)) % nb;
                ca = e / (nd * nc * nb);
                comp(I.l, ca, ax, ay, az);
                comp(J.l, cb, bx, by, bz);
                comp(K.l, cc, cx, cy, cz);
                comp(Lsh.l, cd, dx, dy, dz);
            }
            double acc = 0.0;
            quartet_walk(g, s, I, J, K, Lsh, 0, 0, 0, 0, L, [&](const double* R, double, double, double, double) {
                if (ok) acc += eri_elem(s.eb, s.ek, R, ax, ay, az, bx, by, bz, cx, cy, cz, dx, dy, dz);
            });
            if (!ok) continue;
            const double v = deg * acc * cart_factor(ax, ay, az) * cart_factor(bx, by, bz) * cart_factor(cx, cy, cz) *
                             cart_factor(dx, dy, dz);
            const size_t a = I.ao + ca, b = J.ao + cb, c = K.ao + cc, d = Lsh.ao + cd;
            for (int st = 0; st < g.nset; ++st) {
                const double* Dm = g.dens + st * nn;
                long long* jacc = g.fx + (size_t)(4 * st) * nn;
                long long* kacc = g.fx + (size_t)(4 * st + 2) * nn;
                // J' and K' of set st: a plane of high words, then a plane of low words (k_jk_finalize)
                auto add = [&](long long* base, size_t idx, double x) {
                    long long h, l;
                    fx_split(x, h, l);
                    if (h) atomicAdd((unsigned long long*)&base[idx], (unsigned long long)h);
                    if (l) atomicAdd((unsigned long long*)&base[nn + idx], (unsigned long long)l);
                };
                add(jacc, a * n + b, 2.0 * v * Dm[c * n + d]);
                add(jacc, c * n + d, 2.0 * v * Dm[a * n + b]);
                add(kacc, a * n + c, v * Dm[b * n + d]);
                add(kacc, b * n + c, v * Dm[a * n + d]);
                add(kacc, a * n + d, v * Dm[b * n + c]);
                add(kacc, b * n + d, v * Dm[a * n + c]);
            }
        }
    }
}

// --------------------------------------------------------- derivative ERIs
// d/dR of 1/2 sum (pq|rs) [Dc_pq Dc_rs - kfac sum_x Dx_pr Dx_qs] over the
// unique quartets: per element G = deg [4 Dc_ab Dc_cd - 2 kfac sum_x (Dx_ac Dx_bd
// + Dx_ad Dx_bc)] times the derivative integrals on centres A, B, C; D = -(A+B+C)
__global__ __launch_bounds__(WG) void k_gen_grad2(GenArgs g) {
    __shared__ Lds s;
    const size_t n = g.nao, nn = n * n;
    for (long long qi = blockIdx.x; qi < g.nq; qi += gridDim.x) {
        const int2 pq = g.quartets[qi];
        const int si = g.pair_i[pq.x], sj = g.pair_j[pq.x], sk = g.pair_i[pq.y], sl = g.pair_j[pq.y];
        const Shell I = shell(g, si), J = shell(g, sj), K = shell(g, sk), Lsh = shell(g, sl);
        // no force between centres on one atom
        if (I.atom == J.atom && I.atom == K.atom && I.atom == Lsh.atom) continue;
        double deg = 1.0;
        if (si == sj) deg *= 0.5;
        if (sk == sl) deg *= 0.5;
        if (pq.x == pq.y) deg *= 0.5;
        const int nb = J.nc, nc = K.nc, nd = Lsh.nc;
        const int ne = I.nc * nb * nc * nd;
        const int L = I.l + J.l + K.l + Lsh.l + 1;
        double gsum[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        for (int e0 = 0; e0 < ne; e0 += WG) {
            const int e = e0 + threadIdx.x;
            const bool ok = e < ne;
            int ca = 0, cb = 0, cc = 0, cd = 0;
            int ax = 0, ay = 0, az = 0, bx = 0, by = 0, bz = 0, cx = 0, cy = 0, cz = 0, dx = 0, dy = 0, dz = 0;
            if (ok) {
                cd = e % nd;
                cc = (e / nd) % nc;
                cb = (e / (nd * nc)) % nb;
                ca = e / (nd * nc * nb);
                comp(I.l, ca, ax, ay, az);
                comp(J.l, cb, bx, by, bz);
                comp(K.l, cc, cx, cy, cz);
                comp(Lsh.l, cd, dx, dy, dz);
            }
            double G = 0.0;
            if (ok) {
                const size_t a = I.ao + ca, b = J.ao + cb, c = K.ao + cc, d = Lsh.ao + cd;
                double ex = 0.0;
                for (int x = 0; x < g.nx; ++x) {
                    const double* X = g.dx + x * nn;
                    ex += X[a * n + c] * X[b * n + d] + X[a * n + d] * X[b * n + c];
                }
                G = deg * (4.0 * g.dc[a * n + b] * g.dc[c * n + d] - 2.0 * g.kfac * ex) * cart_factor(ax, ay, az) *
                    cart_factor(bx, by, bz) * cart_factor(cx, cy, cz) * cart_factor(dx, dy, dz);
            }
            double d9[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
            quartet_walk(g, s, I, J, K, Lsh, 1, 1, 1, 0, L, [&](const double* R, double a, double b, double c, double) {
                if (!ok || G == 0.0) return;
                const int aa[3] = {ax, ay, az}, bb[3] = {bx, by, bz}, cc3[3] = {cx, cy, cz};
                for (int k = 0; k < 3; ++k) {
                    int u[3] = {aa[0], aa[1], aa[2]};
                    u[k] += 1;
                    double v = 2.0 * a * eri_elem(s.eb, s.ek, R, u[0], u[1], u[2], bx, by, bz, cx, cy, cz, dx, dy, dz);
                    if (aa[k] > 0) {
                        u[k] -= 2;
                        v -= aa[k] * eri_elem(s.eb, s.ek, R, u[0], u[1], u[2], bx, by, bz, cx, cy, cz, dx, dy, dz);
                    }
                    d9[k] += v;
                    int w[3] = {bb[0], bb[1], bb[2]};
                    w[k] += 1;
                    v = 2.0 * b * eri_elem(s.eb, s.ek, R, ax, ay, az, w[0], w[1], w[2], cx, cy, cz, dx, dy, dz);
                    if (bb[k] > 0) {
                        w[k] -= 2;
                        v -= bb[k] * eri_elem(s.eb, s.ek, R, ax, ay, az, w[0], w[1], w[2], cx, cy, cz, dx, dy, dz);
                    }
                    d9[3 + k] += v;
                    int z[3] = {cc3[0], cc3[1], cc3[2]};
                    z[k] += 1;
                    v = 2.0 * c * eri_elem(s.eb, s.ek, R, ax, ay, az, bx, by, bz, z[0], z[1], z[2], dx, dy, dz);
                    if (cc3[k] > 0) {
                        z[k] -= 2;
                        v -= cc3[k] * eri_elem(s.eb, s.ek, R, ax, ay, az, bx, by, bz, z[0], z[1], z[2], dx, dy, dz);
                    }
                    d9[6 + k] += v;
                }
            });
            for (int k = 0; k < 9; ++k) gsum[k] += G * d9[k];
        }
        for (int k = 0; k < 9; ++k) gsum[k] = wave_sum(gsum[k]);
        if (threadIdx.x == 0) {
            for (int k = 0; k < 3; ++k) {
                fx_add(g.fx, 3 * I.atom + k, gsum[k]);
                fx_add(g.fx, 3 * J.atom + k, gsum[3 + k]);
                fx_add(g.fx, 3 * K.atom + k, gsum[6 + k]);
                fx_add(g.fx, 3 * Lsh.atom + k, -(gsum[k] + gsum[3 + k] + gsum[6 + k]));
            }
        }
    }
}

// ------------------------------------------------------------ one-electron
// Overlap-type 1-D factors of component pair (i, j) from the table rows:
// s = E^{ij}_0, t = kinetic, m = dipole moment about the origin (all without sqrt(pi/p))
__device__ __forceinline__ void one_d(const double* e, int d, int i, int j, double b, double P, double& s, double& t,
                                      double& m) {
    s = erow(e, d, i, j)[0];
    t = -2.0 * b * b * erow(e, d, i, j + 2)[0] + b * (2 * j + 1) * s;
    if (j >= 2) t -= 0.5 * j * (j - 1) * erow(e, d, i, j - 2)[0];
    m = erow(e, d, i, j)[1] + P * s;
}

// S, T, V and the dipole integrals M of shell pair P (i >= j), both orientations written
__global__ __launch_bounds__(WG) void k_gen_oneel(GenArgs g) {
    __shared__ Lds s;
    const int P = blockIdx.x;
    const Shell I = shell(g, g.pair_i[P]), J = shell(g, g.pair_j[P]);
    const int ne = I.nc * J.nc;
    const double* A = g.pos + 3 * I.atom;
    const double* B = g.pos + 3 * J.atom;
    const size_t n = g.nao, nn = n * n;
    for (int e0 = 0; e0 < ne; e0 += WG) {
        const int e = e0 + threadIdx.x;
        const bool ok = e < ne;
        int ax = 0, ay = 0, az = 0, bx = 0, by = 0, bz = 0;
        if (ok) {
            comp(I.l, e / J.nc, ax, ay, az);
            comp(J.l, e % J.nc, bx, by, bz);
        }
        double S = 0, T = 0, V = 0, M[3] = {0, 0, 0};
        for (int ia = I.p0; ia < I.p1; ++ia)
            for (int ib = J.p0; ib < J.p1; ++ib) {
                const double a = g.prim_exp[ia], b = g.prim_exp[ib], p = a + b;
                __syncthreads();
                if (threadIdx.x < 3)
                    e1d(a, b, A[threadIdx.x] - B[threadIdx.x], I.l, J.l + 2, s.eb + threadIdx.x * GE);
                __syncthreads();
                const double cab = g.prim_coef[ia] * g.prim_coef[ib];
                double Pc[3];
                for (int k = 0; k < 3; ++k) Pc[k] = (a * A[k] + b * B[k]) / p;
                if (ok) {
                    const double sp = sqrt(MD_PI / p);
                    const int aa[3] = {ax, ay, az}, bb[3] = {bx, by, bz};
                    double s1[3], t1[3], m1[3];
                    for (int d = 0; d < 3; ++d) one_d(s.eb, d, aa[d], bb[d], b, Pc[d], s1[d], t1[d], m1[d]);
                    const double c3 = cab * sp * sp * sp;
                    S += c3 * s1[0] * s1[1] * s1[2];
                    T += c3 * (t1[0] * s1[1] * s1[2] + s1[0] * t1[1] * s1[2] + s1[0] * s1[1] * t1[2]);
                    M[0] += c3 * m1[0] * s1[1] * s1[2];
                    M[1] += c3 * s1[0] * m1[1] * s1[2];
                    M[2] += c3 * s1[0] * s1[1] * m1[2];
                }
                for (int C = 0; C < g.natm; ++C) {
                    const double* Cp = g.pos + 3 * C;
                    const double* R = r_table(s, I.l + J.l, p, Pc[0] - Cp[0], Pc[1] - Cp[1], Pc[2] - Cp[2],
                                              -g.charges[C] * 2.0 * MD_PI / p * cab, g.boys);
                    if (ok) V += nuc_elem(s.eb, R, ax, ay, az, bx, by, bz, 0, 0, 0);
                    __syncthreads();
                }
            }
        if (!ok) continue;
        const double f = cart_factor(ax, ay, az) * cart_factor(bx, by, bz);
        const size_t i = I.ao + e / J.nc, j = J.ao + e % J.nc;
        const double vals[6] = {S * f, T * f, V * f, M[0] * f, M[1] * f, M[2] * f};
        double* outs[6] = {g.S, g.T, g.V, g.M, g.M + nn, g.M + 2 * nn};
        for (int k = 0; k < 6; ++k) {
            outs[k][i * n + j] = vals[k];
            outs[k][j * n + i] = vals[k];
        }
    }
}

// d/dR of sum_ab [D_ab (T_ab + V_ab) - W_ab S_ab] over shell pair P (i >= j):
// parts bit 1 the kinetic term, 2 the nuclear attraction, 4 the overlap term
__global__ __launch_bounds__(WG) void k_gen_grad1(GenArgs g) {
    __shared__ Lds s;
    const int P = blockIdx.x;
    const Shell I = shell(g, g.pair_i[P]), J = shell(g, g.pair_j[P]);
    const int ne = I.nc * J.nc;
    const double* A = g.pos + 3 * I.atom;
    const double* B = g.pos + 3 * J.atom;
    const size_t n = g.nao;
    const double w = g.pair_i[P] == g.pair_j[P] ? 1.0 : 2.0;
    const bool kin = g.parts & 1, nuc = g.parts & 2, ovl = g.parts & 4;
    double gA[3] = {0, 0, 0}, gB[3] = {0, 0, 0};
    for (int e0 = 0; e0 < ne; e0 += WG) {
        const int e = e0 + threadIdx.x;
        const bool ok = e < ne;
        int ax = 0, ay = 0, az = 0, bx = 0, by = 0, bz = 0;
        double Dab = 0.0, Wab = 0.0;
        if (ok) {
            comp(I.l, e / J.nc, ax, ay, az);
            comp(J.l, e % J.nc, bx, by, bz);
            const double f = w * cart_factor(ax, ay, az) * cart_factor(bx, by, bz);
            const size_t i = I.ao + e / J.nc, j = J.ao + e % J.nc;
            Dab = f * g.D[i * n + j];
            Wab = f * g.W[i * n + j];
        }
        const int aa[3] = {ax, ay, az}, bb[3] = {bx, by, bz};
        for (int C = -1; C < (nuc ? g.natm : 0); ++C) {
            // C = -1: overlap and kinetic; C >= 0: attraction to nucleus C
            double gC[3] = {0, 0, 0};
            for (int ia = I.p0; ia < I.p1; ++ia)
                for (int ib = J.p0; ib < J.p1; ++ib) {
                    const double a = g.prim_exp[ia], b = g.prim_exp[ib], p = a + b;
                    __syncthreads();
                    if (threadIdx.x < 3)
                        e1d(a, b, A[threadIdx.x] - B[threadIdx.x], I.l + 1, J.l + (C < 0 ? 2 : 1),
                            s.eb + threadIdx.x * GE);
                    __syncthreads();
                    const double cab = g.prim_coef[ia] * g.prim_coef[ib];
                    double Pc[3];
                    for (int k = 0; k < 3; ++k) Pc[k] = (a * A[k] + b * B[k]) / p;
                    if (C < 0) {
                        if (!ok || !(kin || ovl)) continue;
                        const double sp = sqrt(MD_PI / p);
                        const double c3 = cab * sp * sp * sp;
                        double s0[3], t0[3], m0[3], sp1[3], tp1[3], sm1[3], tm1[3];
                        for (int d = 0; d < 3; ++d) {
                            one_d(s.eb, d, aa[d], bb[d], b, Pc[d], s0[d], t0[d], m0[d]);
                            one_d(s.eb, d, aa[d] + 1, bb[d], b, Pc[d], sp1[d], tp1[d], m0[d]);
                            sm1[d] = tm1[d] = 0.0;
                            if (aa[d] > 0) one_d(s.eb, d, aa[d] - 1, bb[d], b, Pc[d], sm1[d], tm1[d], m0[d]);
                        }
                        for (int k = 0; k < 3; ++k) {
                            // d/dA_k of the 1-D factor k, the others unchanged
                            const double ds = 2.0 * a * sp1[k] - aa[k] * sm1[k];
                            const double dt = 2.0 * a * tp1[k] - aa[k] * tm1[k];
                            const int k1 = (k + 1) % 3, k2 = (k + 2) % 3;
                            const double dS = c3 * ds * s0[k1] * s0[k2];
                            const double dT = c3 * (dt * s0[k1] * s0[k2] + ds * t0[k1] * s0[k2] + ds * s0[k1] * t0[k2]);
                            double v = 0.0;
                            if (kin) v += Dab * dT;
                            if (ovl) v -= Wab * dS;
                            gA[k] += v;
                            gB[k] -= v;
                        }
                        continue;
                    }
                    const double* Cp = g.pos + 3 * C;
                    const double* R = r_table(s, I.l + J.l + 1, p, Pc[0] - Cp[0], Pc[1] - Cp[1], Pc[2] - Cp[2],
                                              -g.charges[C] * 2.0 * MD_PI / p * cab, g.boys);
                    if (ok && Dab != 0.0)
                        for (int k = 0; k < 3; ++k) {
                            int u[3] = {ax, ay, az};
                            u[k] += 1;
                            double v = 2.0 * a * nuc_elem(s.eb, R, u[0], u[1], u[2], bx, by, bz, 0, 0, 0);
                            if (aa[k] > 0) {
                                u[k] -= 2;
                                v -= aa[k] * nuc_elem(s.eb, R, u[0], u[1], u[2], bx, by, bz, 0, 0, 0);
                            }
                            gA[k] += Dab * v;
                            int z[3] = {bx, by, bz};
                            z[k] += 1;
                            v = 2.0 * b * nuc_elem(s.eb, R, ax, ay, az, z[0], z[1], z[2], 0, 0, 0);
                            if (bb[k] > 0) {
                                z[k] -= 2;
                                v -= bb[k] * nuc_elem(s.eb, R, ax, ay, az, z[0], z[1], z[2], 0, 0, 0);
                            }
                            gB[k] += Dab * v;
                            // R depends on P - C: d/dC_k = -d/dP_k, and d/dP_k R_{tuv} = R_{t+1,..}
                            gC[k] -= Dab * nuc_elem(s.eb, R, ax, ay, az, bx, by, bz, k == 0, k == 1, k == 2);
                        }
                    __syncthreads();
                }
            if (C >= 0) {
                for (int k = 0; k < 3; ++k) gC[k] = wave_sum(gC[k]);
                if (threadIdx.x == 0)
                    for (int k = 0; k < 3; ++k)
                        if (gC[k] != 0.0) fx_add(g.fx, 3 * C + k, gC[k]);
            }
        }
    }
    for (int k = 0; k < 3; ++k) {
        gA[k] = wave_sum(gA[k]);
        gB[k] = wave_sum(gB[k]);
    }
    if (threadIdx.x == 0)
        for (int k = 0; k < 3; ++k) {
            fx_add(g.fx, 3 * I.atom + k, gA[k]);
            fx_add(g.fx, 3 * J.atom + k, gB[k]);
        }
}

}  // namespace

void launch_gen_schwarz(const GenArgs& g, int npair, hipStream_t st) {
    if (npair > 0) k_gen_schwarz<<<dim3(npair), dim3(WG), 0, st>>>(g);
}

void launch_gen_jk(const GenArgs& g, hipStream_t st) {
    if (g.nq > 0) k_gen_jk<<<dim3((unsigned)(g.nq < (1 << 16) ? g.nq : (1 << 16))), dim3(WG), 0, st>>>(g);
}

void launch_gen_grad2(const GenArgs& g, hipStream_t st) {
    if (g.nq > 0) k_gen_grad2<<<dim3((unsigned)(g.nq < (1 << 16) ? g.nq : (1 << 16))), dim3(WG), 0, st>>>(g);
}

void launch_gen_oneel(const GenArgs& g, int npair, hipStream_t st) {
    if (npair > 0) k_gen_oneel<<<dim3(npair), dim3(WG), 0, st>>>(g);
}

void launch_gen_grad1(const GenArgs& g, int npair, hipStream_t st) {
    if (npair > 0) k_gen_grad1<<<dim3(npair), dim3(WG), 0, st>>>(g);
}

}  // namespace aimdhf
