// Batched plans: many independent small molecules (replicas) in one plan. Their
// screened quartets are evaluated into the class-binned store by the ordinary
// per-class store kernels (one launch per class for all replicas); this file
// expands the stored blocks into one dense (n_r^4) integral tensor per replica
// and digests those into J and K for every replica in a single launch.
//
// Matrices of a batch live in a padded layout: replica r's (nset x) ld x ld
// block starts at (r * nset + x) * ld^2; rows and columns past n_r stay zero.

#include "aimdhf_internal.h"

namespace aimdhf {

namespace {

__device__ __forceinline__ int ncart_of(int l) { return (l + 1) * (l + 2) / 2; }

// one thread per stored integral: class by binary search over the class
// offsets, then the integral's four AOs; the value goes to all eight
// permutation-equivalent places of its replica's tensor (the same value where
// places coincide)
__global__ __launch_bounds__(256) void k_batch_scatter(BatchArgs b, long long total) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= total) return;
    int k = 0;
    for (int step = 16; step > 0; step >>= 1)
        if (k + step <= NQCLASS_B - 1 && b.store_base[k + step] <= t) k += step;
    const long long local = t - b.store_base[k];
    const int blk = b.blk[k], ncd = b.ncd[k];
    const long long q = b.qoff[k] + local / blk;
    const int e = (int)(local % blk), ab = e / ncd, cd = e % ncd;
    const int2 pq = b.list[q];
    const int A = b.pair_a[pq.x], B = b.pair_b[pq.x], C = b.pair_a[pq.y], D = b.pair_b[pq.y];
    const int nb = ncart_of(b.shell_l[B]), nd = ncart_of(b.shell_l[D]);
    const long long m = b.shell_ao[A] + ab / nb, n = b.shell_ao[B] + ab % nb;
    const long long l = b.shell_ao[C] + cd / nd, s = b.shell_ao[D] + cd % nd;
    const int r = b.shell_rep[A];
    const long long N = b.rep_nao[r];
    double* T = b.tens + b.rep_toff[r];
    const double v = b.store[t];
    const long long mn = m * N + n, nm = n * N + m, ls = l * N + s, sl = s * N + l, NN = N * N;
    T[mn * NN + ls] = v;
    T[nm * NN + ls] = v;
    T[mn * NN + sl] = v;
    T[nm * NN + sl] = v;
    T[ls * NN + mn] = v;
    T[sl * NN + mn] = v;
    T[ls * NN + nm] = v;
    T[sl * NN + nm] = v;
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// one wavefront per output element (r, m, c):
//   J_x[m][c] = sum_{l,s} (m c|l s) D_x[l][s]     (the contiguous slab T[m][c])
//   K_x[m][c] = sum_{l,s} (m l|c s) D_x[l][s]     (rows T[m][l][c][.], lanes over s)
// fixed per-lane order and a butterfly sum: bitwise reproducible
__global__ __launch_bounds__(256) void k_batch_jk(BatchArgs b, int nset, const double* __restrict__ dens,
                                                  double* __restrict__ J, double* __restrict__ K, long long nitems) {
    const long long item = ((long long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    if (item >= nitems) return;
    int r = 0;
    for (int step = 1 << 20; step > 0; step >>= 1)
        if (r + step < b.nrep && b.rep_item[r + step] <= item) r += step;
    const long long loc = item - b.rep_item[r];
    const int N = b.rep_nao[r];
    const int m = (int)(loc / N), c = (int)(loc % N);
    const double* T = b.tens + b.rep_toff[r];
    const size_t ld = b.ld, ld2 = ld * ld;
    const double* D0 = dens + (size_t)r * nset * ld2;
    double ja[2] = {0.0, 0.0}, ka[2] = {0.0, 0.0};
    const double* slab = T + ((size_t)m * N + c) * N * N;
    for (int e = lane; e < N * N; e += 64) {
        const double t = slab[e];
        const size_t de = (size_t)(e / N) * ld + e % N;
        ja[0] += t * D0[de];
        if (nset == 2) ja[1] += t * D0[ld2 + de];
    }
    for (int s = lane; s < N; s += 64)
        for (int l = 0; l < N; ++l) {
            const double t = T[(((size_t)m * N + l) * N + c) * N + s];
            const size_t de = (size_t)l * ld + s;
            ka[0] += t * D0[de];
            if (nset == 2) ka[1] += t * D0[ld2 + de];
        }
    for (int x = 0; x < nset; ++x) {
        const double jv = wave_sum(ja[x]), kv = wave_sum(ka[x]);
        if (lane == 0) {
            const size_t o = ((size_t)r * nset + x) * ld2 + (size_t)m * ld + c;
            J[o] = jv;
            K[o] = kv;
        }
    }
}

// Trace-correcting purification of nmat symmetric ld x ld matrices (ld <= 64), one
// workgroup per matrix held in LDS for all steps: X0 = (e_max - F) / (e_max -
// e_min) with F's Gershgorin bounds, then X <- X^2 if tr X > nocc else 2X - X^2,
// an iterate with |tr X - tr X^2| < tol held (more steps would amplify its
// rounding errors), and two final unheld steps that square the remaining error.
// err[i] receives the idempotency error |tr X - tr X^2| before those two steps.
constexpr int PUR_THREADS = 256;
constexpr int PUR_MAXLD = 64;
constexpr int PUR_PER_THREAD = PUR_MAXLD * PUR_MAXLD / PUR_THREADS;

__device__ __forceinline__ double block_sum(double v, double* red) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    return red[0] + red[1] + red[2] + red[3];
}

__global__ __launch_bounds__(PUR_THREADS) void k_batch_purify(int ld, const double* __restrict__ F,
                                                              const double* __restrict__ nocc, double* __restrict__ P,
                                                              int steps, double tol, double* __restrict__ err) {
    __shared__ double X[PUR_MAXLD * PUR_MAXLD];
    __shared__ double red[4], lo[4], hi[4];
    const int mat = blockIdx.x, t = threadIdx.x, n2 = ld * ld;
    const double* Fm = F + (size_t)mat * n2;
    // Gershgorin bounds, thread per row (min / max do not depend on the order)
    double mn = 1e300, mx = -1e300;
    for (int i = t; i < ld; i += PUR_THREADS) {
        double r = 0.0;
        for (int j = 0; j < ld; ++j) r += j == i ? 0.0 : fabs(Fm[i * ld + j]);
        mn = fmin(mn, Fm[i * ld + i] - r);
        mx = fmax(mx, Fm[i * ld + i] + r);
    }
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        mn = fmin(mn, __shfl_xor(mn, o, 64));
        mx = fmax(mx, __shfl_xor(mx, o, 64));
    }
    if ((t & 63) == 0) {
        lo[t >> 6] = mn;
        hi[t >> 6] = mx;
    }
    __syncthreads();
    const double emin = fmin(fmin(lo[0], lo[1]), fmin(lo[2], lo[3]));
    const double emax = fmax(fmax(hi[0], hi[1]), fmax(hi[2], hi[3]));
    const double inv = 1.0 / (emax - emin);
    for (int e = t; e < n2; e += PUR_THREADS) X[e] = ((e / ld == e % ld ? emax : 0.0) - Fm[e]) * inv;
    __syncthreads();
    const double ne = nocc[mat];
    double x2[PUR_PER_THREAD];
    double last = 0.0;
    for (int it = 0; it < steps + 2; ++it) {
        double dx = 0.0, dx2 = 0.0;
#pragma unroll
        for (int k = 0; k < PUR_PER_THREAD; ++k) {
            const int e = t + k * PUR_THREADS;
            double v = 0.0;
            if (e < n2) {
                const int i = e / ld, j = e % ld;
                for (int q = 0; q < ld; ++q) v += X[i * ld + q] * X[q * ld + j];
                if (i == j) {
                    dx += X[e];
                    dx2 += v;
                }
            }
            x2[k] = v;
        }
        const double tr = block_sum(dx, red), tr2 = block_sum(dx2, red);
        if (it == steps) last = fabs(tr - tr2);
        const bool hold = it < steps && fabs(tr - tr2) < tol;
        const bool square = tr > ne;
        __syncthreads();
        if (!hold)
#pragma unroll
            for (int k = 0; k < PUR_PER_THREAD; ++k) {
                const int e = t + k * PUR_THREADS;
                if (e < n2) X[e] = square ? x2[k] : 2.0 * X[e] - x2[k];
            }
        __syncthreads();
    }
    double* Pm = P + (size_t)mat * n2;
    for (int e = t; e < n2; e += PUR_THREADS) Pm[e] = 0.5 * (X[e] + X[(e % ld) * ld + e / ld]);
    if (t == 0) err[mat] = last;
}

// Pulay coefficients of nrep DIIS problems, thread per replica: the m x m system
// G c = r (G the Gram matrix of the error differences, r the right-hand side;
// rows of unused slots are identity rows with r = 0) by Gaussian elimination with
// partial pivoting; a singular pivot (below 1e-14 of the largest diagonal) leaves
// its unknown 0, as the reference's least-squares solve drops null directions.
__global__ void k_batch_diis_solve(int nrep, int m, const double* __restrict__ G, const double* __restrict__ r,
                                   double* __restrict__ c) {
    const int rep = blockIdx.x * blockDim.x + threadIdx.x;
    if (rep >= nrep) return;
    constexpr int MM = 12;
    double A[MM][MM + 1];
    const double* g = G + (size_t)rep * m * m;
    double scale = 0.0;
    for (int i = 0; i < m; ++i) {
        for (int j = 0; j < m; ++j) A[i][j] = g[i * m + j];
        A[i][m] = r[(size_t)rep * m + i];
        scale = fmax(scale, fabs(A[i][i]));
    }
    int piv_of[MM];
    for (int col = 0; col < m; ++col) {
        int best = col;
        for (int i = col + 1; i < m; ++i)
            if (fabs(A[i][col]) > fabs(A[best][col])) best = i;
        if (best != col)
            for (int j = 0; j <= m; ++j) {
                const double tmp = A[col][j];
                A[col][j] = A[best][j];
                A[best][j] = tmp;
            }
        piv_of[col] = fabs(A[col][col]) > 1e-14 * scale;
        if (!piv_of[col]) continue;
        for (int i = col + 1; i < m; ++i) {
            const double f = A[i][col] / A[col][col];
            for (int j = col; j <= m; ++j) A[i][j] -= f * A[col][j];
        }
    }
    double x[MM];
    for (int i = m - 1; i >= 0; --i) {
        if (!piv_of[i]) {
            x[i] = 0.0;
            continue;
        }
        double v = A[i][m];
        for (int j = i + 1; j < m; ++j) v -= A[i][j] * x[j];
        x[i] = v / A[i][i];
    }
    for (int i = 0; i < m; ++i) c[(size_t)rep * m + i] = x[i];
}

}  // namespace

void launch_batch_purify(int nmat, int ld, const double* F, const double* nocc, double* P, int steps, double tol,
                         double* err, hipStream_t st) {
    if (nmat > 0) k_batch_purify<<<dim3(nmat), dim3(PUR_THREADS), 0, st>>>(ld, F, nocc, P, steps, tol, err);
}

void launch_batch_diis_solve(int nrep, int m, const double* G, const double* r, double* c, hipStream_t st) {
    if (nrep > 0) k_batch_diis_solve<<<dim3((nrep + 63) / 64), dim3(64), 0, st>>>(nrep, m, G, r, c);
}

void launch_batch_scatter(const BatchArgs& b, long long total, hipStream_t st) {
    if (total > 0) k_batch_scatter<<<dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st>>>(b, total);
}

void launch_batch_jk(const BatchArgs& b, int nset, const double* dens, double* J, double* K, long long nitems,
                     hipStream_t st) {
    if (nitems > 0)
        k_batch_jk<<<dim3((unsigned)((nitems + 3) / 4)), dim3(256), 0, st>>>(b, nset, dens, J, K, nitems);
}

}  // namespace aimdhf
