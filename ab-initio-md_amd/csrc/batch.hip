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

// G lanes (64, 32 or 16) per output element (r, m, c), lanes over (l, s):
//   J_x[m][c] = sum_{l,s} (m c|l s) D_x[l][s]     (the contiguous slab T[m][c])
//   K_x[m][c] = sum_{l,s} (m l|c s) D_x[l][s]     (rows T[m][l][c][.])
// both against the same density element; fixed per-lane order and a butterfly
// sum within the G lanes: bitwise reproducible for a given G
template <int G>
__global__ __launch_bounds__(256) void k_batch_jk(BatchArgs b, int nset, const double* __restrict__ dens,
                                                  double* __restrict__ J, double* __restrict__ K, long long nitems) {
    const long long item = ((long long)blockIdx.x * blockDim.x + threadIdx.x) / G;
    const int lane = threadIdx.x % G;
    if (item >= nitems) return;                 // whole groups of G lanes leave together
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
    const double* krow = T + (size_t)m * N * N * N + (size_t)c * N;   // + l N^2 + s
    for (int e = lane; e < N * N; e += G) {
        const int l = e / N, s = e % N;
        const size_t de = (size_t)l * ld + s;
        const double tj = slab[e], tk = krow[(size_t)l * N * N + s];
        const double d0 = D0[de];
        ja[0] += tj * d0;
        ka[0] += tk * d0;
        if (nset == 2) {
            const double d1 = D0[ld2 + de];
            ja[1] += tj * d1;
            ka[1] += tk * d1;
        }
    }
    for (int x = 0; x < nset; ++x) {
        double jv = ja[x], kv = ka[x];
#pragma unroll
        for (int o = G / 2; o >= 1; o >>= 1) {
            jv += __shfl_xor(jv, o, 64);
            kv += __shfl_xor(kv, o, 64);
        }
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
//
// X^2 on the matrix cores: X sits zero-padded to 64 x 64 in LDS (row stride 65
// doubles, so the 16 rows one operand load touches fall in distinct banks);
// wavefront w forms the 16 x 64 row block w of X^2 as four 16 x 16 tiles, each
// 16 steps of v_mfma_f64_16x16x4f64 over k. Operand and result lanes (measured on
// gfx950): A(i, k) from lane i + 16 k, B(k, j) from lane j + 16 k, D(i, j) in
// register i / 4 of lane j + 16 (i % 4). Each result element has one owner lane,
// which applies the update.
constexpr int PUR_THREADS = 256;
constexpr int PUR_MAXLD = 64;
constexpr int PUR_LS = PUR_MAXLD + 1;       // LDS row stride (doubles)

typedef double v4d __attribute__((ext_vector_type(4)));

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
    __shared__ double X[PUR_MAXLD * PUR_LS];
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
    for (int e = t; e < PUR_MAXLD * PUR_LS; e += PUR_THREADS) {
        const int i = e / PUR_LS, j = e % PUR_LS;
        X[e] = (i < ld && j < ld) ? ((i == j ? emax : 0.0) - Fm[i * ld + j]) * inv : 0.0;
    }
    __syncthreads();
    const double ne = nocc[mat];
    const int w = t >> 6, lane = t & 63, li = lane & 15, lk = lane >> 4;
    const int row0 = 16 * w + lk;                   // rows row0 + 4 v (v < 4) of the lane's results
    double last = 0.0;
    for (int it = 0; it < steps + 2; ++it) {
        v4d acc[4];
#pragma unroll
        for (int tj = 0; tj < 4; ++tj) acc[tj] = v4d{0.0, 0.0, 0.0, 0.0};
#pragma unroll 4
        for (int k0 = 0; k0 < PUR_MAXLD; k0 += 4) {
            const double a = X[(16 * w + li) * PUR_LS + k0 + lk];
#pragma unroll
            for (int tj = 0; tj < 4; ++tj) {
                const double b = X[(k0 + lk) * PUR_LS + 16 * tj + li];
                acc[tj] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[tj], 0, 0, 0);
            }
        }
        // traces: the diagonal of X, and of X^2 in the wavefront's diagonal tile
        double dx = 0.0, dx2 = 0.0;
        if (t < PUR_MAXLD) dx = X[t * PUR_LS + t];
#pragma unroll
        for (int v = 0; v < 4; ++v)
            if (row0 + 4 * v == 16 * w + li) dx2 += acc[w][v];
        const double tr = block_sum(dx, red), tr2 = block_sum(dx2, red);
        if (it == steps) last = fabs(tr - tr2);
        const bool hold = it < steps && fabs(tr - tr2) < tol;
        const bool square = tr > ne;
        __syncthreads();
        if (!hold)
#pragma unroll
            for (int tj = 0; tj < 4; ++tj)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    double& x = X[(row0 + 4 * v) * PUR_LS + 16 * tj + li];
                    x = square ? acc[tj][v] : 2.0 * x - acc[tj][v];
                }
        __syncthreads();
    }
    double* Pm = P + (size_t)mat * n2;
    for (int e = t; e < n2; e += PUR_THREADS) {
        const int i = e / ld, j = e % ld;
        Pm[e] = 0.5 * (X[i * PUR_LS + j] + X[j * PUR_LS + i]);
    }
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
                     int lanes, hipStream_t st) {
    if (nitems <= 0) return;
    const long long per = 256 / lanes;          // output elements per workgroup
    const dim3 grid((unsigned)((nitems + per - 1) / per));
    if (lanes == 16)
        k_batch_jk<16><<<grid, dim3(256), 0, st>>>(b, nset, dens, J, K, nitems);
    else if (lanes == 32)
        k_batch_jk<32><<<grid, dim3(256), 0, st>>>(b, nset, dens, J, K, nitems);
    else
        k_batch_jk<64><<<grid, dim3(256), 0, st>>>(b, nset, dens, J, K, nitems);
}

}  // namespace aimdhf
