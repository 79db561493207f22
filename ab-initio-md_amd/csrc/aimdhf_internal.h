// Argument blocks shared by the kernels and the host-side plan code.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>

namespace aimdhf {

struct PairArgs {
    int npair;                  // shell pairs (s1 >= s2), index s1 (s1+1)/2 + s2
    int nbpair;                 // contraction-block pairs
    const int* pair_a;          // shell with the higher l (ties: the higher index)
    const int* pair_b;
    const int* pair_cls;        // pair class (ss, ps, pp, ds, dp, dd) = 0..5
    const int* shell_l;
    const int* shell_atom;
    const int* shell_prim;      // nshell + 1
    const double* prim_exp;
    const double* prim_coef;
    const double* pos;          // (natm, 3) bohr
    const long long* rec_off;   // first record of each pair, in doubles
    const int* rec_size;        // doubles per record of each pair
    double* recs;
    int* npp;                   // kept primitive pairs per pair
    const int* cls_pairs;       // pairs grouped by class
    const int* cls_pair_off;    // 7 offsets into cls_pairs
    const int* ncls_pairs;      // 6 counts
    double* qshell;             // Schwarz bound of each shell pair
    double* qblock;             // of each block pair
    double* qscr;               // screening value of each shell pair (its block pair's bound)
    const int* bp_off;          // nbpair + 1
    const int* bp_list;         // shell pairs of each block pair
    const double* boys;         // 402 x 43 table
};

struct JKArgs {
    int nao;
    int nset;
    const int* pair_a;
    const int* pair_b;
    const int* shell_ao;
    const int* npp;
    const long long* rec_off;
    const double* recs;
    const double* boys;
    const double* dens;         // (nset, nao, nao), symmetric
    long long* fx;              // (nset, 4, nao, nao): J hi, J lo, K hi, K lo
};

struct OneElArgs {
    int npair;                  // shell pairs s1 >= s2
    int nao;
    int natm;
    const int* shell_atom;
    const int* shell_l;
    const int* shell_prim;
    const int* shell_ao;
    const double* prim_exp;
    const double* prim_coef;
    const double* pos;
    const double* charges;
    const double* boys;
};

struct GradArgs {
    int nao;
    int natm;
    int nx;
    double kfac;
    double thr;
    const int* pair_a;
    const int* pair_b;
    const int* shell_ao;
    const int* shell_atom;
    const int* npp;
    const long long* rec_off;
    const double* recs;
    const double* boys;
    const double* pos;
    const double* qscr;
    const double* dc;           // (nao, nao), symmetric
    const double* dx;           // (nx, nao, nao), symmetric
    long long* out;             // (natm, 3) fixed point, hi/lo interleaved
};

void launch_build_pairs(const PairArgs& g, hipStream_t st);
void launch_schwarz(const PairArgs& g, const int* ncls_host, hipStream_t st);
void launch_count_quartets(const PairArgs& g, double thr, int* counts, hipStream_t st);
void launch_fill_quartets(const PairArgs& g, double thr, const long long* offsets, int2* list, hipStream_t st);
void launch_zero64(long long* p, size_t n, hipStream_t st);
void launch_symmetrize(const double* D, int nset, int n, double* out, hipStream_t st);
void launch_jk_one(int k, const JKArgs& a, const int2* list, long long nq, hipStream_t st);
void launch_grad_one_class(int k, const GradArgs& a, const int2* list, long long nq, hipStream_t st);
void launch_jk(const JKArgs& a, const int2* const* lists, const long long* counts, hipStream_t st);
void launch_oneel(const OneElArgs& g, double* S, double* T, double* M, double* V, hipStream_t st);
void launch_grad_one(const OneElArgs& g, const double* D, const double* Wneg, long long* out, hipStream_t st);
void launch_negate(const double* W, size_t n, double* out, hipStream_t st);
void launch_fx_to_double(const long long* fx, int n, double* out, hipStream_t st);
void launch_grad(const GradArgs& a, const int2* const* lists, const long long* counts, hipStream_t st);
double device_absmax(const double* x, size_t n, double* scratch, hipStream_t st);
void launch_jk_finalize(const long long* fx, int nset, int n, double* J, double* K, hipStream_t st);

}  // namespace aimdhf
