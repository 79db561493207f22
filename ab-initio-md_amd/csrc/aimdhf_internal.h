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
    const int* pair_lo;         // batched plans: first pair of each pair's replica (quartets pair
                                // within a replica only); null: every pair Q <= P
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
    int ncls_host[6];           // pairs per class
    int cls_off_host[6];        // first pair of each class in cls_pairs
    const int* cls_pairs;
    const int* pair_a;          // l_A >= l_B orientation
    const int* pair_b;
    const int* shell_atom;
    const int* shell_l;
    const int* shell_prim;
    const int* shell_ao;
    const double* prim_exp;
    const double* prim_coef;
    const double* pos;
    const double* charges;
    const double* boys;
    int parts;                  // one-electron gradient: 1 overlap/kinetic, 2 nuclear attraction, 3 both
    // matrix layout: element (m, n) of a shell pair's matrices sits at
    // rep * ld^2 + m * ld + n (m, n: AO offsets within the replica, shell_ao);
    // the dipole components are mstride apart. A single molecule: rep 0, ld nao.
    int ld;
    long long mstride;
    const int* shell_rep;       // replica of each shell (null: 0)
    const int* rep_atom;        // nrep + 1: atoms of each replica (null: all atoms)
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
    int npair;                  // shell pairs (s1 >= s2, index s1 (s1 + 1) / 2 + s2)
    const double* dmc;          // per shell pair: max |dc| over its block
    const double* dmx;          // (nx, npair): max |dx| over each block
    double fp32_switch;         // quartets with Q_P Q_Q max|Gamma| below: FP32 derivative integrals
    unsigned char* keep;        // per quartet of the launched class: 0 dropped, 1 FP64, 2 FP32
    int* sel64;                 // selected quartets (indices into the class list), FP64
    int* sel32;                 // and FP32
    int* nsel;                  // their counts (2 ints)
    int* bcnt;                  // pre-pass block counts (2 ints per 256 quartets)
    int* boff;                  // and their exclusive prefix sums
    long long* out;             // (natm, 3) fixed point, hi/lo interleaved
    // batched plans (see OneElArgs): dc of replica r at r ld^2, dx set x at (r nx + x) ld^2;
    // shell pair (s, t) of a replica at shell_pair0[s] + tri(s - shell_first[s], t - shell_first[t])
    int ld;
    const int* shell_rep;       // null: one molecule
    const int* shell_first;
    const int* shell_pair0;
};

// owner-computes J/K (J-engine + exchange owners)
struct JEArgs {
    int npair;
    double thr;
    const int* cls_pairs;
    const int* cls_pair_off;
    int ncls_host[6];           // bra pairs per class (launch sizes)
    const double* qscr;
    const double* ket_q;        // sorted kets (q descending within each class)
    int ket_off[7];             // class ranges of the sorted kets
    const long long* kpre;      // primitive-stream start of every sorted ket (+1)
    const long long* krec;      // record offset of every stream entry
    const double* dh;           // Hermite densities of the stream entries
    long long dh_base[6];       // start of each class's Hermite densities
    const double* recs;
    const long long* rec_off;
    const int* npp;
    const double* boys;
    double* jpart;              // (6, npair, nset, 36)
};

struct DhArgs {
    int nao;
    int nset;
    long long pbase[6];         // first stream entry of each class
    long long dh_base[6];
    const long long* krec;
    const int* kpair;
    const int* pair_a;
    const int* pair_b;
    const int* shell_ao;
    const double* recs;
    const double* dens;
    double* dh;
};

struct JAsmArgs {
    int npair;
    int nao;
    int nset;
    int cq_valid[6];
    const int* pair_a;
    const int* pair_b;
    const int* shell_ao;        // nshell + 1
    const double* jpart;
    double* J;
};

struct KOArgs {
    int nao;
    int nset;
    double thr;
    int out_off[10];            // output blocks per (l_X, l_Y) type
    const int* out_x;
    const int* out_y;
    const int* part_off;        // (nshell * 3 + 1)
    const int* part_shell;
    const int* part_pair;
    const double* part_q;
    const int* pair_a;
    const long long* rec_off;
    const double* recs;
    const int* npp;
    const int* shell_ao;
    const double* boys;
    const double* dens;
    double* kpart;              // (9, nset, nao, nao)
};

// integral store (per geometry) and its gather lists
struct StoreArgs {
    const double* recs;
    const long long* rec_off;
    const int* npp;
    const double* boys;
    double* store;              // this class's blocks
};

// device tables of the store's gather segments (see store_tables.hip)
struct TabArgs {
    int nout, npair, nbig, unit_items;
    int out_off[10];
    int cls_pair_off[7];
    const int* out_x;
    const int* out_y;
    const int* shell_l;
    const int* pair_cls;
    const int* cls_pairs;
    const long long* kband;     // (nout * 9, NBAND) entries per segment band
    const long long* jband;     // (npair * 6, NBAND)
    long long* kcount;          // per K segment: entries, FP64 entries, tiled sizes
    long long* kbig;
    long long* ksz64;
    long long* ksz32;
    long long* knu;             // units per (segment, band)
    long long* jcount;          // per J segment (pair order)
    long long* jbig;
    long long* jsz64;
    long long* jsz32;
    long long* jnu;             // units per (class-order position, cq, band)
    long long* kseg;            // exclusive scans, n + 1 entries
    long long* kdoff;
    long long* kdoff32;
    long long* kustart;
    long long* jseg;
    long long* jdoff;
    long long* jdoff32;
    long long* justart;
    int* kunit_os;              // unit tables
    long long* kunit_e;
    int* out_unit;
    int* junit_pc;
    long long* junit_e;
    int* pair_unit;
};

// pair-pair map of the screened list: qmap[hi (hi + 1) / 2 + lo] = index of
// the quartet (hi, lo) within its class list
struct QmapArgs {
    const int2* list;           // class-binned quartet lists, class k from qoff[k]
    long long qoff[22];
    int* qmap;
};

// Schwarz-product bands of the store's gather segments: band b holds the
// entries with q q in [band_lo[b], band_lo[b - 1]) (band 0 unbounded above,
// band_lo[NBAND - 1] = the screening threshold), in band order, so density-
// difference builds skip whole units of the low bands
constexpr int NBAND = 4;

struct ListArgs {
    int npair;
    double thr;
    double band_lo[NBAND];      // descending lower bounds, the last one thr
    const int* out_x;
    const int* out_y;
    const int* shell_l;
    const int* shell_ao;
    const int* pair_a;
    const int* pair_b;
    const int* pair_cls;
    const int* part_off;
    const int* part_shell;
    const int* part_pair;
    const double* part_q;
    const double* qscr;
    const int* ket_pair;
    const double* ket_q;
    int ket_off[7];
    const int* qmap;
    long long qcount[21];       // quartets per class (bounds of the map values)
    int* err;                   // set to 1 if a map value is out of its class range
    const long long* kseg;      // nout * 9 + 1
    const long long* kband;     // nout * 9 * NBAND entries per segment and band
    long long* klist;
    const long long* jseg;      // npair * 6 + 1
    long long* jlist;
};

struct KSArgs {
    int nao;
    int nset;
    int out_off[10];
    const int* out_x;
    const int* out_y;
    const int* shell_ao;
    const long long* kseg;
    const long long* klist;
    const double* store;
    long long store_base[21];
    double* kdata;              // gather-ordered, reoriented copies (per output segment), FP64 part
    const long long* kdoff;     // nout * 9 start of each segment's FP64 copies
    float* kdata32;             // FP32 part (entries below the switch)
    const long long* kdoff32;   // nout * 9 start of each segment's FP32 copies
    const long long* kbig;      // FP64 entries per segment (they come first)
    int unit_base[10];          // work units per output type
    const int* unit_os;         // per unit: output << 5 | fp32 << 4 | subclass
    const long long* unit_e;    // per unit: entry range [e0, e1)
    const int* out_unit;        // per output: unit range (nout + 1)
    double* kpart;              // per unit partial block (nset NX NY doubles, per type region)
    long long part_base[9];     // start of each type's partial region
    const double* dens;         // (nset, nao, nao), symmetric
    double* K;
    // density-difference builds (incremental Fock): a unit whose largest
    // Schwarz product times max |dens| is below dthr contributes nothing;
    // with kdata32b set and max |dens| <= d32, FP64-part units read that FP32
    // copy (same offsets)
    int delta;
    const double* unit_q;       // per unit: largest q_P q_Q of its entries
    const double* dmax;         // device scalar max |dens|
    double dthr;
    double d32;
    float* kdata32b;            // FP32 copy of the FP64 part (mixed precision), or null
};

struct JSArgs {
    int nao;
    int nset;
    int cls_off;                // first pair of the launched class in cls_pairs
    const int* cls_pairs;
    const int* pair_a;
    const int* pair_b;
    const int* shell_ao;
    const long long* jseg;
    const long long* jlist;
    const double* store;
    long long store_base[21];
    double* jdata;              // gather-ordered, oriented copies (per pair segment), FP64 part
    const long long* jdoff;     // npair * 6 start of each segment's FP64 copies
    float* jdata32;             // FP32 part
    const long long* jdoff32;   // npair * 6 start of each segment's FP32 copies
    const long long* jbig;      // FP64 entries per segment (they come first)
    int unit_base[7];           // work units per bra pair class
    const int* unit_pc;         // per unit: pair << 4 | fp32 << 3 | partner class
    const long long* unit_e;    // per unit: entry range [e0, e1)
    const int* pair_unit;       // per cls_pairs position: unit range (npair + 1)
    double* jpart;              // per unit partial block (nset NAB doubles, per class region)
    long long part_base[6];     // start of each class's partial region
    const double* dens;
    double* J;
    // density-difference builds, as in KSArgs
    int delta;
    const double* unit_q;
    const double* dmax;
    double dthr;
    double d32;
    float* jdata32b;            // FP32 copy of the FP64 part (mixed precision), or null
};

// per-unit largest Schwarz products of the store's work units
struct UnitQArgs {
    const int2* list;           // class-binned quartet lists, class k from qoff[k]
    long long qoff[21];
    const double* qscr;         // per pair Schwarz value
    const int* shell_l;
    const int* out_x;
    const int* out_y;
    const int* pair_cls;
    const long long* klist;
    const int* kunit_os;
    const long long* kunit_e;
    int nkunit;
    double* kunit_q;
    const long long* jlist;
    const int* junit_pc;
    const long long* junit_e;
    int njunit;
    double* junit_q;
};

// batched plans: stored blocks -> dense per-replica tensors, and their J/K digestion
constexpr int NQCLASS_B = 21;
struct BatchArgs {
    int nrep;
    int ld;                     // padded matrix dimension (largest replica)
    const int* rep_nao;         // AOs of each replica
    const long long* rep_toff;  // first tensor element of each replica
    const long long* rep_item;  // nrep + 1: prefix sums of rep_nao^2 (J/K work items)
    const int* shell_rep;
    const int* shell_ao;        // AO offset of each shell within its replica
    const int* shell_l;
    const int* pair_a;
    const int* pair_b;
    const int2* list;           // class-binned quartet list
    long long qoff[NQCLASS_B + 1];
    long long store_base[NQCLASS_B + 1];
    int blk[NQCLASS_B];         // integrals per block of each class
    int ncd[NQCLASS_B];         // ket components of each class
    const double* store;
    double* tens;
};
void launch_batch_scatter(const BatchArgs& b, long long total, hipStream_t st);
void launch_batch_jk(const BatchArgs& b, int nset, const double* dens, double* J, double* K, long long nitems,
                     int lanes, hipStream_t st);
void launch_batch_purify(int nmat, int ld, const double* F, const double* nocc, double* P, int steps, double tol,
                         double* err, hipStream_t st);
void launch_batch_diis_solve(int nrep, int m, const double* G, const double* r, double* c, hipStream_t st);

void launch_eri_store(int k, const StoreArgs& a, const int2* list, long long nq, long long n64, hipStream_t st);
long long store_block(int k);
void launch_klist_count(const ListArgs& a, int nout, long long* counts, long long* band, hipStream_t st);
void launch_klist_fill(const ListArgs& a, int nout, hipStream_t st);
void launch_jlist_count(const ListArgs& a, long long* counts, long long* band, hipStream_t st);
void launch_jlist_fill(const ListArgs& a, hipStream_t st);
void launch_kstore(int type, const KSArgs& a, int nout, hipStream_t st);
void launch_jstore(int pclass, const JSArgs& a, int npairs, hipStream_t st);
void launch_kreduce(int type, const KSArgs& a, int nout, hipStream_t st);
void launch_jreduce(int pclass, const JSArgs& a, int npairs, hipStream_t st);
void launch_kcopy(int type, const KSArgs& a, int nout, hipStream_t st);
void launch_jcopy(int pclass, const JSArgs& a, int npairs, hipStream_t st);
void launch_fill_int(int* p, size_t n, int v, hipStream_t st);
void launch_unit_q(const UnitQArgs& a, hipStream_t st);
void launch_qmap(const QmapArgs& a, hipStream_t st);
void launch_scan(const long long* in, long long* out, long long n, long long* bsum, hipStream_t st);
long long scan_max_items();
void launch_seg_sizes(const TabArgs& a, hipStream_t st);
void launch_units(const TabArgs& a, hipStream_t st);
void launch_tab_totals(const TabArgs& a, long long* out, hipStream_t st);
void launch_absmax(const double* x, size_t n, double* out, hipStream_t st);

void launch_tc2_step(int n, const double* X, const double* X2, double* Xout, double nocc, double freeze,
                     hipStream_t st);
void launch_tc2_init(int n, const double* F, double* X0, hipStream_t st);
void launch_tc2_error(int n, const double* X, double* out, hipStream_t st);
void launch_dh(const DhArgs& a, const long long* n_per_class, hipStream_t st);
void launch_jeng(int pclass, const JEArgs& a, int nset, const int* nket_host, hipStream_t st);
void launch_j_assemble(const JAsmArgs& a, hipStream_t st);
void launch_kown(int type, const KOArgs& a, int nout, const int* has_l, hipStream_t st);
void launch_k_assemble(const double* kpart, int nset, int n, const int* ao_shell, double* K, hipStream_t st);
void launch_zero_d(double* p, size_t n, hipStream_t st);

void launch_build_pairs(const PairArgs& g, hipStream_t st);
void launch_schwarz(const PairArgs& g, const int* ncls_host, hipStream_t st);
// quartet lists: per class, NCOSTBIN bins of primitive-quartet cost (heaviest first)
constexpr int NCOSTBIN = 5;
// list bins of a class: the quartets with q_P q_Q >= sw (FP64) in cost bins,
// then those below sw (FP32 in mixed precision) in cost bins
constexpr int NLBIN = 2 * NCOSTBIN;
void launch_count_quartets(const PairArgs& g, double thr, double sw, int* counts, long long* offs, long long* totals,
                           long long* base, hipStream_t st);
int list_bins(double sw);
void launch_fill_quartets(const PairArgs& g, double thr, double sw, const long long* offs, const long long* base, int2* list,
                          hipStream_t st);
// ---------------------------------------------------------------- runtime-l path (l <= GEN_LMAX)
constexpr int GEN_LMAX = 4;
struct GenArgs {
    int nao, natm, nset;
    const int* shell_atom;
    const int* shell_l;
    const int* shell_prim;
    const int* shell_ao;
    const double* prim_exp;
    const double* prim_coef;
    const double* pos;
    const double* charges;
    const double* boys;
    const int* pair_i;          // shell pairs P: pair_i[P] >= pair_j[P]
    const int* pair_j;
    const int2* quartets;       // unique (P, Q), P >= Q
    long long nq;
    const double* dens;         // J/K: nset symmetric densities
    long long* fx;              // fixed-point accumulators (J/K planes, or 3 natm gradient words)
    const double* dc;           // derivative ERIs: Coulomb density, nx exchange densities, factor
    const double* dx;
    int nx;
    double kfac;
    double *S, *T, *V, *M;      // one-electron outputs
    const double* D;            // one-electron gradient densities
    const double* W;
    int parts;                  // one-electron gradient terms: 1 kinetic, 2 nuclear attraction, 4 overlap
    double* q;                  // Schwarz bounds per shell pair
};
void launch_gen_schwarz(const GenArgs& g, int npair, hipStream_t st);
void launch_gen_jk(const GenArgs& g, hipStream_t st);
void launch_gen_grad2(const GenArgs& g, hipStream_t st);
void launch_gen_oneel(const GenArgs& g, int npair, hipStream_t st);
void launch_gen_grad1(const GenArgs& g, int npair, hipStream_t st);

void launch_zero64(long long* p, size_t n, hipStream_t st);
void launch_symmetrize(const double* D, int nset, int n, double* out, hipStream_t st);
void launch_grad_one_class(int k, const GradArgs& a, const int2* list, long long nq, hipStream_t st);
void launch_grad_keep(int k, const GradArgs& a, const int2* list, long long nq, hipStream_t st);
void launch_pair_dmax(const GradArgs& a, const int* shell_l, double* dmc, double* dmx, hipStream_t st);
void launch_oneel(const OneElArgs& g, double* S, double* T, double* M, double* V, hipStream_t st);
// device address of the fixed-size scratch array of the allocation-free entry
// points on the current device (resolved once per device), and its size
void* noalloc_scratch(size_t* bytes);
void launch_grad_one(const OneElArgs& g, const double* D, const double* Wneg, long long* out, hipStream_t st);
void launch_negate(const double* W, size_t n, double* out, hipStream_t st);
void launch_fx_to_double(const long long* fx, int n, double* out, hipStream_t st);
void launch_grad(const GradArgs& a, const int2* const* lists, const long long* counts, hipStream_t st);
double device_absmax(const double* x, size_t n, double* scratch, hipStream_t st);
void launch_jk_finalize(const long long* fx, int nset, int n, double* J, double* K, hipStream_t st);

}  // namespace aimdhf
