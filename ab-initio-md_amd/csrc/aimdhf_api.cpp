// Host side of the C ABI (include/aimdhf.h): plan state, device buffers,
// launch sequencing. Compiled with hipcc together with the kernels.

#include "../../include/aimdhf.h"
#include "aimdhf_internal.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

struct HipError {
    int code;
};

#define HIPCHECK(x)                                                                  \
    do {                                                                             \
        hipError_t e_ = (x);                                                         \
        if (e_ != hipSuccess) {                                                      \
            g_err = std::string(#x) + ": " + hipGetErrorString(e_);                  \
            throw HipError{e_ == hipErrorOutOfMemory ? AIMDHF_E_NOMEM : AIMDHF_E_HIP}; \
        }                                                                            \
    } while (0)

// Device buffer that grows on demand and is freed with the plan.
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    // grows with 1/8 slack: per-geometry sizes drift along a trajectory, and each
    // reallocation (hipFree synchronises the device) would stall the pipeline
    void ensure(size_t n) {
        if (n <= bytes) return;
        if (p) HIPCHECK(hipFree(p));
        p = nullptr;
        bytes = 0;
        const size_t want = std::max<size_t>(n + n / 8, 256);
        HIPCHECK(hipMalloc(&p, want));
        bytes = want;
    }
    void release() {
        if (p) hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    template <class T>
    T* as() const { return (T*)p; }
};

template <class T>
void upload(DevBuf& b, const std::vector<T>& v) {
    b.ensure(sizeof(T) * std::max<size_t>(v.size(), 1));
    if (!v.empty()) HIPCHECK(hipMemcpy(b.p, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice));
}

int ncart(int l) { return (l + 1) * (l + 2) / 2; }
int nherm(int L) { return (L + 1) * (L + 2) * (L + 3) / 6; }
constexpr int NQCLASS = 21;
constexpr int REC_HDR = 8;

// Boys grid table F_n(k/10), n < 43 (aimd/qc/boys.py:57-75): series for the
// highest order, then the downward recursion.
std::vector<double> boys_table() {
    const int NG = 402, NR = 43, top = NR - 1;
    std::vector<double> tab((size_t)NG * NR);
    std::vector<double> col(NR);
    for (int k = 0; k < NG; ++k) {
        double x = k * 0.1;
        double term = 1.0 / (2 * top + 1), total = term;
        for (int m = 1; m < 400; ++m) {
            term = term * (2.0 * x) / (double)(2 * top + 2 * m + 1);
            total += term;
        }
        double ex = std::exp(-x);
        col[top] = ex * total;
        for (int n = top - 1; n >= 0; --n) col[n] = (2.0 * x * col[n + 1] + ex) / (double)(2 * n + 1);
        for (int n = 0; n < NR; ++n) tab[(size_t)k * NR + n] = col[n];
    }
    return tab;
}

// Optional per-launch timing with HIP events on the caller's stream.
// Kinds as listed at aimdhf_profile_enable in aimdhf.h (0..63).
struct Profiler {
    static constexpr int NKIND = 98;
    bool on = false;
    struct Rec {
        int kind;
        hipEvent_t a, b;
    };
    std::vector<hipEvent_t> pool;
    std::vector<Rec> recs;
    double ms[NKIND] = {0};
    long long n[NKIND] = {0};
    hipEvent_t get() {
        if (!pool.empty()) {
            hipEvent_t e = pool.back();
            pool.pop_back();
            return e;
        }
        hipEvent_t e;
        HIPCHECK(hipEventCreate(&e));
        return e;
    }
    template <class F>
    void time(int kind, hipStream_t st, F&& launch) {
        if (!on) {
            launch();
            return;
        }
        Rec r{kind, get(), get()};
        HIPCHECK(hipEventRecord(r.a, st));
        launch();
        HIPCHECK(hipEventRecord(r.b, st));
        recs.push_back(r);
    }
    void collect() {
        for (Rec& r : recs) {
            HIPCHECK(hipEventSynchronize(r.b));
            float t = 0.0f;
            HIPCHECK(hipEventElapsedTime(&t, r.a, r.b));
            ms[r.kind] += t;
            n[r.kind] += 1;
            pool.push_back(r.a);
            pool.push_back(r.b);
        }
        recs.clear();
    }
    // host wall-clock phase (milliseconds since t0), recorded when enabled
    using clock = std::chrono::steady_clock;
    void host(int kind, clock::time_point t0) {
        if (!on) return;
        ms[kind] += std::chrono::duration<double, std::milli>(clock::now() - t0).count();
        n[kind] += 1;
    }
    void reset() {
        collect();
        for (int i = 0; i < NKIND; ++i) { ms[i] = 0.0; n[i] = 0; }
    }
    ~Profiler() {
        for (Rec& r : recs) { (void)hipEventDestroy(r.a); (void)hipEventDestroy(r.b); }
        for (hipEvent_t e : pool) (void)hipEventDestroy(e);
    }
};

// Fan-out of independent launches over the caller's stream and NAUX extra
// streams (default priority, created on first use): fork() makes the extra
// streams wait for the caller's work so far, on(i) picks a stream round-robin,
// join() makes the caller's stream wait for everything launched on them.
struct Fan {
    static constexpr int NAUX = 3;
    hipStream_t aux[NAUX] = {};
    hipEvent_t fork_ev = nullptr, join_ev[NAUX] = {};
    hipStream_t base = nullptr;
    bool serial = false;        // every launch on the caller's stream
    void init() {
        if (fork_ev) return;
        for (int i = 0; i < NAUX; ++i) {
            HIPCHECK(hipStreamCreateWithFlags(&aux[i], hipStreamNonBlocking));
            HIPCHECK(hipEventCreateWithFlags(&join_ev[i], hipEventDisableTiming));
        }
        HIPCHECK(hipEventCreateWithFlags(&fork_ev, hipEventDisableTiming));
    }
    void fork(hipStream_t st) {
        base = st;
        if (serial) return;
        init();
        HIPCHECK(hipEventRecord(fork_ev, st));
        for (int i = 0; i < NAUX; ++i) HIPCHECK(hipStreamWaitEvent(aux[i], fork_ev, 0));
    }
    hipStream_t on(int i) const { return serial || i % (NAUX + 1) == 0 ? base : aux[i % (NAUX + 1) - 1]; }
    void join() {
        if (serial) return;
        for (int i = 0; i < NAUX; ++i) {
            HIPCHECK(hipEventRecord(join_ev[i], aux[i]));
            HIPCHECK(hipStreamWaitEvent(base, join_ev[i], 0));
        }
    }
    ~Fan() {
        for (int i = 0; i < NAUX; ++i) {
            if (aux[i]) (void)hipStreamDestroy(aux[i]);
            if (join_ev[i]) (void)hipEventDestroy(join_ev[i]);
        }
        if (fork_ev) (void)hipEventDestroy(fork_ev);
    }
};

}  // namespace

struct aimdhf_plan {
    Profiler prof;
    Fan fan;
    int natm = 0, nshell = 0, nao = 0, npair = 0, nblock = 0, nbpair = 0, lmax = 0;
    std::vector<int> shell_atom, shell_l, shell_prim, shell_ao, shell_block;
    std::vector<double> prim_exp, prim_coef, charges, pos;
    std::vector<int> pair_a, pair_b, pair_cls, rec_size, cls_pairs, cls_pair_off, ncls;
    std::vector<long long> rec_off;
    std::vector<int> bp_off, bp_list;
    long long rec_total = 0;
    bool have_geometry = false;
    double threshold = 0.0;
    std::vector<long long> qcount;   // per class
    std::vector<long long> q64count; // per class, quartets with q_P q_Q >= the FP32 switch (listed first)
    std::vector<long long> g64count;
    long long qtotal = 0;

    DevBuf d_shell_atom, d_shell_l, d_shell_prim, d_shell_ao, d_prim_exp, d_prim_coef, d_pos;
    DevBuf d_pair_a, d_pair_b, d_pair_cls, d_rec_size, d_rec_off, d_cls_pairs, d_cls_pair_off, d_ncls;
    DevBuf d_bp_off, d_bp_list, d_boys;
    DevBuf d_recs, d_npp, d_qshell, d_qblock, d_qscr, d_counts, d_offsets, d_list;
    DevBuf d_fx, d_dsym, d_charges, d_glist, d_gfx, d_gden, d_scratch;
    std::vector<long long> gcount;
    // owner-computes J/K: static output lists and per-geometry streams / partner lists
    std::vector<int> out_x, out_y, ao_shell;
    int out_off[10] = {0};
    int has_l[3] = {0, 0, 0};
    DevBuf d_out_x, d_out_y, d_ao_shell;
    std::vector<int> ket_pair, ket_off, part_off, part_shell, part_pair, kpair;
    std::vector<double> ket_q, part_q;
    std::vector<long long> kpre, krec;
    long long pbase[7] = {0}, dh_base2[7] = {0};
    int nket[6] = {0};
    DevBuf d_ket_q, d_kpre, d_krec, d_kpair, d_part_off, d_part_shell, d_part_pair, d_part_q;
    DevBuf d_dh, d_jpart, d_kpart, d_ket_pair;
    // integral store: mode (0 auto, 1 direct, 2 store), byte limit (0: half the free device memory)
    int jk_mode = 0;
    long long store_limit = 0;
    bool stored = false;
    long long store_base[NQCLASS + 1] = {0};
    long long store_bytes = 0, k_entries = 0, j_entries = 0;
    DevBuf d_gkeep, d_gsel, d_nsel, d_gblk;
    // precision: 0 FP64 everywhere; 1 mixed (FP32 below the switches)
    int precision = 0;
    double jk_switch = 0.0, grad_switch = 0.0;
    DevBuf d_store, d_qmap, d_kcnt, d_jcnt, d_kseg, d_jseg, d_klist, d_jlist, d_lerr;
    DevBuf d_kdata, d_jdata, d_kdoff, d_jdoff, d_kdata32, d_jdata32, d_kdoff32, d_jdoff32, d_kbig, d_jbig;
    long long kdata_total = 0, jdata_total = 0, kdata32_total = 0, jdata32_total = 0;
    // work units of the store kernels: at most unit_items lane items each
    int unit_items = 1024;
    int kunit_base[10] = {0}, junit_base[7] = {0};
    DevBuf d_kunit_os, d_kunit_e, d_out_unit, d_kpart2, d_junit_pc, d_junit_e, d_pair_unit, d_jpart2;
    // density-difference builds: per-unit largest Schwarz products, max |dD|,
    // and in mixed precision FP32 copies of the FP64 parts
    DevBuf d_kunit_q, d_junit_q, d_dmax, d_kdata32b, d_jdata32b;
    bool have_32b = false;
    // largest max |dD| at which mixed-precision difference builds digest every
    // unit in FP32 (larger differences use the FP64 parts of the store)
    double delta_fp32_max = 1e-3;
    DevBuf d_tab;                   // store-table temporaries (sizes, unit counts, scan blocks)
    DevBuf d_gdmax;                 // per shell pair density maxima for the gradient screening
    long long* h_pin = nullptr;     // pinned host words for the totals and the error flag
    // pinned host staging of the per-geometry list transfers: asynchronous copies on
    // the caller's stream, which do not wait for the integrals on the library streams
    char* h_stage = nullptr;
    size_t h_stage_cap = 0, h_stage_used = 0;
    // batched plans: independent molecules (replicas) sharing one plan; shells and
    // atoms grouped by replica, pairs and quartets only within a replica, matrices
    // in the padded layout (replica r's ld x ld block at r ld^2), two-electron
    // integrals as one dense n_r^4 tensor per replica
    bool batched = false;
    int nrep = 1, bld = 0;
    int batch_jk_lanes = 64;        // lanes per output element of the batched J/K digestion
    std::vector<int> rep_shell, rep_atom, rep_nao, shell_rep, shell_first, shell_pair0, pair_lo;
    std::vector<long long> rep_toff, rep_item;
    long long tens_total = 0;
    DevBuf d_shell_rep, d_shell_first, d_shell_pair0, d_pair_lo, d_rep_atom, d_rep_nao, d_rep_toff, d_rep_item, d_tens;
    // elements of one stack of matrices (nao^2, or nrep ld^2 for a batch)
    size_t mat_elems() const { return batched ? (size_t)nrep * bld * bld : (size_t)nao * nao; }
    // device address and size of the fixed scratch array of the allocation-free
    // entry point on this plan's device (resolved at plan creation)
    void* noalloc_ptr = nullptr;
    size_t noalloc_bytes = 0;
    // runtime-l path (a shell with l = 3 or 4): every shell pair i >= j in index
    // order; the energy quartets in d_list / qtotal, the last gradient list in
    // d_glist / gen_gq; bounds per shell pair in d_qshell, per block pair in d_qblock
    // arguments of the last store-mode J/K build (kernel replays for timing)
    aimdhf::KSArgs last_ka{};
    aimdhf::JSArgs last_ja{};
    bool have_last_store = false;
    bool generic = false;
    std::vector<int> gpair_i, gpair_j, gpair_bp, gen_npp;
    std::vector<double> gen_qb;
    long long gen_gq = 0;
    DevBuf d_gpair_i, d_gpair_j;

    aimdhf::GenArgs gen_args() const {
        aimdhf::GenArgs g{};
        g.nao = nao;
        g.natm = natm;
        g.nset = 1;
        g.shell_atom = d_shell_atom.as<int>();
        g.shell_l = d_shell_l.as<int>();
        g.shell_prim = d_shell_prim.as<int>();
        g.shell_ao = d_shell_ao.as<int>();
        g.prim_exp = d_prim_exp.as<double>();
        g.prim_coef = d_prim_coef.as<double>();
        g.pos = d_pos.as<double>();
        g.charges = d_charges.as<double>();
        g.boys = d_boys.as<double>();
        g.pair_i = d_gpair_i.as<int>();
        g.pair_j = d_gpair_j.as<int>();
        g.q = d_qshell.as<double>();
        return g;
    }

    // unique shell quartets (P >= Q) whose block-pair bounds give q_P q_Q >= thr, the
    // heaviest first, into dst; returns their number
    long long gen_build_list(double thr, DevBuf& dst) {
        std::vector<int> ord(npair);
        for (int i = 0; i < npair; ++i) ord[i] = i;
        std::stable_sort(ord.begin(), ord.end(),
                         [&](int a, int b) { return gen_qb[gpair_bp[a]] > gen_qb[gpair_bp[b]]; });
        std::vector<int2> out;
        for (int x = 0; x < npair; ++x) {
            const double qa = gen_qb[gpair_bp[ord[x]]];
            for (int y = 0; y < npair; ++y) {
                if (!(qa * gen_qb[gpair_bp[ord[y]]] >= thr)) break;
                const int P = ord[x], Q = ord[y];
                if (P >= Q) out.push_back(make_int2(P, Q));
            }
        }
        auto cost = [&](const int2& q) {
            const int a = gpair_i[q.x], b = gpair_j[q.x], c = gpair_i[q.y], d = gpair_j[q.y];
            const long long np = (long long)gen_npp[q.x] * gen_npp[q.y];
            return np * ncart(shell_l[a]) * ncart(shell_l[b]) * ncart(shell_l[c]) * ncart(shell_l[d]);
        };
        std::stable_sort(out.begin(), out.end(), [&](const int2& a, const int2& b) { return cost(a) > cost(b); });
        upload(dst, out);
        return (long long)out.size();
    }

    // make room for `bytes` of staging (no transfer may be pending) and start over
    void stage_reset(size_t bytes) {
        if (bytes > h_stage_cap) {
            if (h_stage) HIPCHECK(hipHostFree(h_stage));
            h_stage = nullptr;
            h_stage_cap = 0;
            const size_t want = bytes + bytes / 8 + 4096;
            HIPCHECK(hipHostMalloc((void**)&h_stage, want));
            h_stage_cap = want;
        }
        h_stage_used = 0;
    }
    char* stage_take(size_t bytes) {
        const size_t at = h_stage_used;
        h_stage_used += (bytes + 255) & ~(size_t)255;
        if (h_stage_used > h_stage_cap) {
            g_err = "staging buffer overflow";
            throw HipError{AIMDHF_E_STATE};
        }
        return h_stage + at;
    }
    template <class T>
    void upload_async(DevBuf& b, const std::vector<T>& v, hipStream_t st) {
        b.ensure(sizeof(T) * std::max<size_t>(v.size(), 1));
        if (v.empty()) return;
        char* h = stage_take(sizeof(T) * v.size());
        std::memcpy(h, v.data(), sizeof(T) * v.size());
        HIPCHECK(hipMemcpyAsync(b.p, h, sizeof(T) * v.size(), hipMemcpyHostToDevice, st));
    }

    ~aimdhf_plan() {
        for (DevBuf* b : {&d_shell_atom, &d_shell_l, &d_shell_prim, &d_shell_ao, &d_prim_exp, &d_prim_coef,
                          &d_pos, &d_pair_a, &d_pair_b, &d_pair_cls, &d_rec_size, &d_rec_off, &d_cls_pairs,
                          &d_cls_pair_off, &d_ncls, &d_bp_off, &d_bp_list, &d_boys, &d_recs, &d_npp,
                          &d_qshell, &d_qblock, &d_qscr, &d_counts, &d_offsets, &d_list, &d_fx, &d_dsym,
                          &d_charges, &d_glist, &d_gfx, &d_gden, &d_scratch, &d_out_x, &d_out_y, &d_ao_shell,
                          &d_ket_q, &d_kpre, &d_krec, &d_kpair, &d_part_off, &d_part_shell, &d_part_pair,
                          &d_part_q, &d_dh, &d_jpart, &d_kpart, &d_ket_pair, &d_store, &d_qmap, &d_kcnt,
                          &d_jcnt, &d_kseg, &d_jseg, &d_klist, &d_jlist, &d_lerr, &d_gkeep, &d_gsel, &d_nsel, &d_gblk,
                          &d_kdata, &d_jdata, &d_kdoff, &d_jdoff, &d_kdata32, &d_jdata32, &d_kdoff32, &d_jdoff32,
                          &d_kbig, &d_jbig, &d_kunit_os, &d_kunit_e, &d_out_unit,
                          &d_kpart2, &d_junit_pc, &d_junit_e, &d_pair_unit, &d_jpart2, &d_kunit_q, &d_junit_q,
                          &d_dmax, &d_kdata32b, &d_jdata32b, &d_tab, &d_gdmax, &d_shell_rep, &d_shell_first,
                          &d_shell_pair0, &d_pair_lo, &d_rep_atom, &d_rep_nao, &d_rep_toff, &d_rep_item, &d_tens,
                          &d_gpair_i, &d_gpair_j})
            b->release();
        if (h_pin) (void)hipHostFree(h_pin);
        if (h_stage) (void)hipHostFree(h_stage);
    }

    aimdhf::PairArgs pair_args() const {
        aimdhf::PairArgs g;
        g.npair = npair;
        g.nbpair = nbpair;
        g.pair_a = d_pair_a.as<int>();
        g.pair_b = d_pair_b.as<int>();
        g.pair_cls = d_pair_cls.as<int>();
        g.shell_l = d_shell_l.as<int>();
        g.shell_atom = d_shell_atom.as<int>();
        g.shell_prim = d_shell_prim.as<int>();
        g.prim_exp = d_prim_exp.as<double>();
        g.prim_coef = d_prim_coef.as<double>();
        g.pos = d_pos.as<double>();
        g.rec_off = d_rec_off.as<long long>();
        g.rec_size = d_rec_size.as<int>();
        g.recs = d_recs.as<double>();
        g.npp = d_npp.as<int>();
        g.cls_pairs = d_cls_pairs.as<int>();
        g.cls_pair_off = d_cls_pair_off.as<int>();
        g.ncls_pairs = d_ncls.as<int>();
        g.qshell = d_qshell.as<double>();
        g.qblock = d_qblock.as<double>();
        g.qscr = d_qscr.as<double>();
        g.bp_off = d_bp_off.as<int>();
        g.bp_list = d_bp_list.as<int>();
        g.boys = d_boys.as<double>();
        g.pair_lo = batched ? d_pair_lo.as<int>() : nullptr;
        return g;
    }

    const int2* list_of_class(int k) const {
        long long off = 0;
        for (int i = 0; i < k; ++i) off += qcount[i];
        return d_list.as<int2>() + off;
    }

    // Class-binned list of unique shell quartets with q_P q_Q >= thr into buf;
    // counts[k] receives the per-class sizes. Synchronises the stream.
    // Screened quartet list binned by class, precision part (q_P q_Q >= sw
    // first) and cost; counts per class, n64 per class the part >= sw.
    void build_list(double thr, double sw, DevBuf& buf, std::vector<long long>& counts, std::vector<long long>& n64,
                    hipStream_t st) {
        aimdhf::PairArgs g = pair_args();
        const int nb = aimdhf::list_bins(sw), nv = NQCLASS * nb;
        d_counts.ensure(sizeof(int) * (size_t)npair * nv);
        d_offsets.ensure(sizeof(long long) * ((size_t)npair * nv + 2 * nv));
        long long* offs = d_offsets.as<long long>();
        long long* totals = offs + (size_t)npair * nv;
        long long* base = totals + nv;
        aimdhf::launch_count_quartets(g, thr, sw, d_counts.as<int>(), offs, totals, base, st);
        HIPCHECK(hipGetLastError());
        std::vector<long long> tot(nv);
        HIPCHECK(hipMemcpyAsync(tot.data(), totals, sizeof(long long) * nv, hipMemcpyDeviceToHost, st));
        HIPCHECK(hipStreamSynchronize(st));
        counts.assign(NQCLASS, 0);
        n64.assign(NQCLASS, 0);
        long long total = 0;
        for (int k = 0; k < NQCLASS; ++k)
            for (int b = 0; b < nb; ++b) {
                counts[k] += tot[k * nb + b];
                if (b < aimdhf::NCOSTBIN) n64[k] += tot[k * nb + b];
            }
        for (long long c : counts) total += c;
        buf.ensure(sizeof(int2) * std::max<long long>(total, 1));
        aimdhf::launch_fill_quartets(g, thr, sw, offs, base, buf.as<int2>(), st);
        HIPCHECK(hipGetLastError());
    }

    // Sorted ket streams (J-engine) and per-shell partner lists (exchange
    // owners) from the screening values of the current geometry. Synchronous.
    // direct: also the ket primitive records and Hermite-density layout the
    // direct-mode J-engine streams (not needed when the integrals are stored)
    void build_owner_lists(double thr, bool direct, hipStream_t st) {
        HIPCHECK(hipStreamSynchronize(st));
        stage_reset(12 * (size_t)npair + 512);
        double* hq = (double*)stage_take(sizeof(double) * npair);
        int* hn = (int*)stage_take(sizeof(int) * npair);
        HIPCHECK(hipMemcpyAsync(hq, d_qscr.p, sizeof(double) * npair, hipMemcpyDeviceToHost, st));
        HIPCHECK(hipMemcpyAsync(hn, d_npp.p, sizeof(int) * npair, hipMemcpyDeviceToHost, st));
        HIPCHECK(hipStreamSynchronize(st));
        std::vector<double> q(hq, hq + npair);
        std::vector<int> np(hn, hn + npair);
        double qmax = 0.0;
        for (double v : q) qmax = std::max(qmax, v);
        // kets by class, q descending (ties: pair index)
        ket_pair.clear();
        ket_q.clear();
        ket_off.assign(7, 0);
        for (int c = 0; c < 6; ++c) {
            std::vector<int> ps(cls_pairs.begin() + cls_pair_off[c], cls_pairs.begin() + cls_pair_off[c + 1]);
            std::stable_sort(ps.begin(), ps.end(), [&](int a, int b) { return q[a] > q[b] || (q[a] == q[b] && a < b); });
            for (int P : ps) {
                ket_pair.push_back(P);
                ket_q.push_back(q[P]);
            }
            ket_off[c + 1] = (int)ket_pair.size();
            nket[c] = ket_off[c + 1] - ket_off[c];
        }
        kpre.assign(ket_pair.size() + 1, 0);
        krec.clear();
        kpair.clear();
        for (size_t m = 0; direct && m < ket_pair.size(); ++m) {
            int P = ket_pair[m];
            kpre[m + 1] = kpre[m] + np[P];
            for (int k = 0; k < np[P]; ++k) {
                krec.push_back(rec_off[P] + (long long)k * rec_size[P]);
                kpair.push_back(P);
            }
        }
        long long db = 0;
        for (int c = 0; c < 6; ++c) {
            int lc = c == 0 ? 0 : c <= 2 ? 1 : 2, ld = c - lc * (lc + 1) / 2;
            pbase[c] = kpre[ket_off[c]];
            dh_base2[c] = db;
            db += (kpre[ket_off[c + 1]] - kpre[ket_off[c]]) * 2 * nherm(lc + ld);
        }
        pbase[6] = kpre[ket_off[6]];
        dh_base2[6] = db;
        // partner lists: shell X -> shells J of each l with q_XJ qmax >= thr (the exact
        // test: no quartet of a pair failing it can pass q_P q_Q >= thr), q descending
        part_off.assign(nshell * 3 + 1, 0);
        part_shell.clear();
        part_pair.clear();
        part_q.clear();
        std::vector<std::pair<double, int>> tmp;
        for (int X = 0; X < nshell; ++X)
            for (int l = 0; l < 3; ++l) {
                tmp.clear();
                for (int J = 0; J < nshell; ++J) {
                    if (shell_l[J] != l) continue;
                    int P = X >= J ? X * (X + 1) / 2 + J : J * (J + 1) / 2 + X;
                    if (!(q[P] * qmax < thr)) tmp.push_back({q[P], J});
                }
                std::stable_sort(tmp.begin(), tmp.end(), [](const std::pair<double, int>& a, const std::pair<double, int>& b) {
                    return a.first > b.first || (a.first == b.first && a.second < b.second);
                });
                for (auto& e : tmp) {
                    int J = e.second;
                    part_shell.push_back(J);
                    part_pair.push_back(X >= J ? X * (X + 1) / 2 + J : J * (J + 1) / 2 + X);
                    part_q.push_back(e.first);
                }
                part_off[X * 3 + l + 1] = (int)part_shell.size();
            }
        size_t bytes = 12 * ket_q.size() + 4 * part_off.size() + 16 * part_shell.size() + 8 * 1024;
        if (direct) bytes += 8 * kpre.size() + 12 * krec.size();
        stage_reset(bytes);   // the downloads above have completed
        upload_async(d_ket_q, ket_q, st);
        upload_async(d_ket_pair, ket_pair, st);
        if (direct) {
            upload_async(d_kpre, kpre, st);
            upload_async(d_krec, krec, st);
            upload_async(d_kpair, kpair, st);
            d_dh.ensure(sizeof(double) * std::max<long long>(db, 1));
        }
        upload_async(d_part_off, part_off, st);
        upload_async(d_part_shell, part_shell, st);
        upload_async(d_part_pair, part_pair, st);
        upload_async(d_part_q, part_q, st);
    }

    // Bytes the integral store of the current list would take: the blocks, up
    // to six gather-ordered copies of them (four exchange, two Coulomb), the
    // pair-pair map and the gather lists (about six entries per quartet).
    long long store_estimate() const {
        long long dbl = 0;
        for (int k = 0; k < NQCLASS; ++k) dbl += qcount[k] * aimdhf::store_block(k);
        long long keys = (long long)npair * (npair + 1) / 2;
        return 8 * 7 * dbl + 4 * keys + 8 * 6 * qtotal;
    }

    aimdhf::KSArgs ks_args(int nset, const double* dens, double* K) const {
        aimdhf::KSArgs ka;
        ka.nao = nao;
        ka.nset = nset;
        for (int t = 0; t < 10; ++t) ka.out_off[t] = out_off[t];
        ka.out_x = d_out_x.as<int>();
        ka.out_y = d_out_y.as<int>();
        ka.shell_ao = d_shell_ao.as<int>();
        ka.kseg = d_kseg.as<long long>();
        ka.klist = d_klist.as<long long>();
        ka.store = d_store.as<double>();
        for (int k = 0; k < NQCLASS; ++k) ka.store_base[k] = store_base[k];
        ka.kdata = d_kdata.as<double>();
        ka.kdoff = d_kdoff.as<long long>();
        ka.kdata32 = d_kdata32.as<float>();
        ka.kdoff32 = d_kdoff32.as<long long>();
        ka.kbig = d_kbig.as<long long>();
        for (int t = 0; t < 10; ++t) ka.unit_base[t] = kunit_base[t];
        long long pb = 0;
        for (int t = 0; t < 9; ++t) {
            ka.part_base[t] = pb;
            pb += (long long)(kunit_base[t + 1] - kunit_base[t]) * std::max(nset, 1) * ncart(t / 3) * ncart(t % 3);
        }
        ka.unit_os = d_kunit_os.as<int>();
        ka.unit_e = d_kunit_e.as<long long>();
        ka.out_unit = d_out_unit.as<int>();
        ka.kpart = d_kpart2.as<double>();
        ka.dens = dens;
        ka.K = K;
        ka.delta = 0;
        ka.unit_q = d_kunit_q.as<double>();
        ka.dmax = d_dmax.as<double>();
        ka.dthr = 0.0;
        ka.d32 = 0.0;
        ka.kdata32b = have_32b ? d_kdata32b.as<float>() : nullptr;
        return ka;
    }

    aimdhf::JSArgs js_args(int nset, const double* dens, double* J) const {
        aimdhf::JSArgs ja;
        ja.nao = nao;
        ja.nset = nset;
        ja.cls_off = 0;
        ja.cls_pairs = d_cls_pairs.as<int>();
        ja.pair_a = d_pair_a.as<int>();
        ja.pair_b = d_pair_b.as<int>();
        ja.shell_ao = d_shell_ao.as<int>();
        ja.jseg = d_jseg.as<long long>();
        ja.jlist = d_jlist.as<long long>();
        ja.store = d_store.as<double>();
        for (int k = 0; k < NQCLASS; ++k) ja.store_base[k] = store_base[k];
        ja.jdata = d_jdata.as<double>();
        ja.jdoff = d_jdoff.as<long long>();
        ja.jdata32 = d_jdata32.as<float>();
        ja.jdoff32 = d_jdoff32.as<long long>();
        ja.jbig = d_jbig.as<long long>();
        for (int c = 0; c < 7; ++c) ja.unit_base[c] = junit_base[c];
        static const int ncp[6] = {1, 3, 9, 6, 18, 36};
        long long pb = 0;
        for (int c = 0; c < 6; ++c) {
            ja.part_base[c] = pb;
            pb += (long long)(junit_base[c + 1] - junit_base[c]) * std::max(nset, 1) * ncp[c];
        }
        ja.unit_pc = d_junit_pc.as<int>();
        ja.unit_e = d_junit_e.as<long long>();
        ja.pair_unit = d_pair_unit.as<int>();
        ja.jpart = d_jpart2.as<double>();
        ja.dens = dens;
        ja.J = J;
        ja.delta = 0;
        ja.unit_q = d_junit_q.as<double>();
        ja.dmax = d_dmax.as<double>();
        ja.dthr = 0.0;
        ja.d32 = 0.0;
        ja.jdata32b = have_32b ? d_jdata32b.as<float>() : nullptr;
        return ja;
    }

    // Evaluate the screened quartets into the store and build the J and K
    // gather lists. Synchronises the stream.
    void build_store(hipStream_t st) {
        long long off = 0;
        for (int k = 0; k < NQCLASS; ++k) {
            store_base[k] = off;
            off += qcount[k] * aimdhf::store_block(k);
        }
        store_base[NQCLASS] = off;
        const long long keys = (long long)npair * (npair + 1) / 2;
        d_store.ensure(sizeof(double) * std::max<long long>(off, 1));
        d_qmap.ensure(sizeof(int) * std::max<long long>(keys, 1));
        d_lerr.ensure(sizeof(int));
        // the integrals run on the library streams while the caller's stream builds
        // the gather lists (they need only the pair-pair map); joined before the copies
        fan.fork(st);
        {
            aimdhf::StoreArgs sa;
            sa.recs = d_recs.as<double>();
            sa.rec_off = d_rec_off.as<long long>();
            sa.npp = d_npp.as<int>();
            sa.boys = d_boys.as<double>();
            for (int k = 0, i = 0; k < NQCLASS; ++k) {
                if (!qcount[k]) continue;
                sa.store = d_store.as<double>() + store_base[k];
                hipStream_t s2 = fan.on(1 + i++ % Fan::NAUX);
                prof.time(47, s2, [&] { aimdhf::launch_eri_store(k, sa, list_of_class(k), qcount[k], q64count[k], s2); });
            }
        }
        {
            // partner and ket lists on the host while the integrals run
            auto t0 = Profiler::clock::now();
            build_owner_lists(threshold, false, st);
            prof.host(86, t0);
        }
        prof.time(48, st, [&] {
            aimdhf::launch_fill_int(d_qmap.as<int>(), (size_t)keys, -1, st);
            aimdhf::launch_fill_int(d_lerr.as<int>(), 1, 0, st);
            aimdhf::QmapArgs qa;
            qa.list = d_list.as<int2>();
            qa.qoff[0] = 0;
            for (int k = 0; k < NQCLASS; ++k) qa.qoff[k + 1] = qa.qoff[k] + qcount[k];
            qa.qmap = d_qmap.as<int>();
            aimdhf::launch_qmap(qa, st);
        });
        HIPCHECK(hipGetLastError());
        aimdhf::ListArgs la;
        la.npair = npair;
        la.thr = threshold;
        // mixed precision: bands two decades apart from the FP64 / FP32 switch down, the
        // first held in FP64; FP64: one band (finer units only slow the complete builds)
        const bool split = precision == 1 && jk_switch > threshold;
        const int nbig = split ? 1 : aimdhf::NBAND;   // leading bands held in FP64
        for (int b = 0; b < aimdhf::NBAND - 1; ++b)
            la.band_lo[b] = split ? std::max(jk_switch * std::pow(1e-2, b), threshold) : threshold;
        la.band_lo[aimdhf::NBAND - 1] = threshold;
        la.out_x = d_out_x.as<int>();
        la.out_y = d_out_y.as<int>();
        la.shell_l = d_shell_l.as<int>();
        la.shell_ao = d_shell_ao.as<int>();
        la.pair_a = d_pair_a.as<int>();
        la.pair_b = d_pair_b.as<int>();
        la.pair_cls = d_pair_cls.as<int>();
        la.part_off = d_part_off.as<int>();
        la.part_shell = d_part_shell.as<int>();
        la.part_pair = d_part_pair.as<int>();
        la.part_q = d_part_q.as<double>();
        la.qscr = d_qscr.as<double>();
        la.ket_pair = d_ket_pair.as<int>();
        la.ket_q = d_ket_q.as<double>();
        for (int c = 0; c < 7; ++c) la.ket_off[c] = ket_off[c];
        la.qmap = d_qmap.as<int>();
        for (int k = 0; k < NQCLASS; ++k) la.qcount[k] = qcount[k];
        la.err = d_lerr.as<int>();
        const int nout = out_off[9];
        constexpr int NB = aimdhf::NBAND;
        const long long nks = (long long)nout * 9, njs = (long long)npair * 6;
        if (nks * NB + 1 > aimdhf::scan_max_items() || njs * NB + 1 > aimdhf::scan_max_items()) {
            g_err = "integral store: too many gather segments for the table scans";
            throw HipError{AIMDHF_E_UNSUPPORTED};
        }
        // device tables: per-band counts -> sizes -> scans -> units; only totals come back
        d_kcnt.ensure(sizeof(long long) * (NB + 1) * std::max<long long>(nks, 1));   // bands, then totals
        d_jcnt.ensure(sizeof(long long) * (NB + 1) * std::max<long long>(njs, 1));
        d_kseg.ensure(sizeof(long long) * (nks + 1));
        d_jseg.ensure(sizeof(long long) * (njs + 1));
        d_kdoff.ensure(sizeof(long long) * (nks + 1));
        d_kdoff32.ensure(sizeof(long long) * (nks + 1));
        d_jdoff.ensure(sizeof(long long) * (njs + 1));
        d_jdoff32.ensure(sizeof(long long) * (njs + 1));
        d_kbig.ensure(sizeof(long long) * std::max<long long>(nks, 1));
        d_jbig.ensure(sizeof(long long) * std::max<long long>(njs, 1));
        const long long nku_slots = nks * NB + 1, nju_slots = njs * NB + 1;
        const long long nbsum = (std::max(nku_slots, nju_slots) + 4095) / 4096 + 1;
        const long long tab_n = 3 * nks + (nks * NB) + nku_slots + 3 * njs + (njs * NB) + nju_slots + nbsum + 32;
        d_tab.ensure(sizeof(long long) * tab_n);
        d_out_unit.ensure(sizeof(int) * (nout + 1));
        d_pair_unit.ensure(sizeof(int) * (npair + 1));
        aimdhf::TabArgs ta;
        ta.nout = nout;
        ta.npair = npair;
        ta.nbig = nbig;
        ta.unit_items = unit_items;
        for (int t = 0; t < 10; ++t) ta.out_off[t] = out_off[t];
        for (int c = 0; c < 7; ++c) ta.cls_pair_off[c] = cls_pair_off[c];
        ta.out_x = d_out_x.as<int>();
        ta.out_y = d_out_y.as<int>();
        ta.shell_l = d_shell_l.as<int>();
        ta.pair_cls = d_pair_cls.as<int>();
        ta.cls_pairs = d_cls_pairs.as<int>();
        ta.kband = d_kcnt.as<long long>();
        ta.jband = d_jcnt.as<long long>();
        long long* tp = d_tab.as<long long>();
        auto take = [&](long long n) { long long* r = tp; tp += n; return r; };
        ta.kcount = take(nks);
        ta.ksz64 = take(nks);
        ta.ksz32 = take(nks);
        ta.knu = take(nks * NB);
        ta.kustart = take(nku_slots);
        ta.jcount = take(njs);
        ta.jsz64 = take(njs);
        ta.jsz32 = take(njs);
        ta.jnu = take(njs * NB);
        ta.justart = take(nju_slots);
        long long* bsum = take(nbsum);
        long long* tot = take(32);      // totals and check flag of launch_tab_totals
        ta.kbig = d_kbig.as<long long>();
        ta.jbig = d_jbig.as<long long>();
        ta.kseg = d_kseg.as<long long>();
        ta.kdoff = d_kdoff.as<long long>();
        ta.kdoff32 = d_kdoff32.as<long long>();
        ta.jseg = d_jseg.as<long long>();
        ta.jdoff = d_jdoff.as<long long>();
        ta.jdoff32 = d_jdoff32.as<long long>();
        ta.out_unit = d_out_unit.as<int>();
        ta.pair_unit = d_pair_unit.as<int>();
        ta.kunit_os = nullptr;
        ta.kunit_e = nullptr;
        ta.junit_pc = nullptr;
        ta.junit_e = nullptr;
        la.kband = d_kcnt.as<long long>();
        if (!h_pin) HIPCHECK(hipHostMalloc((void**)&h_pin, sizeof(long long) * 64));
        auto th = Profiler::clock::now();
        prof.time(48, st, [&] {
            aimdhf::launch_klist_count(la, nout, d_kcnt.as<long long>() + NB * nks, d_kcnt.as<long long>(), st);
            aimdhf::launch_jlist_count(la, d_jcnt.as<long long>() + NB * njs, d_jcnt.as<long long>(), st);
            aimdhf::launch_seg_sizes(ta, st);
            aimdhf::launch_scan(ta.kcount, ta.kseg, nks, bsum, st);
            aimdhf::launch_scan(ta.ksz64, ta.kdoff, nks, bsum, st);
            aimdhf::launch_scan(ta.ksz32, ta.kdoff32, nks, bsum, st);
            aimdhf::launch_scan(ta.knu, ta.kustart, nks * NB, bsum, st);
            aimdhf::launch_scan(ta.jcount, ta.jseg, njs, bsum, st);
            aimdhf::launch_scan(ta.jsz64, ta.jdoff, njs, bsum, st);
            aimdhf::launch_scan(ta.jsz32, ta.jdoff32, njs, bsum, st);
            aimdhf::launch_scan(ta.jnu, ta.justart, njs * NB, bsum, st);
            aimdhf::launch_tab_totals(ta, tot, st);
        });
        HIPCHECK(hipGetLastError());
        HIPCHECK(hipMemcpyAsync(h_pin, tot, sizeof(long long) * 26, hipMemcpyDeviceToHost, st));
        HIPCHECK(hipStreamSynchronize(st));
        if (h_pin[25]) {
            g_err = "integral store: inconsistent gather-segment tables";
            throw HipError{AIMDHF_E_STATE};
        }
        prof.host(89, th);
        th = Profiler::clock::now();
        k_entries = h_pin[0];
        j_entries = h_pin[1];
        const long long ko = h_pin[2], ko32 = h_pin[3], jo = h_pin[4], jo32 = h_pin[5];
        const long long nku = h_pin[6], nju = h_pin[7];
        for (int t = 0; t < 10; ++t) kunit_base[t] = (int)h_pin[8 + t];
        for (int c = 0; c < 7; ++c) junit_base[c] = (int)h_pin[18 + c];
        kdata_total = ko;
        jdata_total = jo;
        kdata32_total = ko32;
        jdata32_total = jo32;
        d_klist.ensure(sizeof(long long) * std::max<long long>(k_entries, 1));
        d_jlist.ensure(sizeof(long long) * std::max<long long>(j_entries, 1));
        d_kunit_os.ensure(sizeof(int) * std::max<long long>(nku, 1));
        d_kunit_e.ensure(sizeof(long long) * 2 * std::max<long long>(nku, 1));
        d_junit_pc.ensure(sizeof(int) * std::max<long long>(nju, 1));
        d_junit_e.ensure(sizeof(long long) * 2 * std::max<long long>(nju, 1));
        d_kpart2.ensure(sizeof(double) * 2 * 36 * std::max<long long>(nku, 1));
        d_jpart2.ensure(sizeof(double) * 2 * 36 * std::max<long long>(nju, 1));
        d_kdata.ensure(sizeof(double) * std::max<long long>(ko, 1));
        d_jdata.ensure(sizeof(double) * std::max<long long>(jo, 1));
        have_32b = precision == 1;
        if (have_32b) {
            d_kdata32b.ensure(sizeof(float) * std::max<long long>(ko, 1));
            d_jdata32b.ensure(sizeof(float) * std::max<long long>(jo, 1));
        }
        d_kdata32.ensure(sizeof(float) * std::max<long long>(ko32, 1));
        d_jdata32.ensure(sizeof(float) * std::max<long long>(jo32, 1));
        ta.kunit_os = d_kunit_os.as<int>();
        ta.kunit_e = d_kunit_e.as<long long>();
        ta.junit_pc = d_junit_pc.as<int>();
        ta.junit_e = d_junit_e.as<long long>();
        la.kseg = d_kseg.as<long long>();
        la.klist = d_klist.as<long long>();
        la.jseg = d_jseg.as<long long>();
        la.jlist = d_jlist.as<long long>();
        prof.host(90, th);
        th = Profiler::clock::now();
        prof.time(48, st, [&] {
            aimdhf::launch_klist_fill(la, nout, st);
            aimdhf::launch_jlist_fill(la, st);
            aimdhf::launch_units(ta, st);
        });
        HIPCHECK(hipGetLastError());
        HIPCHECK(hipMemcpyAsync(h_pin + 32, d_lerr.p, sizeof(int), hipMemcpyDeviceToHost, st));
        HIPCHECK(hipStreamSynchronize(st));
        if (*(const int*)(h_pin + 32)) {
            g_err = "integral store: a gather-list quartet is missing from the screened list";
            throw HipError{AIMDHF_E_STATE};
        }
        prof.host(91, th);
        th = Profiler::clock::now();
        prof.host(92, th);
        fan.join();
        prof.time(15, st, [&] {
            fan.fork(st);
            int i = 0;
            aimdhf::KSArgs ka = ks_args(1, nullptr, nullptr);
            for (int t = 0; t < 9; ++t) {
                const int no = out_off[t + 1] - out_off[t];
                if (no > 0) aimdhf::launch_kcopy(t, ka, no, fan.on(i++));
            }
            aimdhf::JSArgs ja = js_args(1, nullptr, nullptr);
            for (int c = 0; c < 6; ++c) {
                if (!ncls[c]) continue;
                ja.cls_off = cls_pair_off[c];
                aimdhf::launch_jcopy(c, ja, ncls[c], fan.on(i++));
            }
            fan.join();
        });
        HIPCHECK(hipGetLastError());
        have_unit_q = false;   // unit bounds: on the first density-difference build
        store_bytes = 8 * (off + ko + jo) + 4 * (ko32 + jo32) + (have_32b ? 4 * (ko + jo) : 0) + 4 * keys +
                      8 * (k_entries + j_entries + 4 * nks + 4 * njs);
    }

    // Largest Schwarz product of every work unit of the integral store, which
    // density-difference builds screen with; computed once per store, on the first
    // such build (stream-ordered, no synchronisation).
    bool have_unit_q = false;
    void ensure_unit_q(hipStream_t st) {
        if (have_unit_q) return;
        const int nku = kunit_base[9], nju = junit_base[6];
        d_kunit_q.ensure(sizeof(double) * std::max(nku, 1));
        d_junit_q.ensure(sizeof(double) * std::max(nju, 1));
        d_dmax.ensure(sizeof(double));
        aimdhf::UnitQArgs ua;
        ua.list = d_list.as<int2>();
        long long qo = 0;
        for (int k = 0; k < NQCLASS; ++k) {
            ua.qoff[k] = qo;
            qo += qcount[k];
        }
        ua.qscr = d_qscr.as<double>();
        ua.shell_l = d_shell_l.as<int>();
        ua.out_x = d_out_x.as<int>();
        ua.out_y = d_out_y.as<int>();
        ua.pair_cls = d_pair_cls.as<int>();
        ua.klist = d_klist.as<long long>();
        ua.kunit_os = d_kunit_os.as<int>();
        ua.kunit_e = d_kunit_e.as<long long>();
        ua.nkunit = nku;
        ua.kunit_q = d_kunit_q.as<double>();
        ua.jlist = d_jlist.as<long long>();
        ua.junit_pc = d_junit_pc.as<int>();
        ua.junit_e = d_junit_e.as<long long>();
        ua.njunit = nju;
        ua.junit_q = d_junit_q.as<double>();
        prof.time(48, st, [&] { aimdhf::launch_unit_q(ua, st); });
        HIPCHECK(hipGetLastError());
        have_unit_q = true;
    }

    aimdhf::OneElArgs oneel_args() const {
        aimdhf::OneElArgs g;
        g.npair = npair;
        g.nao = nao;
        g.natm = natm;
        for (int c = 0; c < 6; ++c) {
            g.ncls_host[c] = ncls[c];
            g.cls_off_host[c] = cls_pair_off[c];
        }
        g.cls_pairs = d_cls_pairs.as<int>();
        g.pair_a = d_pair_a.as<int>();
        g.pair_b = d_pair_b.as<int>();
        g.shell_atom = d_shell_atom.as<int>();
        g.shell_l = d_shell_l.as<int>();
        g.shell_prim = d_shell_prim.as<int>();
        g.shell_ao = d_shell_ao.as<int>();
        g.prim_exp = d_prim_exp.as<double>();
        g.prim_coef = d_prim_coef.as<double>();
        g.pos = d_pos.as<double>();
        g.charges = d_charges.as<double>();
        g.boys = d_boys.as<double>();
        g.parts = 3;
        g.ld = batched ? bld : nao;
        g.mstride = (long long)mat_elems();
        g.shell_rep = batched ? d_shell_rep.as<int>() : nullptr;
        g.rep_atom = batched ? d_rep_atom.as<int>() : nullptr;
        return g;
    }

    aimdhf::BatchArgs batch_args() const {
        aimdhf::BatchArgs b;
        b.nrep = nrep;
        b.ld = bld;
        b.rep_nao = d_rep_nao.as<int>();
        b.rep_toff = d_rep_toff.as<long long>();
        b.rep_item = d_rep_item.as<long long>();
        b.shell_rep = d_shell_rep.as<int>();
        b.shell_ao = d_shell_ao.as<int>();
        b.shell_l = d_shell_l.as<int>();
        b.pair_a = d_pair_a.as<int>();
        b.pair_b = d_pair_b.as<int>();
        b.list = d_list.as<int2>();
        b.qoff[0] = 0;
        for (int k = 0; k < NQCLASS; ++k) {
            b.qoff[k + 1] = b.qoff[k] + qcount[k];
            b.store_base[k] = store_base[k];
            b.blk[k] = (int)aimdhf::store_block(k);
            int hi = 0;
            while ((hi + 1) * (hi + 2) / 2 <= k) ++hi;
            const int lo = k - hi * (hi + 1) / 2, lc = lo == 0 ? 0 : lo <= 2 ? 1 : 2, ld_ = lo - lc * (lc + 1) / 2;
            b.ncd[k] = ncart(lc) * ncart(ld_);
        }
        b.store_base[NQCLASS] = store_base[NQCLASS];
        b.store = d_store.as<double>();
        b.tens = d_tens.as<double>();
        return b;
    }

    // Batch geometry: every screened quartet of every replica into the store (one
    // launch per class, FP64), then into the replicas' dense tensors
    void build_batch_tensors(hipStream_t st) {
        long long off = 0;
        for (int k = 0; k < NQCLASS; ++k) {
            store_base[k] = off;
            off += qcount[k] * aimdhf::store_block(k);
        }
        store_base[NQCLASS] = off;
        d_store.ensure(sizeof(double) * std::max<long long>(off, 1));
        d_tens.ensure(sizeof(double) * std::max<long long>(tens_total, 1));
        aimdhf::StoreArgs sa;
        sa.recs = d_recs.as<double>();
        sa.rec_off = d_rec_off.as<long long>();
        sa.npp = d_npp.as<int>();
        sa.boys = d_boys.as<double>();
        prof.time(47, st, [&] {
            for (int k = 0; k < NQCLASS; ++k) {
                if (!qcount[k]) continue;
                sa.store = d_store.as<double>() + store_base[k];
                aimdhf::launch_eri_store(k, sa, list_of_class(k), qcount[k], q64count[k], st);
            }
        });
        prof.time(15, st, [&] {
            aimdhf::launch_zero_d(d_tens.as<double>(), (size_t)tens_total, st);
            aimdhf::launch_batch_scatter(batch_args(), off, st);
        });
        HIPCHECK(hipGetLastError());
    }
};

static int pair_class(int la, int lb) { return la * (la + 1) / 2 + lb; }

extern "C" {

int aimdhf_last_error(char* buf, size_t len) {
    if (buf && len) {
        std::strncpy(buf, g_err.c_str(), len - 1);
        buf[len - 1] = 0;
    }
    return AIMDHF_OK;
}

int aimdhf_lmax(void) { return aimdhf::GEN_LMAX; }

}  // extern "C"

// Plan of one molecule (nrep = 0) or of nrep replicas (shells and atoms grouped
// by replica: rep_shell / rep_atom hold nrep + 1 offsets)
static int create_plan(int natm, int nshell, const int64_t* shell_atom, const int64_t* shell_l,
                       const int64_t* shell_prim, const int64_t* shell_block, const double* prim_exp,
                       const double* prim_coef, const double* charges, int nrep, const int64_t* rep_atom,
                       const int64_t* rep_shell, aimdhf_plan** out) {
    if (!out) return fail(AIMDHF_E_ARG, "out is NULL");
    *out = nullptr;
    if (natm <= 0 || nshell <= 0) return fail(AIMDHF_E_ARG, "natm and nshell must be positive");
    if (!shell_atom || !shell_l || !shell_prim || !shell_block || !prim_exp || !prim_coef || !charges)
        return fail(AIMDHF_E_ARG, "NULL basis array");
    if (nrep > 0) {
        if (!rep_atom || !rep_shell) return fail(AIMDHF_E_ARG, "NULL replica offsets");
        if (rep_atom[0] != 0 || rep_shell[0] != 0 || rep_atom[nrep] != natm || rep_shell[nrep] != nshell)
            return fail(AIMDHF_E_ARG, "replica offsets must run from 0 to natm and nshell");
        for (int r = 0; r < nrep; ++r)
            if (rep_atom[r + 1] <= rep_atom[r] || rep_shell[r + 1] <= rep_shell[r])
                return fail(AIMDHF_E_ARG, "every replica needs at least one atom and one shell");
        for (int r = 0; r < nrep; ++r)
            for (int64_t sh = rep_shell[r]; sh < rep_shell[r + 1]; ++sh)
                if (shell_atom[sh] < rep_atom[r] || shell_atom[sh] >= rep_atom[r + 1])
                    return fail(AIMDHF_E_ARG, "a shell sits on an atom of another replica");
    }
    aimdhf_plan* p = new (std::nothrow) aimdhf_plan();
    if (!p) return fail(AIMDHF_E_NOMEM, "host allocation failed");
    try {
        p->natm = natm;
        p->nshell = nshell;
        p->batched = nrep > 0;
        p->nrep = nrep > 0 ? nrep : 1;
        // replica of every shell and the first shell of its replica (one replica: 0)
        p->shell_rep.assign(nshell, 0);
        p->shell_first.assign(nshell, 0);
        p->rep_shell.assign(p->nrep + 1, 0);
        p->rep_atom.assign(p->nrep + 1, 0);
        p->rep_shell[p->nrep] = nshell;
        p->rep_atom[p->nrep] = natm;
        if (p->batched)
            for (int r = 0; r < nrep; ++r) {
                p->rep_shell[r] = (int)rep_shell[r];
                p->rep_atom[r] = (int)rep_atom[r];
                for (int64_t sh = rep_shell[r]; sh < rep_shell[r + 1]; ++sh) {
                    p->shell_rep[sh] = r;
                    p->shell_first[sh] = (int)rep_shell[r];
                }
            }
        int nprim = (int)shell_prim[nshell];
        p->shell_atom.resize(nshell);
        p->shell_l.resize(nshell);
        p->shell_prim.resize(nshell + 1);
        p->shell_block.resize(nshell);
        p->shell_ao.resize(nshell + 1);
        p->shell_ao[0] = 0;
        for (int s = 0; s < nshell; ++s) {
            if (shell_l[s] < 0 || shell_l[s] > aimdhf::GEN_LMAX || (nrep > 0 && shell_l[s] > 2)) {
                delete p;
                return fail(AIMDHF_E_UNSUPPORTED,
                            "angular momentum l = " + std::to_string(shell_l[s]) + " is not supported" +
                                (nrep > 0 ? std::string(" in batched plans (l <= 2)")
                                          : " (l <= " + std::to_string(aimdhf::GEN_LMAX) + ")"));
            }
            if (shell_atom[s] < 0 || shell_atom[s] >= natm || shell_prim[s + 1] <= shell_prim[s]) {
                delete p;
                return fail(AIMDHF_E_ARG, "inconsistent shell arrays");
            }
            p->shell_atom[s] = (int)shell_atom[s];
            p->shell_l[s] = (int)shell_l[s];
            p->shell_prim[s] = (int)shell_prim[s];
            p->shell_block[s] = (int)shell_block[s];
            p->shell_ao[s + 1] = p->shell_ao[s] + ncart(p->shell_l[s]);
            p->lmax = std::max(p->lmax, p->shell_l[s]);
        }
        p->shell_prim[nshell] = nprim;
        p->nao = p->shell_ao[nshell];
        if (p->lmax > 2) {
            // runtime-l kernels for every integral: shell pairs, bounds, no pair records
            p->generic = true;
            p->nblock = p->shell_block[nshell - 1] + 1;
            p->nbpair = p->nblock * (p->nblock + 1) / 2;
            p->prim_exp.assign(prim_exp, prim_exp + nprim);
            p->prim_coef.assign(prim_coef, prim_coef + nprim);
            p->charges.assign(charges, charges + natm);
            p->npair = nshell * (nshell + 1) / 2;
            for (int i = 0; i < nshell; ++i)
                for (int j = 0; j <= i; ++j) {
                    p->gpair_i.push_back(i);
                    p->gpair_j.push_back(j);
                    const int b1 = p->shell_block[i], b2 = p->shell_block[j];
                    p->gpair_bp.push_back(b1 * (b1 + 1) / 2 + b2);
                }
            p->qcount.assign(NQCLASS, 0);
            p->gcount.assign(NQCLASS, 0);
            upload(p->d_shell_atom, p->shell_atom);
            upload(p->d_shell_l, p->shell_l);
            upload(p->d_shell_prim, p->shell_prim);
            upload(p->d_shell_ao, p->shell_ao);
            upload(p->d_prim_exp, p->prim_exp);
            upload(p->d_prim_coef, p->prim_coef);
            upload(p->d_charges, p->charges);
            upload(p->d_boys, boys_table());
            upload(p->d_gpair_i, p->gpair_i);
            upload(p->d_gpair_j, p->gpair_j);
            p->d_scratch.ensure(sizeof(double) * 1024);
            p->d_pos.ensure(sizeof(double) * 3 * natm);
            p->d_qshell.ensure(sizeof(double) * p->npair);
            p->d_qblock.ensure(sizeof(double) * p->nbpair);
            *out = p;
            return AIMDHF_OK;
        }
        if (p->batched) {
            // AO offsets within each replica; replica sizes, padded dimension, tensors
            p->rep_nao.assign(p->nrep, 0);
            for (int r = 0; r < p->nrep; ++r) {
                const int a0 = p->shell_ao[p->rep_shell[r]];
                p->rep_nao[r] = p->shell_ao[p->rep_shell[r + 1]] - a0;
                p->bld = std::max(p->bld, p->rep_nao[r]);
            }
            std::vector<int> local(nshell + 1);
            for (int sh = 0; sh < nshell; ++sh) local[sh] = p->shell_ao[sh] - p->shell_ao[p->shell_first[sh]];
            local[nshell] = 0;
            p->shell_ao = local;
            p->rep_toff.assign(p->nrep + 1, 0);
            p->rep_item.assign(p->nrep + 1, 0);
            for (int r = 0; r < p->nrep; ++r) {
                const long long n = p->rep_nao[r];
                p->rep_toff[r + 1] = p->rep_toff[r] + n * n * n * n;
                p->rep_item[r + 1] = p->rep_item[r] + n * n;
            }
            p->tens_total = p->rep_toff[p->nrep];
        }
        p->nblock = p->shell_block[nshell - 1] + 1;
        p->prim_exp.assign(prim_exp, prim_exp + nprim);
        p->prim_coef.assign(prim_coef, prim_coef + nprim);
        p->charges.assign(charges, charges + natm);
        // shell pairs (s1 >= s2) in the reference's pair order, oriented l_A >= l_B; in a
        // batch only the pairs within a replica, replica by replica
        p->npair = 0;
        for (int r = 0; r < p->nrep; ++r) {
            const int ns = p->rep_shell[r + 1] - p->rep_shell[r];
            p->npair += ns * (ns + 1) / 2;
        }
        p->pair_a.resize(p->npair);
        p->pair_b.resize(p->npair);
        p->pair_cls.resize(p->npair);
        p->rec_size.resize(p->npair);
        p->rec_off.resize(p->npair + 1);
        long long off = 0;
        int P = 0;
        std::vector<std::vector<int>> bylcls(6);
        // block pairs within each replica, replica by replica (one molecule: i(i+1)/2+j)
        std::vector<int> rep_b0(p->nrep), rep_bp0(p->nrep + 1, 0), rep_p0(p->nrep + 1, 0);
        for (int r = 0; r < p->nrep; ++r) {
            rep_b0[r] = p->shell_block[p->rep_shell[r]];
            const int nb = p->shell_block[p->rep_shell[r + 1] - 1] - rep_b0[r] + 1;
            rep_bp0[r + 1] = rep_bp0[r] + nb * (nb + 1) / 2;
            const int ns = p->rep_shell[r + 1] - p->rep_shell[r];
            rep_p0[r + 1] = rep_p0[r] + ns * (ns + 1) / 2;
        }
        p->nbpair = rep_bp0[p->nrep];
        std::vector<std::vector<int>> bpl(p->nbpair);
        p->shell_pair0.assign(nshell, 0);
        p->pair_lo.assign(p->npair, 0);
        for (int s1 = 0; s1 < nshell; ++s1) p->shell_pair0[s1] = rep_p0[p->shell_rep[s1]];
        for (int s1 = 0; s1 < nshell; ++s1)
            for (int s2 = p->shell_first[s1]; s2 <= s1; ++s2, ++P) {
                p->pair_lo[P] = p->shell_pair0[s1];
                int a = s1, b = s2;
                if (p->shell_l[s2] > p->shell_l[s1]) std::swap(a, b);
                int la = p->shell_l[a], lb = p->shell_l[b];
                p->pair_a[P] = a;
                p->pair_b[P] = b;
                p->pair_cls[P] = pair_class(la, lb);
                int rs = REC_HDR + nherm(la + lb) * ncart(la) * ncart(lb);
                p->rec_size[P] = rs;
                p->rec_off[P] = off;
                off += (long long)rs * (shell_prim[s1 + 1] - shell_prim[s1]) * (shell_prim[s2 + 1] - shell_prim[s2]);
                bylcls[p->pair_cls[P]].push_back(P);
                const int r = p->shell_rep[s1];
                int b1 = p->shell_block[s1] - rep_b0[r], b2 = p->shell_block[s2] - rep_b0[r];
                bpl[rep_bp0[r] + b1 * (b1 + 1) / 2 + b2].push_back(P);
            }
        p->rec_off[p->npair] = off;
        p->rec_total = off;
        p->cls_pair_off.assign(7, 0);
        p->ncls.assign(6, 0);
        for (int c = 0; c < 6; ++c) {
            p->cls_pair_off[c + 1] = p->cls_pair_off[c] + (int)bylcls[c].size();
            p->ncls[c] = (int)bylcls[c].size();
            p->cls_pairs.insert(p->cls_pairs.end(), bylcls[c].begin(), bylcls[c].end());
        }
        p->bp_off.assign(p->nbpair + 1, 0);
        for (int i = 0; i < p->nbpair; ++i) {
            p->bp_off[i + 1] = p->bp_off[i] + (int)bpl[i].size();
            p->bp_list.insert(p->bp_list.end(), bpl[i].begin(), bpl[i].end());
        }
        // exchange output blocks (X >= Y) grouped by (l_X, l_Y); a batch digests its dense
        // tensors instead and needs neither these nor the AO -> shell map
        if (!p->batched) {
            std::vector<std::vector<int>> ox(9), oy(9);
            for (int X = 0; X < nshell; ++X)
                for (int Y = 0; Y <= X; ++Y) {
                    int t = p->shell_l[X] * 3 + p->shell_l[Y];
                    ox[t].push_back(X);
                    oy[t].push_back(Y);
                }
            for (int t = 0; t < 9; ++t) {
                p->out_off[t + 1] = p->out_off[t] + (int)ox[t].size();
                p->out_x.insert(p->out_x.end(), ox[t].begin(), ox[t].end());
                p->out_y.insert(p->out_y.end(), oy[t].begin(), oy[t].end());
            }
            for (int s2 = 0; s2 < nshell; ++s2) p->has_l[p->shell_l[s2]] = 1;
            p->ao_shell.resize(p->shell_ao[nshell]);
            for (int s2 = 0; s2 < nshell; ++s2)
                for (int i = p->shell_ao[s2]; i < p->shell_ao[s2 + 1]; ++i) p->ao_shell[i] = s2;
        }
        // device copies
        upload(p->d_out_x, p->out_x);
        upload(p->d_out_y, p->out_y);
        upload(p->d_ao_shell, p->ao_shell);
        upload(p->d_shell_atom, p->shell_atom);
        upload(p->d_shell_l, p->shell_l);
        upload(p->d_shell_prim, p->shell_prim);
        upload(p->d_shell_ao, p->shell_ao);
        upload(p->d_prim_exp, p->prim_exp);
        upload(p->d_prim_coef, p->prim_coef);
        upload(p->d_pair_a, p->pair_a);
        upload(p->d_pair_b, p->pair_b);
        upload(p->d_pair_cls, p->pair_cls);
        upload(p->d_rec_size, p->rec_size);
        upload(p->d_rec_off, p->rec_off);
        upload(p->d_cls_pairs, p->cls_pairs);
        upload(p->d_cls_pair_off, p->cls_pair_off);
        upload(p->d_ncls, p->ncls);
        upload(p->d_bp_off, p->bp_off);
        upload(p->d_bp_list, p->bp_list);
        upload(p->d_boys, boys_table());
        upload(p->d_charges, p->charges);
        if (p->batched) {
            upload(p->d_shell_rep, p->shell_rep);
            upload(p->d_shell_first, p->shell_first);
            upload(p->d_shell_pair0, p->shell_pair0);
            upload(p->d_pair_lo, p->pair_lo);
            upload(p->d_rep_atom, p->rep_atom);
            upload(p->d_rep_nao, p->rep_nao);
            upload(p->d_rep_toff, p->rep_toff);
            upload(p->d_rep_item, p->rep_item);
        }
        p->d_scratch.ensure(sizeof(double) * 1024);
        p->d_pos.ensure(sizeof(double) * 3 * natm);
        p->d_recs.ensure(sizeof(double) * std::max<long long>(off, 1));
        p->d_npp.ensure(sizeof(int) * p->npair);
        p->d_qshell.ensure(sizeof(double) * p->npair);
        p->d_qscr.ensure(sizeof(double) * p->npair);
        p->d_qblock.ensure(sizeof(double) * p->nbpair);
        p->qcount.assign(NQCLASS, 0);
        p->noalloc_ptr = aimdhf::noalloc_scratch(&p->noalloc_bytes);
        if (!p->noalloc_ptr) {
            delete p;
            return fail(AIMDHF_E_HIP, "hipGetSymbolAddress of the fixed scratch array failed");
        }
    } catch (const HipError& e) {
        delete p;
        return e.code;
    } catch (const std::bad_alloc&) {
        delete p;
        return fail(AIMDHF_E_NOMEM, "host allocation failed");
    }
    *out = p;
    return AIMDHF_OK;
}

// ------------------------------------------------------------ runtime-l path
static int set_geometry_generic(aimdhf_plan* p, const double* positions, double thr, hipStream_t st) {
    try {
        p->pos.assign(positions, positions + 3 * p->natm);
        HIPCHECK(hipMemcpyAsync(p->d_pos.p, positions, sizeof(double) * 3 * p->natm, hipMemcpyHostToDevice, st));
        aimdhf::GenArgs g = p->gen_args();
        p->prof.time(42, st, [&] { aimdhf::launch_gen_schwarz(g, p->npair, st); });
        HIPCHECK(hipGetLastError());
        std::vector<double> qs(p->npair);
        HIPCHECK(hipMemcpyAsync(qs.data(), p->d_qshell.p, sizeof(double) * p->npair, hipMemcpyDeviceToHost, st));
        HIPCHECK(hipStreamSynchronize(st));
        auto t0 = Profiler::clock::now();
        // block-pair bounds: the largest shell-pair bound of the block pair
        p->gen_qb.assign(p->nbpair, 0.0);
        for (int P = 0; P < p->npair; ++P) p->gen_qb[p->gpair_bp[P]] = std::max(p->gen_qb[p->gpair_bp[P]], qs[P]);
        upload(p->d_qblock, p->gen_qb);
        // kept primitive pairs (mu |AB|^2 <= 50) per shell pair
        p->gen_npp.assign(p->npair, 0);
        for (int P = 0; P < p->npair; ++P) {
            const int a = p->gpair_i[P], b = p->gpair_j[P];
            double r2 = 0.0;
            for (int d = 0; d < 3; ++d) {
                const double x = p->pos[3 * p->shell_atom[a] + d] - p->pos[3 * p->shell_atom[b] + d];
                r2 += x * x;
            }
            int n = 0;
            for (int i = p->shell_prim[a]; i < p->shell_prim[a + 1]; ++i)
                for (int j = p->shell_prim[b]; j < p->shell_prim[b + 1]; ++j) {
                    const double x = p->prim_exp[i], y = p->prim_exp[j];
                    if (x * y / (x + y) * r2 <= 50.0) ++n;
                }
            p->gen_npp[P] = n;
        }
        p->qtotal = p->gen_build_list(thr, p->d_list);
        p->threshold = thr;
        p->stored = false;
        p->store_bytes = p->k_entries = p->j_entries = 0;
        p->prof.host(85, t0);
        p->have_geometry = true;
    } catch (const HipError& e) {
        p->have_geometry = false;
        return e.code;
    }
    return AIMDHF_OK;
}

static int jk_generic(aimdhf_plan* p, int nset, const double* dens, double* J, double* K, hipStream_t st) {
    try {
        const size_t nn = (size_t)p->nao * p->nao;
        p->d_dsym.ensure(sizeof(double) * nset * nn);
        p->d_fx.ensure(sizeof(long long) * 4 * nset * nn);
        aimdhf::launch_symmetrize(dens, nset, p->nao, p->d_dsym.as<double>(), st);
        aimdhf::launch_zero64(p->d_fx.as<long long>(), 4 * nset * nn, st);
        aimdhf::GenArgs g = p->gen_args();
        g.nset = nset;
        g.dens = p->d_dsym.as<double>();
        g.fx = p->d_fx.as<long long>();
        g.quartets = p->d_list.as<int2>();
        g.nq = p->qtotal;
        p->prof.time(95, st, [&] { aimdhf::launch_gen_jk(g, st); });
        aimdhf::launch_jk_finalize(p->d_fx.as<long long>(), nset, p->nao, J, K, st);
        HIPCHECK(hipGetLastError());
    } catch (const HipError& e) {
        return e.code;
    }
    return AIMDHF_OK;
}

// one-electron gradient terms `parts` (1 kinetic, 2 nuclear attraction, 4 overlap) into grad
static void grad_one_generic(aimdhf_plan* p, const double* D, const double* W, int parts, double* grad, hipStream_t st) {
    const size_t nn = (size_t)p->nao * p->nao;
    p->d_gden.ensure(sizeof(double) * 2 * nn);
    double* Ds = p->d_gden.as<double>();
    double* Ws = Ds + nn;
    aimdhf::launch_symmetrize(D, 1, p->nao, Ds, st);
    aimdhf::launch_symmetrize(W, 1, p->nao, Ws, st);
    p->d_gfx.ensure(sizeof(long long) * 6 * p->natm);
    aimdhf::launch_zero64(p->d_gfx.as<long long>(), 6 * p->natm, st);
    aimdhf::GenArgs g = p->gen_args();
    g.D = Ds;
    g.W = Ws;
    g.parts = parts;
    g.fx = p->d_gfx.as<long long>();
    p->prof.time(44, st, [&] { aimdhf::launch_gen_grad1(g, p->npair, st); });
    aimdhf::launch_fx_to_double(p->d_gfx.as<long long>(), 3 * p->natm, grad, st);
    HIPCHECK(hipGetLastError());
}

static int grad_two_generic(aimdhf_plan* p, const double* Dc, int nx, const double* Dx, double kfac, double thr,
                            double* grad, hipStream_t st) {
    try {
        const size_t nn = (size_t)p->nao * p->nao;
        p->d_gden.ensure(sizeof(double) * (1 + nx) * nn);
        double* dc = p->d_gden.as<double>();
        double* dx = dc + nn;
        aimdhf::launch_symmetrize(Dc, 1, p->nao, dc, st);
        aimdhf::launch_symmetrize(Dx, nx, p->nao, dx, st);
        // quartets with q_P q_Q max|Gamma| >= thr for the largest possible |Gamma| element
        const double mc = aimdhf::device_absmax(dc, nn, p->d_scratch.as<double>(), st);
        const double mx = aimdhf::device_absmax(dx, nx * nn, p->d_scratch.as<double>(), st);
        if (mc < 0.0 || mx < 0.0) return fail(AIMDHF_E_HIP, "max reduction failed");
        const double gbound = 4.0 * (mc * mc + std::fabs(kfac) * nx * mx * mx);
        auto t0 = Profiler::clock::now();
        p->gen_gq = p->gen_build_list(gbound > 0.0 ? thr / gbound : 0.0, p->d_glist);
        p->prof.host(88, t0);
        p->d_gfx.ensure(sizeof(long long) * 6 * p->natm);
        aimdhf::launch_zero64(p->d_gfx.as<long long>(), 6 * p->natm, st);
        aimdhf::GenArgs g = p->gen_args();
        g.dc = dc;
        g.dx = dx;
        g.nx = nx;
        g.kfac = kfac;
        g.fx = p->d_gfx.as<long long>();
        g.quartets = p->d_glist.as<int2>();
        g.nq = p->gen_gq;
        p->prof.time(96, st, [&] { aimdhf::launch_gen_grad2(g, st); });
        aimdhf::launch_fx_to_double(p->d_gfx.as<long long>(), 3 * p->natm, grad, st);
        HIPCHECK(hipGetLastError());
    } catch (const HipError& e) {
        return e.code;
    }
    return AIMDHF_OK;
}

extern "C" {

int aimdhf_plan_create(int natm, int nshell, const int64_t* shell_atom, const int64_t* shell_l,
                       const int64_t* shell_prim, const int64_t* shell_block, const double* prim_exp,
                       const double* prim_coef, const double* charges, aimdhf_plan** out) {
    return create_plan(natm, nshell, shell_atom, shell_l, shell_prim, shell_block, prim_exp, prim_coef, charges, 0,
                       nullptr, nullptr, out);
}

int aimdhf_batch_create(int nrep, const int64_t* rep_atom, const int64_t* rep_shell, int natm, int nshell,
                        const int64_t* shell_atom, const int64_t* shell_l, const int64_t* shell_prim,
                        const int64_t* shell_block, const double* prim_exp, const double* prim_coef,
                        const double* charges, aimdhf_plan** out) {
    if (nrep <= 0) return fail(AIMDHF_E_ARG, "a batch needs at least one replica");
    return create_plan(natm, nshell, shell_atom, shell_l, shell_prim, shell_block, prim_exp, prim_coef, charges, nrep,
                       rep_atom, rep_shell, out);
}

int aimdhf_batch_layout(aimdhf_plan* p, int* nrep, int* ld, int64_t* tensor_bytes) {
    if (!p) return fail(AIMDHF_E_ARG, "NULL plan");
    if (nrep) *nrep = p->batched ? p->nrep : 0;
    if (ld) *ld = p->batched ? p->bld : p->nao;
    if (tensor_bytes) *tensor_bytes = p->batched ? 8 * p->tens_total : 0;
    return AIMDHF_OK;
}

int aimdhf_plan_destroy(aimdhf_plan* plan) {
    delete plan;
    return AIMDHF_OK;
}

int aimdhf_set_geometry(aimdhf_plan* p, const double* positions, double thr, void* stream) {
    if (!p || !positions) return fail(AIMDHF_E_ARG, "NULL plan or positions");
    for (int i = 0; i < 3 * p->natm; ++i)
        if (!std::isfinite(positions[i])) return fail(AIMDHF_E_ARG, "atomic positions must be finite (got NaN or inf)");
    if (!(thr >= 0.0)) return fail(AIMDHF_E_ARG, "schwarz threshold must be >= 0");
    hipStream_t st = (hipStream_t)stream;
    p->have_last_store = false;     // the store's buffers may move
    if (p->generic) return set_geometry_generic(p, positions, thr, st);
    try {
        p->pos.assign(positions, positions + 3 * p->natm);
        HIPCHECK(hipMemcpyAsync(p->d_pos.p, positions, sizeof(double) * 3 * p->natm, hipMemcpyHostToDevice, st));
        aimdhf::PairArgs g = p->pair_args();
        p->prof.time(42, st, [&] {
            aimdhf::launch_build_pairs(g, st);
            aimdhf::launch_schwarz(g, p->ncls.data(), st);
        });
        HIPCHECK(hipGetLastError());
        auto t0 = Profiler::clock::now();
        p->build_list(thr, p->precision == 1 && p->jk_switch > thr ? p->jk_switch : 0.0, p->d_list, p->qcount,
                      p->q64count, st);
        p->qtotal = 0;
        for (long long c : p->qcount) p->qtotal += c;
        p->threshold = thr;
        p->prof.host(85, t0);
        p->stored = false;
        p->store_bytes = p->k_entries = p->j_entries = 0;
        if (p->batched) {
            auto t1 = Profiler::clock::now();
            p->build_batch_tensors(st);
            p->prof.host(87, t1);
            p->have_geometry = true;
            return AIMDHF_OK;
        }
        bool store = false;
        if (p->jk_mode != 1 && p->nao < (1 << 14)) {
            long long limit = p->store_limit;
            if (limit <= 0) {
                size_t fr = 0, tot = 0;
                HIPCHECK(hipMemGetInfo(&fr, &tot));
                limit = (long long)(fr / 2);
            }
            store = p->jk_mode == 2 || p->store_estimate() <= limit;
        }
        if (!store) {
            t0 = Profiler::clock::now();
            p->build_owner_lists(thr, true, st);
            p->prof.host(86, t0);
        } else {
            auto t1 = Profiler::clock::now();
            p->build_store(st);
            p->prof.host(87, t1);
            p->stored = true;
        }
        p->have_geometry = true;
    } catch (const HipError& e) {
        p->have_geometry = false;
        return e.code;
    }
    return AIMDHF_OK;
}

int aimdhf_schwarz_bounds(aimdhf_plan* p, double* out, int64_t cap) {
    if (!p || !out) return fail(AIMDHF_E_ARG, "NULL argument");
    if (!p->have_geometry) return fail(AIMDHF_E_STATE, "no geometry set");
    if (cap < p->nbpair) return fail(AIMDHF_E_ARG, "output too small");
    try {
        HIPCHECK(hipMemcpy(out, p->d_qblock.p, sizeof(double) * p->nbpair, hipMemcpyDeviceToHost));
    } catch (const HipError& e) {
        return e.code;
    }
    return AIMDHF_OK;
}

int aimdhf_quartet_count(aimdhf_plan* p, int64_t* count) {
    if (!p || !count) return fail(AIMDHF_E_ARG, "NULL argument");
    if (!p->have_geometry) return fail(AIMDHF_E_STATE, "no geometry set");
    *count = p->qtotal;
    return AIMDHF_OK;
}

int aimdhf_quartets(aimdhf_plan* p, int64_t* out, int64_t cap, int64_t* count) {
    if (!p || !count) return fail(AIMDHF_E_ARG, "NULL argument");
    if (!p->have_geometry) return fail(AIMDHF_E_STATE, "no geometry set");
    *count = p->qtotal;
    if (!out || cap <= 0) return AIMDHF_OK;
    try {
        long long n = std::min<long long>(cap, p->qtotal);
        std::vector<int2> tmp(n);
        if (n) HIPCHECK(hipMemcpy(tmp.data(), p->d_list.p, sizeof(int2) * n, hipMemcpyDeviceToHost));
        for (long long i = 0; i < n; ++i) {
            int a = tmp[i].x, b = tmp[i].y;
            out[2 * i] = std::max(a, b);
            out[2 * i + 1] = std::min(a, b);
        }
    } catch (const HipError& e) {
        return e.code;
    }
    return AIMDHF_OK;
}

int aimdhf_grad_quartets(aimdhf_plan* p, int64_t* out, int64_t cap, int64_t* count) {
    if (!p || !count) return fail(AIMDHF_E_ARG, "NULL argument");
    if (!p->have_geometry) return fail(AIMDHF_E_STATE, "no geometry set");
    try {
        if (p->generic) {
            std::vector<int2> lst(p->gen_gq);
            if (p->gen_gq)
                HIPCHECK(hipMemcpy(lst.data(), p->d_glist.p, sizeof(int2) * p->gen_gq, hipMemcpyDeviceToHost));
            for (long long i = 0; out && i < std::min<long long>(cap, p->gen_gq); ++i) {
                out[2 * i] = lst[i].x;
                out[2 * i + 1] = lst[i].y;
            }
            *count = p->gen_gq;
            return AIMDHF_OK;
        }
        long long tot = 0;
        for (long long c : p->gcount) tot += c;
        std::vector<int2> lst(tot);
        std::vector<unsigned char> keep(tot);
        if (tot) {
            HIPCHECK(hipMemcpy(lst.data(), p->d_glist.p, sizeof(int2) * tot, hipMemcpyDeviceToHost));
            HIPCHECK(hipMemcpy(keep.data(), p->d_gkeep.p, tot, hipMemcpyDeviceToHost));
        }
        long long n = 0;
        for (long long i = 0; i < tot; ++i) {
            if (!keep[i]) continue;
            if (out && n < cap) {
                out[2 * n] = std::max(lst[i].x, lst[i].y);
                out[2 * n + 1] = std::min(lst[i].x, lst[i].y);
            }
            ++n;
        }
        *count = n;
    } catch (const HipError& e) {
        return e.code;
    }
    return AIMDHF_OK;
}

// J and K from the integral store; delta: dens is a density difference and
// units below dthr (largest Schwarz product times max |dens|) are skipped
static int jk_from_store(aimdhf_plan* p, int nset, const double* dens, double* J, double* K, bool delta, double dthr,
                         hipStream_t st) {
    try {
        const size_t nn = (size_t)p->nao * p->nao;
        p->d_dsym.ensure(sizeof(double) * nset * nn);
        if (delta) p->ensure_unit_q(st);
        p->prof.time(45, st, [&] {
            aimdhf::launch_symmetrize(dens, nset, p->nao, p->d_dsym.as<double>(), st);
            if (delta) aimdhf::launch_absmax(p->d_dsym.as<double>(), nset * nn, p->d_dmax.as<double>(), st);
        });
        // the 15 J and K launches are independent (disjoint outputs): spread over the fan's streams
        p->fan.fork(st);
        int i = 0;
        aimdhf::KSArgs ka = p->ks_args(nset, p->d_dsym.as<double>(), K);
        aimdhf::JSArgs ja = p->js_args(nset, p->d_dsym.as<double>(), J);
        ka.delta = ja.delta = delta ? 1 : 0;
        ka.dthr = ja.dthr = dthr;
        ka.d32 = ja.d32 = p->delta_fp32_max;
        p->last_ka = ka;
        p->last_ja = ja;
        p->have_last_store = true;
        for (int t = 8; t >= 0; --t) {
            int nout = p->out_off[t + 1] - p->out_off[t];
            if (nout <= 0) continue;
            hipStream_t s2 = p->fan.on(i++);
            p->prof.time(55 + t, s2, [&] { aimdhf::launch_kstore(t, ka, nout, s2); });
            p->prof.time(93, s2, [&] { aimdhf::launch_kreduce(t, ka, nout, s2); });
        }
        for (int c = 5; c >= 0; --c) {
            if (!p->ncls[c]) continue;
            ja.cls_off = p->cls_pair_off[c];
            hipStream_t s2 = p->fan.on(i++);
            p->prof.time(49 + c, s2, [&] { aimdhf::launch_jstore(c, ja, p->ncls[c], s2); });
            p->prof.time(94, s2, [&] { aimdhf::launch_jreduce(c, ja, p->ncls[c], s2); });
        }
        p->fan.join();
        HIPCHECK(hipGetLastError());
    } catch (const HipError& e) {
        return e.code;
    }
    return AIMDHF_OK;
}

int aimdhf_jk_delta(aimdhf_plan* p, int nset, const double* ddens, double* dJ, double* dK, double threshold,
                    void* stream) {
    if (!p || !ddens || !dJ || !dK) return fail(AIMDHF_E_ARG, "NULL argument");
    if (nset < 1 || nset > 2) return fail(AIMDHF_E_ARG, "nset must be 1 or 2");
    if (!(threshold >= 0.0) || !std::isfinite(threshold)) return fail(AIMDHF_E_ARG, "threshold must be finite and >= 0");
    if (!p->have_geometry) return fail(AIMDHF_E_STATE, "no geometry set");
    if (p->batched || !p->stored) return aimdhf_jk(p, nset, ddens, dJ, dK, stream);
    return jk_from_store(p, nset, ddens, dJ, dK, true, threshold, (hipStream_t)stream);
}

// J and K of every replica of a batch from its dense tensor; dens, J, K in the
// padded layout (nrep, nset, ld, ld)
static int jk_batch(aimdhf_plan* p, int nset, const double* dens, double* J, double* K, hipStream_t st) {
    try {
        const size_t n = (size_t)nset * p->mat_elems();
        p->d_dsym.ensure(sizeof(double) * n);
        p->prof.time(45, st, [&] {
            aimdhf::launch_symmetrize(dens, p->nrep * nset, p->bld, p->d_dsym.as<double>(), st);
            aimdhf::launch_zero_d(J, n, st);
            aimdhf::launch_zero_d(K, n, st);
        });
        p->prof.time(49, st, [&] {
            aimdhf::launch_batch_jk(p->batch_args(), nset, p->d_dsym.as<double>(), J, K, p->rep_item[p->nrep],
                                    p->batch_jk_lanes, st);
        });
        HIPCHECK(hipGetLastError());
    } catch (const HipError& e) {
        return e.code;
    }
    return AIMDHF_OK;
}

int aimdhf_jk(aimdhf_plan* p, int nset, const double* dens, double* J, double* K, void* stream) {
    if (!p || !dens || !J || !K) return fail(AIMDHF_E_ARG, "NULL argument");
    if (nset < 1 || nset > 2) return fail(AIMDHF_E_ARG, "nset must be 1 or 2");
    if (!p->have_geometry) return fail(AIMDHF_E_STATE, "no geometry set");
    hipStream_t st = (hipStream_t)stream;
    if (p->batched) return jk_batch(p, nset, dens, J, K, st);
    if (p->generic) return jk_generic(p, nset, dens, J, K, st);
    if (p->stored) return jk_from_store(p, nset, dens, J, K, false, 0.0, st);
    try {
        const size_t nn = (size_t)p->nao * p->nao;
        p->d_dsym.ensure(sizeof(double) * nset * nn);
        p->d_jpart.ensure(sizeof(double) * 6 * (size_t)p->npair * nset * 36);
        p->d_kpart.ensure(sizeof(double) * 9 * nset * nn);
        p->prof.time(45, st, [&] {
            aimdhf::launch_symmetrize(dens, nset, p->nao, p->d_dsym.as<double>(), st);
            aimdhf::launch_zero_d(p->d_kpart.as<double>(), 9 * nset * nn, st);
        });
        // J-engine: Hermite densities of the ket streams, then one kernel per (bra class, ket class)
        aimdhf::DhArgs dh;
        dh.nao = p->nao;
        dh.nset = nset;
        long long ncls_entries[6];
        for (int c = 0; c < 6; ++c) {
            dh.pbase[c] = p->pbase[c];
            dh.dh_base[c] = p->dh_base2[c];
            ncls_entries[c] = p->pbase[c + 1] - p->pbase[c];
        }
        dh.krec = p->d_krec.as<long long>();
        dh.kpair = p->d_kpair.as<int>();
        dh.pair_a = p->d_pair_a.as<int>();
        dh.pair_b = p->d_pair_b.as<int>();
        dh.shell_ao = p->d_shell_ao.as<int>();
        dh.recs = p->d_recs.as<double>();
        dh.dens = p->d_dsym.as<double>();
        dh.dh = p->d_dh.as<double>();
        p->prof.time(46, st, [&] { aimdhf::launch_dh(dh, ncls_entries, st); });
        aimdhf::JEArgs je;
        je.npair = p->npair;
        je.thr = p->threshold;
        je.cls_pairs = p->d_cls_pairs.as<int>();
        je.cls_pair_off = p->d_cls_pair_off.as<int>();
        for (int c = 0; c < 6; ++c) je.ncls_host[c] = p->ncls[c];
        je.qscr = p->d_qscr.as<double>();
        je.ket_q = p->d_ket_q.as<double>();
        for (int c = 0; c < 7; ++c) je.ket_off[c] = p->ket_off[c];
        je.kpre = p->d_kpre.as<long long>();
        je.krec = p->d_krec.as<long long>();
        je.dh = p->d_dh.as<double>();
        for (int c = 0; c < 6; ++c) je.dh_base[c] = p->dh_base2[c];
        je.recs = p->d_recs.as<double>();
        je.rec_off = p->d_rec_off.as<long long>();
        je.npp = p->d_npp.as<int>();
        je.boys = p->d_boys.as<double>();
        je.jpart = p->d_jpart.as<double>();
        for (int c = 0; c < 6; ++c)
            if (p->ncls[c] > 0) p->prof.time(c, st, [&] { aimdhf::launch_jeng(c, je, nset, p->nket, st); });
        aimdhf::JAsmArgs ja;
        ja.npair = p->npair;
        ja.nao = p->nao;
        ja.nset = nset;
        for (int c = 0; c < 6; ++c) ja.cq_valid[c] = p->nket[c] > 0;
        ja.pair_a = p->d_pair_a.as<int>();
        ja.pair_b = p->d_pair_b.as<int>();
        ja.shell_ao = p->d_shell_ao.as<int>();
        ja.jpart = p->d_jpart.as<double>();
        ja.J = J;
        p->prof.time(45, st, [&] { aimdhf::launch_j_assemble(ja, st); });
        // exchange owners: one kernel per (l_X, l_Y, l_J, l_L)
        aimdhf::KOArgs ko;
        ko.nao = p->nao;
        ko.nset = nset;
        ko.thr = p->threshold;
        for (int t = 0; t < 10; ++t) ko.out_off[t] = p->out_off[t];
        ko.out_x = p->d_out_x.as<int>();
        ko.out_y = p->d_out_y.as<int>();
        ko.part_off = p->d_part_off.as<int>();
        ko.part_shell = p->d_part_shell.as<int>();
        ko.part_pair = p->d_part_pair.as<int>();
        ko.part_q = p->d_part_q.as<double>();
        ko.pair_a = p->d_pair_a.as<int>();
        ko.rec_off = p->d_rec_off.as<long long>();
        ko.recs = p->d_recs.as<double>();
        ko.npp = p->d_npp.as<int>();
        ko.shell_ao = p->d_shell_ao.as<int>();
        ko.boys = p->d_boys.as<double>();
        ko.dens = p->d_dsym.as<double>();
        ko.kpart = p->d_kpart.as<double>();
        for (int t = 0; t < 9; ++t) {
            int nout = p->out_off[t + 1] - p->out_off[t];
            if (nout > 0) p->prof.time(6 + t, st, [&] { aimdhf::launch_kown(t, ko, nout, p->has_l, st); });
        }
        p->prof.time(45, st, [&] {
            aimdhf::launch_k_assemble(p->d_kpart.as<double>(), nset, p->nao, p->d_ao_shell.as<int>(), K, st);
        });
        HIPCHECK(hipGetLastError());
    } catch (const HipError& e) {
        return e.code;
    }
    return AIMDHF_OK;
}

int aimdhf_one_electron(aimdhf_plan* p, double* S, double* T, double* V, double* M, void* stream) {
    if (!p || !S || !T || !V || !M) return fail(AIMDHF_E_ARG, "NULL argument");
    if (!p->have_geometry) return fail(AIMDHF_E_STATE, "no geometry set");
    hipStream_t st = (hipStream_t)stream;
    try {
        if (p->batched) {
            // padded layout: only the replicas' own blocks are written
            const size_t n = p->mat_elems();
            aimdhf::launch_zero_d(S, n, st);
            aimdhf::launch_zero_d(T, n, st);
            aimdhf::launch_zero_d(V, n, st);
            aimdhf::launch_zero_d(M, 3 * n, st);
        }
        if (p->generic) {
            aimdhf::GenArgs g = p->gen_args();
            g.S = S;
            g.T = T;
            g.V = V;
            g.M = M;
            p->prof.time(43, st, [&] { aimdhf::launch_gen_oneel(g, p->npair, st); });
        } else {
            p->prof.time(43, st, [&] { aimdhf::launch_oneel(p->oneel_args(), S, T, M, V, st); });
        }
        HIPCHECK(hipGetLastError());
    } catch (const HipError& e) {
        return e.code;
    }
    return AIMDHF_OK;
}

int aimdhf_grad_one(aimdhf_plan* p, const double* D, const double* W, double* grad, void* stream) {
    if (!p || !D || !W || !grad) return fail(AIMDHF_E_ARG, "NULL argument");
    if (!p->have_geometry) return fail(AIMDHF_E_STATE, "no geometry set");
    hipStream_t st = (hipStream_t)stream;
    try {
        if (p->generic) {
            grad_one_generic(p, D, W, 7, grad, st);
            return AIMDHF_OK;
        }
        const size_t nn = p->mat_elems();
        const int nmat = p->batched ? p->nrep : 1, ld = p->batched ? p->bld : p->nao;
        p->d_gden.ensure(sizeof(double) * 2 * nn);
        double* Ds = p->d_gden.as<double>();
        double* Wn = Ds + nn;
        aimdhf::launch_symmetrize(D, nmat, ld, Ds, st);
        aimdhf::launch_symmetrize(W, nmat, ld, Wn, st);
        aimdhf::launch_negate(Wn, nn, Wn, st);
        p->d_gfx.ensure(sizeof(long long) * 6 * p->natm);
        aimdhf::launch_zero64(p->d_gfx.as<long long>(), 6 * p->natm, st);
        p->prof.time(44, st, [&] { aimdhf::launch_grad_one(p->oneel_args(), Ds, Wn, p->d_gfx.as<long long>(), st); });
        aimdhf::launch_fx_to_double(p->d_gfx.as<long long>(), 3 * p->natm, grad, st);
        HIPCHECK(hipGetLastError());
    } catch (const HipError& e) {
        return e.code;
    }
    return AIMDHF_OK;
}

int aimdhf_grad_one_noalloc(aimdhf_plan* p, const double* D, const double* W, double* grad, void* stream) {
    if (!p || !D || !W || !grad) return fail(AIMDHF_E_ARG, "NULL argument");
    if (p->batched) return fail(AIMDHF_E_UNSUPPORTED, "aimdhf_grad_one_noalloc takes single-molecule plans");
    if (p->generic) return fail(AIMDHF_E_UNSUPPORTED, "aimdhf_grad_one_noalloc takes bases with l <= 2");
    const size_t nn = (size_t)p->nao * p->nao;
    const size_t mats = (sizeof(double) * 2 * nn + 255) & ~(size_t)255;
    const size_t need = mats + sizeof(long long) * 6 * (size_t)p->natm;
    if (!p->noalloc_ptr || need > p->noalloc_bytes)
        return fail(AIMDHF_E_ARG, "aimdhf_grad_one_noalloc: nao = " + std::to_string(p->nao) + ", natm = " +
                                      std::to_string(p->natm) + " need " + std::to_string(need) +
                                      " bytes of scratch, more than its fixed " + std::to_string(p->noalloc_bytes) +
                                      " (16 nao^2 + 48 natm bytes must fit; use aimdhf_grad_one)");
    if (!p->have_geometry) return fail(AIMDHF_E_STATE, "no geometry set");
    hipStream_t st = (hipStream_t)stream;
    char* base = (char*)p->noalloc_ptr;
    double* Ds = (double*)base;
    double* Wn = Ds + nn;
    long long* fx = (long long*)(base + mats);
    aimdhf::launch_symmetrize(D, 1, p->nao, Ds, st);
    aimdhf::launch_symmetrize(W, 1, p->nao, Wn, st);
    aimdhf::launch_negate(Wn, nn, Wn, st);
    aimdhf::launch_zero64(fx, 6 * (size_t)p->natm, st);     // a fill kernel: the array keeps the last call's sums
    aimdhf::launch_grad_one(p->oneel_args(), Ds, Wn, fx, st);
    aimdhf::launch_fx_to_double(fx, 3 * p->natm, grad, st);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(AIMDHF_E_HIP, std::string("grad_one_noalloc: ") + hipGetErrorString(e));
    return AIMDHF_OK;
}

int aimdhf_grad_one_terms(aimdhf_plan* p, const double* D, const double* W, double* g_kin, double* g_nuc,
                          double* g_ovl, void* stream) {
    if (!p || !D || !W || !g_kin || !g_nuc || !g_ovl) return fail(AIMDHF_E_ARG, "NULL argument");
    if (!p->have_geometry) return fail(AIMDHF_E_STATE, "no geometry set");
    hipStream_t st = (hipStream_t)stream;
    try {
        if (p->generic) {
            grad_one_generic(p, D, W, 1, g_kin, st);
            grad_one_generic(p, D, W, 2, g_nuc, st);
            grad_one_generic(p, D, W, 4, g_ovl, st);
            return AIMDHF_OK;
        }
        const size_t nn = p->mat_elems();
        const int nmat = p->batched ? p->nrep : 1, ld = p->batched ? p->bld : p->nao;
        p->d_gden.ensure(sizeof(double) * 3 * nn);
        double* Ds = p->d_gden.as<double>();
        double* Wn = Ds + nn;
        double* Z = Wn + nn;
        aimdhf::launch_symmetrize(D, nmat, ld, Ds, st);
        aimdhf::launch_symmetrize(W, nmat, ld, Wn, st);
        aimdhf::launch_negate(Wn, nn, Wn, st);
        aimdhf::launch_zero_d(Z, nn, st);
        p->d_gfx.ensure(sizeof(long long) * 18 * p->natm);
        long long* fx = p->d_gfx.as<long long>();
        aimdhf::launch_zero64(fx, 18 * p->natm, st);
        aimdhf::OneElArgs g = p->oneel_args();
        // kinetic: overlap/kinetic part with W = 0; nuclear attraction alone; overlap:
        // overlap/kinetic part with D = 0 (the -W S term)
        g.parts = 1;
        aimdhf::launch_grad_one(g, Ds, Z, fx, st);
        aimdhf::launch_grad_one(g, Z, Wn, fx + 12 * p->natm, st);
        g.parts = 2;
        aimdhf::launch_grad_one(g, Ds, Z, fx + 6 * p->natm, st);
        aimdhf::launch_fx_to_double(fx, 3 * p->natm, g_kin, st);
        aimdhf::launch_fx_to_double(fx + 6 * p->natm, 3 * p->natm, g_nuc, st);
        aimdhf::launch_fx_to_double(fx + 12 * p->natm, 3 * p->natm, g_ovl, st);
        HIPCHECK(hipGetLastError());
    } catch (const HipError& e) {
        return e.code;
    }
    return AIMDHF_OK;
}

int aimdhf_grad_two(aimdhf_plan* p, const double* Dc, int nx, const double* Dx, double kfac, double thr,
                    double* grad, void* stream) {
    if (!p || !Dc || !Dx || !grad) return fail(AIMDHF_E_ARG, "NULL argument");
    if (nx < 1 || nx > 2) return fail(AIMDHF_E_ARG, "nx must be 1 (RHF) or 2 (UHF)");
    if (!(thr >= 0.0)) return fail(AIMDHF_E_ARG, "gradient screening threshold must be >= 0");
    if (!p->have_geometry) return fail(AIMDHF_E_STATE, "no geometry set");
    hipStream_t st = (hipStream_t)stream;
    if (p->generic) return grad_two_generic(p, Dc, nx, Dx, kfac, thr, grad, st);
    try {
        const size_t nn = p->mat_elems();
        const int nmat = p->batched ? p->nrep : 1, ld = p->batched ? p->bld : p->nao;
        p->d_gden.ensure(sizeof(double) * (1 + nx) * nn);
        double* dc = p->d_gden.as<double>();
        double* dx = dc + nn;
        aimdhf::launch_symmetrize(Dc, nmat, ld, dc, st);
        aimdhf::launch_symmetrize(Dx, nmat * nx, ld, dx, st);
        // candidate quartets: q_P q_Q >= thr / max|Gamma| over every possible element
        double mc = aimdhf::device_absmax(dc, nn, p->d_scratch.as<double>(), st);
        double mx = aimdhf::device_absmax(dx, nx * nn, p->d_scratch.as<double>(), st);
        if (mc < 0.0 || mx < 0.0) return fail(AIMDHF_E_HIP, "max reduction failed");
        double gbound = 4.0 * (mc * mc + std::fabs(kfac) * nx * mx * mx);
        double lthr = gbound > 0.0 ? thr / gbound : 0.0;
        auto t0 = Profiler::clock::now();
        p->build_list(lthr, 0.0, p->d_glist, p->gcount, p->g64count, st);
        p->prof.host(88, t0);
        p->d_gfx.ensure(sizeof(long long) * 6 * p->natm);
        aimdhf::launch_zero64(p->d_gfx.as<long long>(), 6 * p->natm, st);
        aimdhf::GradArgs a;
        a.nao = p->nao;
        a.natm = p->natm;
        a.nx = nx;
        a.kfac = kfac;
        a.thr = thr;
        a.pair_a = p->d_pair_a.as<int>();
        a.pair_b = p->d_pair_b.as<int>();
        a.shell_ao = p->d_shell_ao.as<int>();
        a.shell_atom = p->d_shell_atom.as<int>();
        a.npp = p->d_npp.as<int>();
        a.rec_off = p->d_rec_off.as<long long>();
        a.recs = p->d_recs.as<double>();
        a.boys = p->d_boys.as<double>();
        a.pos = p->d_pos.as<double>();
        a.qscr = p->d_qscr.as<double>();
        a.dc = dc;
        a.dx = dx;
        a.npair = p->npair;
        a.ld = ld;
        a.shell_rep = p->batched ? p->d_shell_rep.as<int>() : nullptr;
        a.shell_first = p->batched ? p->d_shell_first.as<int>() : nullptr;
        a.shell_pair0 = p->batched ? p->d_shell_pair0.as<int>() : nullptr;
        p->d_gdmax.ensure(sizeof(double) * (1 + nx) * std::max(p->npair, 1));
        a.dmc = p->d_gdmax.as<double>();
        a.dmx = a.dmc + p->npair;
        aimdhf::launch_pair_dmax(a, p->d_shell_l.as<int>(), p->d_gdmax.as<double>(), p->d_gdmax.as<double>() + p->npair,
                                 st);
        a.out = p->d_gfx.as<long long>();
        long long gtot = 0;
        for (int k = 0; k < NQCLASS; ++k) gtot += p->gcount[k];
        p->d_gkeep.ensure(std::max<long long>(gtot, 1));
        p->d_gsel.ensure(sizeof(int) * 2 * std::max<long long>(gtot, 1));
        p->d_nsel.ensure(sizeof(int) * 2 * NQCLASS);
        p->d_gblk.ensure(sizeof(int) * 4 * (gtot / 256 + 2 * NQCLASS));
        a.fp32_switch = p->precision == 1 ? p->grad_switch : 0.0;
        long long off = 0;
        p->fan.fork(st);
        for (int k = 0, i = 0; k < NQCLASS; ++k) {
            const int2* lst = p->d_glist.as<int2>() + off;
            hipStream_t s2 = p->gcount[k] > 0 ? p->fan.on(i++) : st;
            a.keep = p->d_gkeep.as<unsigned char>() + off;
            a.sel64 = p->d_gsel.as<int>() + off;
            a.sel32 = p->d_gsel.as<int>() + gtot + off;
            a.nsel = p->d_nsel.as<int>() + 2 * k;
            a.bcnt = p->d_gblk.as<int>() + 4 * (off / 256 + k);
            a.boff = a.bcnt + 2 * ((p->gcount[k] + 255) / 256);
            if (p->gcount[k] > 0) {
                p->prof.time(64 + k, s2, [&] { aimdhf::launch_grad_keep(k, a, lst, p->gcount[k], s2); });
                p->prof.time(21 + k, s2, [&] { aimdhf::launch_grad_one_class(k, a, lst, p->gcount[k], s2); });
            }
            off += p->gcount[k];
        }
        p->fan.join();
        aimdhf::launch_fx_to_double(p->d_gfx.as<long long>(), 3 * p->natm, grad, st);
        HIPCHECK(hipGetLastError());
    } catch (const HipError& e) {
        return e.code;
    }
    return AIMDHF_OK;
}

int aimdhf_set_jk_mode(aimdhf_plan* p, int mode, int64_t max_store_bytes) {
    if (!p) return fail(AIMDHF_E_ARG, "NULL plan");
    if (mode < 0 || mode > 2) return fail(AIMDHF_E_ARG, "jk mode must be 0 (auto), 1 (direct) or 2 (store)");
    if (max_store_bytes < 0) return fail(AIMDHF_E_ARG, "max_store_bytes must be >= 0");
    p->jk_mode = mode;
    p->store_limit = max_store_bytes;
    return AIMDHF_OK;
}

int aimdhf_tc2_step(int n, const double* X, const double* X2, double* Xout, double nocc, double freeze, void* stream) {
    if (n <= 0 || !X || !X2 || !Xout) return fail(AIMDHF_E_ARG, "bad tc2 arguments");
    aimdhf::launch_tc2_step(n, X, X2, Xout, nocc, freeze, (hipStream_t)stream);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(AIMDHF_E_HIP, std::string("tc2 step: ") + hipGetErrorString(e));
    return AIMDHF_OK;
}

int aimdhf_batch_set_jk_lanes(aimdhf_plan* p, int lanes) {
    if (!p) return fail(AIMDHF_E_ARG, "NULL plan");
    if (!p->batched) return fail(AIMDHF_E_STATE, "not a batched plan");
    if (lanes != 16 && lanes != 32 && lanes != 64) return fail(AIMDHF_E_ARG, "lanes must be 16, 32 or 64");
    p->batch_jk_lanes = lanes;
    return AIMDHF_OK;
}

int aimdhf_batch_purify(int nmat, int ld, const double* F, const double* nocc, double* P, int steps, double tol,
                        double* err, void* stream) {
    if (nmat < 0 || ld <= 0 || ld > 64 || steps < 0 || !(tol >= 0.0)) return fail(AIMDHF_E_ARG, "bad purify arguments");
    if (nmat == 0) return AIMDHF_OK;
    if (!F || !nocc || !P || !err) return fail(AIMDHF_E_ARG, "NULL argument");
    aimdhf::launch_batch_purify(nmat, ld, F, nocc, P, steps, tol, err, (hipStream_t)stream);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(AIMDHF_E_HIP, std::string("batch purify: ") + hipGetErrorString(e));
    return AIMDHF_OK;
}

int aimdhf_batch_diis_solve(int nrep, int m, const double* G, const double* r, double* c, void* stream) {
    if (nrep < 0 || m <= 0 || m > 12) return fail(AIMDHF_E_ARG, "bad DIIS solve arguments (1 <= m <= 12)");
    if (nrep == 0) return AIMDHF_OK;
    if (!G || !r || !c) return fail(AIMDHF_E_ARG, "NULL argument");
    aimdhf::launch_batch_diis_solve(nrep, m, G, r, c, (hipStream_t)stream);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(AIMDHF_E_HIP, std::string("batch DIIS solve: ") + hipGetErrorString(e));
    return AIMDHF_OK;
}

int aimdhf_tc2_init(int n, const double* F, double* X0, void* stream) {
    if (n <= 0 || !F || !X0) return fail(AIMDHF_E_ARG, "bad tc2 arguments");
    aimdhf::launch_tc2_init(n, F, X0, (hipStream_t)stream);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(AIMDHF_E_HIP, std::string("tc2 init: ") + hipGetErrorString(e));
    return AIMDHF_OK;
}

int aimdhf_tc2_error(int n, const double* X, double* err, void* stream) {
    if (n <= 0 || !X || !err) return fail(AIMDHF_E_ARG, "bad tc2 arguments");
    aimdhf::launch_tc2_error(n, X, err, (hipStream_t)stream);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(AIMDHF_E_HIP, std::string("tc2 error: ") + hipGetErrorString(e));
    return AIMDHF_OK;
}

int aimdhf_set_streams(aimdhf_plan* p, int concurrent) {
    if (!p) return fail(AIMDHF_E_ARG, "NULL plan");
    p->fan.serial = concurrent == 0;
    return AIMDHF_OK;
}

int aimdhf_set_precision(aimdhf_plan* p, int mode, double jk_switch, double grad_switch, double delta_switch) {
    if (!p) return fail(AIMDHF_E_ARG, "NULL plan");
    if (mode < 0 || mode > 1) return fail(AIMDHF_E_ARG, "precision mode must be 0 (fp64) or 1 (mixed)");
    if (!(jk_switch >= 0.0) || !(grad_switch >= 0.0) || !(delta_switch >= 0.0))
        return fail(AIMDHF_E_ARG, "switch thresholds must be >= 0");
    p->precision = mode;
    p->jk_switch = jk_switch;
    p->grad_switch = grad_switch;
    p->delta_fp32_max = delta_switch;
    return AIMDHF_OK;
}

int aimdhf_set_store_units(aimdhf_plan* p, int items) {
    if (!p) return fail(AIMDHF_E_ARG, "NULL plan");
    if (items < 64 || items > (1 << 16)) return fail(AIMDHF_E_ARG, "store work units take 64 .. 65536 lane items");
    p->unit_items = items;
    return AIMDHF_OK;
}

int aimdhf_jk_info(aimdhf_plan* p, int* stored, int64_t* store_bytes, int64_t* k_entries, int64_t* j_entries,
                   int64_t* gather_bytes) {
    if (!p) return fail(AIMDHF_E_ARG, "NULL plan");
    if (!p->have_geometry) return fail(AIMDHF_E_STATE, "no geometry set");
    if (stored) *stored = p->stored ? 1 : 0;
    if (gather_bytes)
        *gather_bytes = p->stored ? 8 * (p->kdata_total + p->jdata_total) + 4 * (p->kdata32_total + p->jdata32_total) : 0;
    if (store_bytes) *store_bytes = p->store_bytes;
    if (k_entries) *k_entries = p->k_entries;
    if (j_entries) *j_entries = p->j_entries;
    return AIMDHF_OK;
}

int aimdhf_jk_type_info(aimdhf_plan* p, int64_t* k_bytes, int64_t* k_entries, int64_t* j_bytes, int64_t* j_entries,
                        int64_t* unique_bytes) {
    if (!p) return fail(AIMDHF_E_ARG, "NULL plan");
    if (!p->have_geometry) return fail(AIMDHF_E_STATE, "no geometry set");
    try {
        if (unique_bytes)
            for (int k = 0; k < NQCLASS; ++k)
                unique_bytes[k] = p->stored ? 8 * (p->store_base[k + 1] - p->store_base[k]) : 0;
        for (int t = 0; t < 9; ++t) {
            if (k_bytes) k_bytes[t] = 0;
            if (k_entries) k_entries[t] = 0;
        }
        for (int c = 0; c < 6; ++c) {
            if (j_bytes) j_bytes[c] = 0;
            if (j_entries) j_entries[c] = 0;
        }
        if (!p->stored || p->batched) return AIMDHF_OK;
        // K segments are (output block, subclass) with the output blocks grouped by type
        auto at = [&](const DevBuf& b, long long i) {
            long long v = 0;
            HIPCHECK(hipMemcpy(&v, b.as<long long>() + i, sizeof(long long), hipMemcpyDeviceToHost));
            return v;
        };
        for (int t = 0; t < 9; ++t) {
            const long long a = 9LL * p->out_off[t], b = 9LL * p->out_off[t + 1];
            if (k_bytes)
                k_bytes[t] = 8 * (at(p->d_kdoff, b) - at(p->d_kdoff, a)) + 4 * (at(p->d_kdoff32, b) - at(p->d_kdoff32, a));
            if (k_entries) k_entries[t] = at(p->d_kseg, b) - at(p->d_kseg, a);
        }
        // J segments are indexed by pair; each pair's six segments go to its class
        const size_t njs = (size_t)p->npair * 6 + 1;
        std::vector<long long> seg(njs), d64(njs), d32(njs);
        HIPCHECK(hipMemcpy(seg.data(), p->d_jseg.p, sizeof(long long) * njs, hipMemcpyDeviceToHost));
        HIPCHECK(hipMemcpy(d64.data(), p->d_jdoff.p, sizeof(long long) * njs, hipMemcpyDeviceToHost));
        HIPCHECK(hipMemcpy(d32.data(), p->d_jdoff32.p, sizeof(long long) * njs, hipMemcpyDeviceToHost));
        for (int P = 0; P < p->npair; ++P) {
            const int c = p->pair_cls[P];
            const size_t a = (size_t)P * 6, b = a + 6;
            if (j_bytes) j_bytes[c] += 8 * (d64[b] - d64[a]) + 4 * (d32[b] - d32[a]);
            if (j_entries) j_entries[c] += seg[b] - seg[a];
        }
    } catch (const HipError& e) {
        return e.code;
    }
    return AIMDHF_OK;
}

int aimdhf_profile_enable(aimdhf_plan* p, int on) {
    if (!p) return fail(AIMDHF_E_ARG, "NULL plan");
    try {
        p->prof.collect();
    } catch (const HipError& e) {
        return e.code;
    }
    p->prof.on = on != 0;
    return AIMDHF_OK;
}

int aimdhf_profile_replay(aimdhf_plan* p, int kind, int reps, void* stream, double* avg_ms) {
    if (!p || !avg_ms) return fail(AIMDHF_E_ARG, "NULL argument");
    if (reps < 1) return fail(AIMDHF_E_ARG, "reps must be >= 1");
    const bool ks = kind >= 55 && kind < 64, js = kind >= 49 && kind < 55;
    if (!ks && !js) return fail(AIMDHF_E_UNSUPPORTED, "replay covers the store-mode digestion kinds 49..63");
    if (!p->have_geometry || !p->stored || !p->have_last_store)
        return fail(AIMDHF_E_STATE, "replay needs a store-mode J/K build at the current geometry");
    hipStream_t st = (hipStream_t)stream;
    hipEvent_t a = nullptr, b = nullptr;
    try {
        HIPCHECK(hipEventCreate(&a));
        HIPCHECK(hipEventCreate(&b));
        aimdhf::JSArgs ja = p->last_ja;
        const int t = kind - 55, c = kind - 49;
        if (js) ja.cls_off = p->cls_pair_off[c];
        const int nout = ks ? p->out_off[t + 1] - p->out_off[t] : p->ncls[c];
        // the digestion writes only its unit partials: replays leave J and K alone
        HIPCHECK(hipEventRecord(a, st));
        for (int r = 0; r < reps; ++r) {
            if (ks)
                aimdhf::launch_kstore(t, p->last_ka, nout, st);
            else
                aimdhf::launch_jstore(c, ja, nout, st);
        }
        HIPCHECK(hipEventRecord(b, st));
        HIPCHECK(hipEventSynchronize(b));
        float ms = 0.0f;
        HIPCHECK(hipEventElapsedTime(&ms, a, b));
        *avg_ms = ms / reps;
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
    } catch (const HipError& e) {
        if (a) (void)hipEventDestroy(a);
        if (b) (void)hipEventDestroy(b);
        return e.code;
    }
    return AIMDHF_OK;
}

int aimdhf_profile_read(aimdhf_plan* p, double* ms, int64_t* launches, int nkind, int reset) {
    if (!p || !ms || !launches) return fail(AIMDHF_E_ARG, "NULL argument");
    try {
        p->prof.collect();
        for (int i = 0; i < nkind && i < Profiler::NKIND; ++i) {
            ms[i] = p->prof.ms[i];
            launches[i] = p->prof.n[i];
        }
        if (reset) p->prof.reset();
    } catch (const HipError& e) {
        return e.code;
    }
    return AIMDHF_OK;
}

int aimdhf_class_counts(aimdhf_plan* p, int64_t* energy_counts, int64_t* gradient_counts) {
    if (!p) return fail(AIMDHF_E_ARG, "NULL plan");
    for (int k = 0; k < NQCLASS; ++k) {
        if (energy_counts) energy_counts[k] = k < (int)p->qcount.size() ? p->qcount[k] : 0;
        if (gradient_counts) gradient_counts[k] = k < (int)p->gcount.size() ? p->gcount[k] : 0;
    }
    return AIMDHF_OK;
}

int aimdhf_pair_npp(aimdhf_plan* p, int32_t* npp_host, int64_t cap) {
    if (!p || !npp_host) return fail(AIMDHF_E_ARG, "NULL argument");
    if (!p->have_geometry) return fail(AIMDHF_E_STATE, "no geometry set");
    if (cap < p->npair) return fail(AIMDHF_E_ARG, "output too small");
    if (p->generic) {
        for (int P = 0; P < p->npair; ++P) npp_host[P] = p->gen_npp[P];
        return AIMDHF_OK;
    }
    try {
        HIPCHECK(hipMemcpy(npp_host, p->d_npp.p, sizeof(int) * p->npair, hipMemcpyDeviceToHost));
    } catch (const HipError& e) {
        return e.code;
    }
    return AIMDHF_OK;
}

}  // extern "C"
