// Dispatch tables over the per-class / per-type kernel objects (one object
// per template configuration, see the Makefile).
#include "aimdhf_internal.h"

namespace aimdhf {

#define DECLG(k) void launch_grad_class_##k(const GradArgs&, const int2*, long long, hipStream_t); \
    void launch_grad_keep_##k(const GradArgs&, const int2*, long long, hipStream_t);
DECLG(0) DECLG(1) DECLG(2) DECLG(3) DECLG(4) DECLG(5) DECLG(6) DECLG(7) DECLG(8) DECLG(9) DECLG(10)
DECLG(11) DECLG(12) DECLG(13) DECLG(14) DECLG(15) DECLG(16) DECLG(17) DECLG(18) DECLG(19) DECLG(20)
#undef DECLG

void launch_grad_one_class(int k, const GradArgs& a, const int2* list, long long nq, hipStream_t st) {
    typedef void (*Fn)(const GradArgs&, const int2*, long long, hipStream_t);
    static const Fn fns[21] = {
        launch_grad_class_0, launch_grad_class_1, launch_grad_class_2, launch_grad_class_3,
        launch_grad_class_4, launch_grad_class_5, launch_grad_class_6, launch_grad_class_7,
        launch_grad_class_8, launch_grad_class_9, launch_grad_class_10, launch_grad_class_11,
        launch_grad_class_12, launch_grad_class_13, launch_grad_class_14, launch_grad_class_15,
        launch_grad_class_16, launch_grad_class_17, launch_grad_class_18, launch_grad_class_19,
        launch_grad_class_20};
    fns[k](a, list, nq, st);
}

void launch_grad_keep(int k, const GradArgs& a, const int2* list, long long nq, hipStream_t st) {
    typedef void (*Fn)(const GradArgs&, const int2*, long long, hipStream_t);
    static const Fn fns[21] = {
        launch_grad_keep_0, launch_grad_keep_1, launch_grad_keep_2, launch_grad_keep_3,
        launch_grad_keep_4, launch_grad_keep_5, launch_grad_keep_6, launch_grad_keep_7,
        launch_grad_keep_8, launch_grad_keep_9, launch_grad_keep_10, launch_grad_keep_11,
        launch_grad_keep_12, launch_grad_keep_13, launch_grad_keep_14, launch_grad_keep_15,
        launch_grad_keep_16, launch_grad_keep_17, launch_grad_keep_18, launch_grad_keep_19,
        launch_grad_keep_20};
    fns[k](a, list, nq, st);
}

void launch_grad(const GradArgs& a, const int2* const* lists, const long long* counts, hipStream_t st) {
    typedef void (*Fn)(const GradArgs&, const int2*, long long, hipStream_t);
    static const Fn fns[21] = {
        launch_grad_class_0, launch_grad_class_1, launch_grad_class_2, launch_grad_class_3,
        launch_grad_class_4, launch_grad_class_5, launch_grad_class_6, launch_grad_class_7,
        launch_grad_class_8, launch_grad_class_9, launch_grad_class_10, launch_grad_class_11,
        launch_grad_class_12, launch_grad_class_13, launch_grad_class_14, launch_grad_class_15,
        launch_grad_class_16, launch_grad_class_17, launch_grad_class_18, launch_grad_class_19,
        launch_grad_class_20};
    for (int k = 0; k < 21; ++k) fns[k](a, lists[k], counts[k], st);
}

#define DECLJ(k) void launch_jeng_##k(const JEArgs&, int, const int*, hipStream_t);
DECLJ(0) DECLJ(1) DECLJ(2) DECLJ(3) DECLJ(4) DECLJ(5)
#undef DECLJ
#define DECLK(x, y) void launch_kown_##x##_##y(const KOArgs&, int, const int*, hipStream_t);
DECLK(0, 0) DECLK(0, 1) DECLK(0, 2) DECLK(1, 0) DECLK(1, 1) DECLK(1, 2) DECLK(2, 0) DECLK(2, 1) DECLK(2, 2)
#undef DECLK

void launch_jeng(int pclass, const JEArgs& a, int nset, const int* nket_host, hipStream_t st) {
    typedef void (*Fn)(const JEArgs&, int, const int*, hipStream_t);
    static const Fn fns[6] = {launch_jeng_0, launch_jeng_1, launch_jeng_2, launch_jeng_3, launch_jeng_4, launch_jeng_5};
    fns[pclass](a, nset, nket_host, st);
}

void launch_kown(int type, const KOArgs& a, int nout, const int* has_l, hipStream_t st) {
    typedef void (*Fn)(const KOArgs&, int, const int*, hipStream_t);
    static const Fn fns[9] = {launch_kown_0_0, launch_kown_0_1, launch_kown_0_2, launch_kown_1_0, launch_kown_1_1,
                              launch_kown_1_2, launch_kown_2_0, launch_kown_2_1, launch_kown_2_2};
    fns[type](a, nout, has_l, st);
}

#define DECLS(k) void launch_eri_store_##k(const StoreArgs&, const int2*, long long, long long, hipStream_t);
DECLS(0) DECLS(1) DECLS(2) DECLS(3) DECLS(4) DECLS(5) DECLS(6) DECLS(7) DECLS(8) DECLS(9) DECLS(10)
DECLS(11) DECLS(12) DECLS(13) DECLS(14) DECLS(15) DECLS(16) DECLS(17) DECLS(18) DECLS(19) DECLS(20)
#undef DECLS
#define DECLJS(c) void launch_jstore_##c(const JSArgs&, int, hipStream_t); \
    void launch_jreduce_##c(const JSArgs&, int, hipStream_t); \
    void launch_jcopy_##c(const JSArgs&, int, hipStream_t);
DECLJS(0) DECLJS(1) DECLJS(2) DECLJS(3) DECLJS(4) DECLJS(5)
#undef DECLJS
#define DECLKS(x, y) void launch_kstore_##x##_##y(const KSArgs&, int, hipStream_t); \
    void launch_kreduce_##x##_##y(const KSArgs&, int, hipStream_t); \
    void launch_kcopy_##x##_##y(const KSArgs&, int, hipStream_t);
DECLKS(0, 0) DECLKS(0, 1) DECLKS(0, 2) DECLKS(1, 0) DECLKS(1, 1) DECLKS(1, 2) DECLKS(2, 0) DECLKS(2, 1) DECLKS(2, 2)
#undef DECLKS

void launch_eri_store(int k, const StoreArgs& a, const int2* list, long long nq, long long n64, hipStream_t st) {
    typedef void (*Fn)(const StoreArgs&, const int2*, long long, long long, hipStream_t);
    static const Fn fns[21] = {
        launch_eri_store_0, launch_eri_store_1, launch_eri_store_2, launch_eri_store_3, launch_eri_store_4,
        launch_eri_store_5, launch_eri_store_6, launch_eri_store_7, launch_eri_store_8, launch_eri_store_9,
        launch_eri_store_10, launch_eri_store_11, launch_eri_store_12, launch_eri_store_13, launch_eri_store_14,
        launch_eri_store_15, launch_eri_store_16, launch_eri_store_17, launch_eri_store_18, launch_eri_store_19,
        launch_eri_store_20};
    fns[k](a, list, nq, n64, st);
}

void launch_jcopy(int pclass, const JSArgs& a, int npairs, hipStream_t st) {
    typedef void (*Fn)(const JSArgs&, int, hipStream_t);
    static const Fn fns[6] = {launch_jcopy_0, launch_jcopy_1, launch_jcopy_2,
                              launch_jcopy_3, launch_jcopy_4, launch_jcopy_5};
    fns[pclass](a, npairs, st);
}

void launch_kcopy(int type, const KSArgs& a, int nout, hipStream_t st) {
    typedef void (*Fn)(const KSArgs&, int, hipStream_t);
    static const Fn fns[9] = {launch_kcopy_0_0, launch_kcopy_0_1, launch_kcopy_0_2,
                              launch_kcopy_1_0, launch_kcopy_1_1, launch_kcopy_1_2,
                              launch_kcopy_2_0, launch_kcopy_2_1, launch_kcopy_2_2};
    fns[type](a, nout, st);
}

// doubles per stored block of quartet class k (heavier pair as bra)
long long store_block(int k) {
    static const int ncp[6] = {1, 3, 9, 6, 18, 36};
    int hi = 0;
    while ((hi + 1) * (hi + 2) / 2 <= k) ++hi;
    int lo = k - hi * (hi + 1) / 2;
    return (long long)ncp[hi] * ncp[lo];
}

void launch_jstore(int pclass, const JSArgs& a, int npairs, hipStream_t st) {
    typedef void (*Fn)(const JSArgs&, int, hipStream_t);
    static const Fn fns[6] = {launch_jstore_0, launch_jstore_1, launch_jstore_2,
                              launch_jstore_3, launch_jstore_4, launch_jstore_5};
    fns[pclass](a, npairs, st);
}

void launch_kstore(int type, const KSArgs& a, int nout, hipStream_t st) {
    typedef void (*Fn)(const KSArgs&, int, hipStream_t);
    static const Fn fns[9] = {launch_kstore_0_0, launch_kstore_0_1, launch_kstore_0_2,
                              launch_kstore_1_0, launch_kstore_1_1, launch_kstore_1_2,
                              launch_kstore_2_0, launch_kstore_2_1, launch_kstore_2_2};
    fns[type](a, nout, st);
}

void launch_jreduce(int pclass, const JSArgs& a, int npairs, hipStream_t st) {
    typedef void (*Fn)(const JSArgs&, int, hipStream_t);
    static const Fn fns[6] = {launch_jreduce_0, launch_jreduce_1, launch_jreduce_2,
                              launch_jreduce_3, launch_jreduce_4, launch_jreduce_5};
    fns[pclass](a, npairs, st);
}

void launch_kreduce(int type, const KSArgs& a, int nout, hipStream_t st) {
    typedef void (*Fn)(const KSArgs&, int, hipStream_t);
    static const Fn fns[9] = {launch_kreduce_0_0, launch_kreduce_0_1, launch_kreduce_0_2,
                              launch_kreduce_1_0, launch_kreduce_1_1, launch_kreduce_1_2,
                              launch_kreduce_2_0, launch_kreduce_2_1, launch_kreduce_2_2};
    fns[type](a, nout, st);
}

}  // namespace aimdhf
