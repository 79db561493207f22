// Launches the 21 per-class J/K kernels (jk_class.hip, one object per class).
#include "aimdhf_internal.h"

namespace aimdhf {

#define DECL(k) void launch_jk_class_##k(const JKArgs&, const int2*, long long, hipStream_t);
DECL(0) DECL(1) DECL(2) DECL(3) DECL(4) DECL(5) DECL(6) DECL(7) DECL(8) DECL(9) DECL(10)
DECL(11) DECL(12) DECL(13) DECL(14) DECL(15) DECL(16) DECL(17) DECL(18) DECL(19) DECL(20)
#undef DECL

void launch_jk_one(int k, const JKArgs& a, const int2* list, long long nq, hipStream_t st) {
    typedef void (*Fn)(const JKArgs&, const int2*, long long, hipStream_t);
    static const Fn fns[21] = {
        launch_jk_class_0, launch_jk_class_1, launch_jk_class_2, launch_jk_class_3, launch_jk_class_4,
        launch_jk_class_5, launch_jk_class_6, launch_jk_class_7, launch_jk_class_8, launch_jk_class_9,
        launch_jk_class_10, launch_jk_class_11, launch_jk_class_12, launch_jk_class_13, launch_jk_class_14,
        launch_jk_class_15, launch_jk_class_16, launch_jk_class_17, launch_jk_class_18, launch_jk_class_19,
        launch_jk_class_20};
    fns[k](a, list, nq, st);
}

void launch_jk(const JKArgs& a, const int2* const* lists, const long long* counts, hipStream_t st) {
    typedef void (*Fn)(const JKArgs&, const int2*, long long, hipStream_t);
    static const Fn fns[21] = {
        launch_jk_class_0, launch_jk_class_1, launch_jk_class_2, launch_jk_class_3, launch_jk_class_4,
        launch_jk_class_5, launch_jk_class_6, launch_jk_class_7, launch_jk_class_8, launch_jk_class_9,
        launch_jk_class_10, launch_jk_class_11, launch_jk_class_12, launch_jk_class_13, launch_jk_class_14,
        launch_jk_class_15, launch_jk_class_16, launch_jk_class_17, launch_jk_class_18, launch_jk_class_19,
        launch_jk_class_20};
    for (int k = 0; k < 21; ++k) fns[k](a, lists[k], counts[k], st);
}

#define DECLG(k) void launch_grad_class_##k(const GradArgs&, const int2*, long long, hipStream_t);
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

}  // namespace aimdhf
