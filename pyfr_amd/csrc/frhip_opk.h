// Internal header for the structured operator kernels (see gen_ops.py).
#pragma once

#include "frhip_common.h"

namespace frhip {

// Streaming load of an operand that is read exactly once.
template <typename T>
__device__ __forceinline__ T ldg(const T *p)
{
    return __builtin_nontemporal_load(p);
}

// Output modes: 0: c = acc; 1: c = acc + c; 2: c = alpha*acc + beta*c;
// 3: c = alpha*acc.
template <typename T, int BM>
__device__ __forceinline__ void opstore(T *c, T acc, T alpha, T beta)
{
    if constexpr (BM == 0)
        __builtin_nontemporal_store(acc, c);
    else if constexpr (BM == 1)
        *c = acc + *c;
    else if constexpr (BM == 2)
        *c = alpha*acc + beta*(*c);
    else
        __builtin_nontemporal_store(alpha*acc, c);
}

using OpKD = void (*)(int, const double *, long, double *, long,
                      const double *, double, double);
using OpKF = void (*)(int, const float *, long, float *, long,
                      const float *, float, float);

struct OpKernelEntry {
    int m, k, nnz;
    unsigned long long hash;
    const int32_t *lin;      // row-major linear indices of the nonzeros
    const double *vals;      // the baked nonzero values, same order
    const char *desc;
    OpKD fd[4];
    OpKF ff[4];
};

extern const OpKernelEntry g_opk_table[];
extern const int g_opk_count;

using FKD = void (*)(int, const double *, long, const double *, long,
                     double *, long, const double *, long, void *, double,
                     double, int);
using FKF = void (*)(int, const float *, long, const float *, long, float *,
                     long, const float *, long, void *, double, double, int);

// Fused two-input kernels: C = [A0 | A1] @ [B0; B1] with an epilogue
// (0: store, 1: negdivconf, 2: + RK45 stage, 3: + mixed RK45 stage,
// 4 / 5: the last stage of 2 / 3, 10-13: + RK4 stage 1-4 in slots 6-9).
// fd[3], fd[5] are null (FP32 only); gradient kernels fill slot 0 only.
struct FusedKernelEntry {
    int m, k0, k1, nnz;
    const int32_t *lin;      // row-major indices into the m x (k0 + k1) A
    const double *vals;
    const char *desc;
    // by epilogue slot: ep 0-5, then ep 10-17, then ep 20-23 (null where
    // not built)
    FKD fd[18];
    FKF ff[18];
};

extern const FusedKernelEntry g_fk_table[];
extern const int g_fk_count;

}  // namespace frhip
