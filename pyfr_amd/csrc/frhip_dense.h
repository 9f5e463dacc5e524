// Dense (matrix-core) operator plans; internal header.
#pragma once

#include "frhip_common.h"

namespace frhip {

struct DensePlan {
    int m, k0, k1, dtype;
    int mt, ks0, ks1;   // 16-row tiles of A; 4-row k steps of each part
    void *afrag;        // [A0 | A1] in MFMA A-fragment order
    void *arow;         // [A0 | A1] row-major (scalar kernel)
    int variant;        // 0: scalar kernel; 1, 2, 4: MFMA, wavefronts per
                        // block; 16 + (1, 2, 4): operator staged in LDS
};

// BF16-operand plans (FP32 data and accumulation): k steps of 16
struct Bf16Plan {
    int m, k0, k1, mt, kb0, kb1;
    void *afrag;        // [A0 | A1] as bf16 MFMA A fragments (4 per lane)
};

int bf16_plan_create(Bf16Plan **out, int m, int k0, int k1, const double *a0,
                     const double *a1);
void bf16_plan_destroy(Bf16Plan *p);
int bf16_launch(const Bf16Plan *p, int ep, long ncol, const void *b0,
                long ld0, const void *b1, long ld1, void *c, long ldc,
                const void *rcp, long ldr, int nvar, void *xa, double dta,
                double dtb, double alpha, double beta, hipStream_t s);

int dense_plan_create(DensePlan **out, int m, int k0, int k1,
                      const double *a0, const double *a1, int dtype);
void dense_plan_destroy(DensePlan *p);
bool dense_variant_ok(const DensePlan *p, int variant);

// ep: 0 store, 1 negdivconf, 2/4 RK45 stage / last stage, 3/5 mixed RK45
// stage / last stage, 6 c += acc, 7 c = alpha*acc + beta*c, 8 c = alpha*acc.
// Column j of a 5-variable AoSoA array belongs to element
// (j/(64*nvar))*64 + j%64, which indexes rcp (row stride ldr; ldr = 0 for one
// value per element).
int dense_launch(const DensePlan *p, int ep, long ncol, const void *b0,
                 long ld0, const void *b1, long ld1, void *c, long ldc,
                 const void *rcp, long ldr, int nvar, void *xa, double dta,
                 double dtb, double alpha, double beta, hipStream_t s);

// Time the variants on the given inputs (into a scratch output) and keep
// the fastest.  Synchronous; the inputs must be ready.
int dense_tune(DensePlan *p, long ncol, const void *b0, long ld0,
               const void *b1, long ld1, int reps, int *chosen);

}  // namespace frhip
