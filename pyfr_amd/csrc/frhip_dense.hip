// Dense operator application on the gfx950 matrix cores.
//
// The tetrahedron and prism FR operators (M0, M1 - M3*M2, M3, ... of
// TetShape / PriShape, pyfr/shapes.py:476-570) have no exploitable sparsity.
// The reference hands them to rocBLAS (pyfr/backends/hip/rocblas.py:93-192);
// here they run as
//   C(m, ncol) = [A0 | A1](m, k0 + k1) @ [B0; B1] + epilogue
// on v_mfma_{f32,f64}_16x16x4 with one wavefront per 64 consecutive columns
// of the element-major AoSoA arrays (one 64-element block of one variable):
//
//  * the operator is pre-arranged on the host in MFMA A-fragment order, so
//    each (m tile, k step) operand is one coalesced 256/512 B load that all
//    wavefronts share through the L2;
//  * the 64 columns are split into four 16-column tiles interleaved with
//    stride 4 (tile t holds columns 4j + t), so one lane's B operands for
//    the four tiles are four consecutive columns of one row: a single
//    16/32 B vector load, and likewise one vector store per output row;
//  * B (the element data) is read from HBM exactly once; each k part is
//    padded to a multiple of 4 rows with zero operator columns (the padded
//    B rows re-read a valid row, which the zeros cancel).
//
// Epilogues (template EP): 0 c = acc; 1 c = -rcp*acc (negdivconf);
// 2/4 RK45 stage / last stage; 3/5 their mixed-precision forms; 6 c += acc;
// 7 c = alpha*acc + beta*c; 8 c = alpha*acc.  A scalar FMA kernel (one
// thread per column) covers shapes the MFMA kernel does not take and is one
// of the variants the plan-time autotuner times.
#include "frhip_dense.h"

#include <algorithm>
#include <cstring>
#include <vector>

namespace frhip {

template <typename T> struct V4;
template <> struct V4<float> {
    typedef float type __attribute__((ext_vector_type(4)));
};
template <> struct V4<double> {
    typedef double type __attribute__((ext_vector_type(4)));
};

__device__ __forceinline__ V4<float>::type mfma4(float a, float b,
                                                 V4<float>::type c)
{
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ V4<double>::type mfma4(double a, double b,
                                                  V4<double>::type c)
{
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// One output value at linear offset off = row*ldc + col; ridx indexes rcp.
template <typename T, int EP>
__device__ __forceinline__ void epi1(T acc, T *__restrict__ c, long off,
                                     const T *__restrict__ rcp, long ridx,
                                     void *__restrict__ xa, double dta,
                                     double dtb, T alpha, T beta)
{
    if constexpr (EP == 0)
        c[off] = acc;
    else if constexpr (EP == 6)
        c[off] = acc + c[off];
    else if constexpr (EP == 7)
        c[off] = alpha*acc + beta*c[off];
    else if constexpr (EP == 8)
        c[off] = alpha*acc;
    else {
        const T t = -rcp[ridx]*acc;
        if constexpr (EP == 1)
            c[off] = t;
        else if constexpr (EP == 2) {
            T *r1 = static_cast<T *>(xa) + off;
            const T t1 = *r1;
            *r1 = t1 + T(dta)*t;
            c[off] = t1 + T(dtb)*t;
        }
        else if constexpr (EP == 4) {
            T *r1 = static_cast<T *>(xa) + off;
            *r1 = *r1 + T(dtb)*t;
        }
        else {
            double *r1 = static_cast<double *>(xa) + off;
            const double t1 = *r1, td = (double)t;
            const double na = t1 + dtb*td;
            *r1 = na;
            c[off] = EP == 5 ? (T)na : (T)(t1 + dta*td);
        }
    }
}

// ---- MFMA kernel ----------------------------------------------------------
template <typename T, int MT, int EP>
__global__ __launch_bounds__(256) void k_dense_mfma(
    int m, int ks0, int ks1, int k0, int k1, const T *__restrict__ af,
    long ncol, const T *__restrict__ b0, long ld0, const T *__restrict__ b1,
    long ld1, T *__restrict__ c, long ldc, const T *__restrict__ rcp,
    long ldr, int nvar, void *__restrict__ xa, double dta, double dtb,
    T alpha, T beta)
{
    using V = typename V4<T>::type;
    const int lane = threadIdx.x & 63;
    const long c0 = ((long)blockIdx.x*(blockDim.x >> 6) +
                     (threadIdx.x >> 6))*64;
    if (c0 >= ncol) return;
    const int kr = lane >> 4, cq = (lane & 15)*4;
    const int KS = ks0 + ks1;

    V acc[MT][4];
#pragma unroll
    for (int mt = 0; mt < MT; ++mt)
#pragma unroll
        for (int t = 0; t < 4; ++t)
            acc[mt][t] = V{0, 0, 0, 0};

    // B rows of k step ks for this lane (rows past a part's end re-read its
    // last row; the matching operator columns are zero)
    auto brow = [&](int ks) -> const V * {
        if (ks < ks0) {
            const int r = min(4*ks + kr, k0 - 1);
            return reinterpret_cast<const V *>(b0 + r*ld0 + c0 + cq);
        }
        const int r = min(4*(ks - ks0) + kr, k1 - 1);
        return reinterpret_cast<const V *>(b1 + r*ld1 + c0 + cq);
    };

    const T *afl = af + lane;
    V bc = *brow(0);
    for (int ks = 0; ks < KS; ++ks) {
        const V bn = (ks + 1 < KS) ? *brow(ks + 1) : bc;
        T a[MT];
#pragma unroll
        for (int mt = 0; mt < MT; ++mt)
            a[mt] = afl[((long)mt*KS + ks)*64];
#pragma unroll
        for (int mt = 0; mt < MT; ++mt)
#pragma unroll
            for (int t = 0; t < 4; ++t)
                acc[mt][t] = mfma4(a[mt], bc[t], acc[mt][t]);
        bc = bn;
    }

    // Lane (kr, cq) holds rows 16 mt + 4 kr + r of columns c0 + cq + t
    const long e0 = (c0/(64L*nvar))*64 + cq;
#pragma unroll
    for (int mt = 0; mt < MT; ++mt)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int row = 16*mt + 4*kr + r;
            if (row >= m) continue;
            const long off = row*ldc + c0 + cq;
            if constexpr (EP == 0) {
                const V v = {acc[mt][0][r], acc[mt][1][r], acc[mt][2][r],
                             acc[mt][3][r]};
                *reinterpret_cast<V *>(c + off) = v;
            }
            else if constexpr (EP == 1) {
                const V rv = {rcp[row*ldr + e0], rcp[row*ldr + e0 + 1],
                              rcp[row*ldr + e0 + 2], rcp[row*ldr + e0 + 3]};
                const V v = {-rv[0]*acc[mt][0][r], -rv[1]*acc[mt][1][r],
                             -rv[2]*acc[mt][2][r], -rv[3]*acc[mt][3][r]};
                *reinterpret_cast<V *>(c + off) = v;
            }
            else {
#pragma unroll
                for (int t = 0; t < 4; ++t)
                    epi1<T, EP>(acc[mt][t][r], c, off + t, rcp,
                                row*ldr + e0 + t, xa, dta, dtb, alpha, beta);
            }
        }
}

// ---- scalar kernel --------------------------------------------------------
template <typename T, int EP>
__global__ __launch_bounds__(256) void k_dense_fma(
    int m, int k0, int k1, const T *__restrict__ a, long ncol,
    const T *__restrict__ b0, long ld0, const T *__restrict__ b1, long ld1,
    T *__restrict__ c, long ldc, const T *__restrict__ rcp, long ldr,
    int nvar, void *__restrict__ xa, double dta, double dtb, T alpha, T beta)
{
    const long j = (long)blockIdx.x*blockDim.x + threadIdx.x;
    if (j >= ncol) return;
    const long e = (j/(64L*nvar))*64 + (j & 63);
    const int k = k0 + k1;
    for (int r = 0; r < m; ++r) {
        const T *ar = a + (long)r*k;
        T acc = T(0);
        for (int q = 0; q < k0; ++q)
            acc = fma(ar[q], b0[q*ld0 + j], acc);
        for (int q = 0; q < k1; ++q)
            acc = fma(ar[k0 + q], b1[q*ld1 + j], acc);
        epi1<T, EP>(acc, c, r*ldc + j, rcp, r*ldr + e, xa, dta, dtb, alpha,
                    beta);
    }
}

// ---- dispatch -------------------------------------------------------------
template <typename T, int EP>
static int launch_ep(const DensePlan *p, int variant, long ncol,
                     const void *b0, long ld0, const void *b1, long ld1,
                     void *c, long ldc, const void *rcp, long ldr, int nvar,
                     void *xa, double dta, double dtb, double alpha,
                     double beta, hipStream_t s)
{
    const T *pb0 = (const T *)b0, *pb1 = (const T *)b1;
    const T *prcp = (const T *)rcp;
    if (variant == 0) {
        hipLaunchKernelGGL((k_dense_fma<T, EP>), dim3(cdiv(ncol, 256)),
                           dim3(256), 0, s, p->m, p->k0, p->k1,
                           (const T *)p->arow, ncol, pb0, ld0, pb1, ld1,
                           (T *)c, ldc, prcp, ldr, nvar, xa, dta, dtb,
                           (T)alpha, (T)beta);
        return 0;
    }
    const int wpb = variant;                 // wavefronts per block
    dim3 grid(cdiv(ncol/64, wpb)), block(64*wpb);
#define FR_MT(MTV)                                                          \
    case MTV:                                                               \
        hipLaunchKernelGGL((k_dense_mfma<T, MTV, EP>), grid, block, 0, s,   \
                           p->m, p->ks0, p->ks1, p->k0, p->k1,              \
                           (const T *)p->afrag, ncol, pb0, ld0, pb1, ld1,   \
                           (T *)c, ldc, prcp, ldr, nvar, xa, dta, dtb,      \
                           (T)alpha, (T)beta);                              \
        return 0;
    switch (p->mt) {
        FR_MT(1) FR_MT(2) FR_MT(3) FR_MT(4) FR_MT(5) FR_MT(6) FR_MT(7)
    }
#undef FR_MT
    set_error("dense operator too tall for the MFMA kernel");
    return 2;
}

template <typename T>
static int launch_t(const DensePlan *p, int variant, int ep, long ncol,
                    const void *b0, long ld0, const void *b1, long ld1,
                    void *c, long ldc, const void *rcp, long ldr, int nvar,
                    void *xa, double dta, double dtb, double alpha,
                    double beta, hipStream_t s)
{
#define FR_EP(E)                                                            \
    case E:                                                                 \
        return launch_ep<T, E>(p, variant, ncol, b0, ld0, b1, ld1, c, ldc,  \
                               rcp, ldr, nvar, xa, dta, dtb, alpha, beta, s);
    switch (ep) {
        FR_EP(0) FR_EP(1) FR_EP(2) FR_EP(4) FR_EP(6) FR_EP(7) FR_EP(8)
        case 3: case 5:
            if constexpr (sizeof(T) == 4) {
                if (ep == 3)
                    return launch_ep<T, 3>(p, variant, ncol, b0, ld0, b1,
                                           ld1, c, ldc, rcp, ldr, nvar, xa,
                                           dta, dtb, alpha, beta, s);
                return launch_ep<T, 5>(p, variant, ncol, b0, ld0, b1, ld1,
                                       c, ldc, rcp, ldr, nvar, xa, dta, dtb,
                                       alpha, beta, s);
            }
            break;
    }
#undef FR_EP
    set_error("invalid dense epilogue");
    return 2;
}

// Whether the MFMA kernel's vector accesses are valid for these operands.
static bool mfma_ok(const DensePlan *p, long ncol, const void *b0, long ld0,
                    const void *b1, long ld1, const void *c, long ldc)
{
    const size_t va = 4*(p->dtype == FRHIP_F64 ? 8 : 4);
    auto al = [va](const void *q) {
        return q == nullptr || (reinterpret_cast<uintptr_t>(q) % va) == 0;
    };
    return p->mt >= 1 && p->mt <= 7 && ncol % 64 == 0 && ld0 % 4 == 0 &&
           (p->k1 == 0 || ld1 % 4 == 0) && ldc % 4 == 0 && al(b0) &&
           al(b1) && al(c);
}

int dense_plan_create(DensePlan **out, int m, int k0, int k1,
                      const double *a0, const double *a1, int dtype)
{
    FR_CHECK(m > 0 && k0 > 0 && k1 >= 0, "invalid operator shape");
    const int k = k0 + k1;
    const int ks0 = (k0 + 3)/4, ks1 = (k1 + 3)/4, KS = ks0 + ks1;
    const int mt = (m + 15)/16;
    const size_t esz = dtype == FRHIP_F64 ? 8 : 4;
    auto aval = [&](int r, int q) {       // q < k0: A0 column, else A1
        return q < k0 ? a0[(long)r*k0 + q] : a1[(long)r*k1 + q - k0];
    };

    std::vector<double> frag((size_t)mt*KS*64, 0.0);
    for (int t = 0; t < mt; ++t)
        for (int s = 0; s < KS; ++s)
            for (int l = 0; l < 64; ++l) {
                const int r = 16*t + (l & 15);
                const int kp = 4*(s < ks0 ? s : s - ks0) + (l >> 4);
                double v = 0.0;
                if (r < m) {
                    if (s < ks0 && kp < k0) v = aval(r, kp);
                    else if (s >= ks0 && kp < k1) v = aval(r, k0 + kp);
                }
                frag[((size_t)t*KS + s)*64 + l] = v;
            }
    std::vector<double> row((size_t)m*k);
    for (int r = 0; r < m; ++r)
        for (int q = 0; q < k; ++q) row[(size_t)r*k + q] = aval(r, q);

    auto upload = [&](const std::vector<double> &h, void **d) -> int {
        std::vector<char> buf(h.size()*esz);
        for (size_t i = 0; i < h.size(); ++i) {
            if (esz == 8) std::memcpy(&buf[i*8], &h[i], 8);
            else {
                const float f = (float)h[i];
                std::memcpy(&buf[i*4], &f, 4);
            }
        }
        FR_HIP(hipMalloc(d, buf.size()));
        FR_HIP(hipMemcpy(*d, buf.data(), buf.size(), hipMemcpyHostToDevice));
        return 0;
    };

    DensePlan *p = new DensePlan{m, k0, k1, dtype, mt, ks0, ks1, nullptr,
                                 nullptr, mt <= 7 ? 4 : 0};
    if (upload(frag, &p->afrag) || upload(row, &p->arow)) {
        dense_plan_destroy(p);
        return 1;
    }
    *out = p;
    return 0;
}

void dense_plan_destroy(DensePlan *p)
{
    if (!p) return;
    if (p->afrag) (void)hipFree(p->afrag);
    if (p->arow) (void)hipFree(p->arow);
    delete p;
}

int dense_launch(const DensePlan *p, int ep, long ncol, const void *b0,
                 long ld0, const void *b1, long ld1, void *c, long ldc,
                 const void *rcp, long ldr, int nvar, void *xa, double dta,
                 double dtb, double alpha, double beta, hipStream_t s)
{
    int variant = p->variant;
    if (variant != 0 && !mfma_ok(p, ncol, b0, ld0, b1, ld1, c, ldc))
        variant = 0;
    const int rc = p->dtype == FRHIP_F64
        ? launch_t<double>(p, variant, ep, ncol, b0, ld0, b1, ld1, c, ldc,
                           rcp, ldr, nvar, xa, dta, dtb, alpha, beta, s)
        : launch_t<float>(p, variant, ep, ncol, b0, ld0, b1, ld1, c, ldc,
                          rcp, ldr, nvar, xa, dta, dtb, alpha, beta, s);
    if (rc) return rc;
    FR_LAUNCHED();
    return 0;
}

int dense_tune(DensePlan *p, long ncol, const void *b0, long ld0,
               const void *b1, long ld1, int reps, int *chosen)
{
    FR_CHECK(reps > 0, "reps must be positive");
    if (!mfma_ok(p, ncol, b0, ld0, b1, ld1, nullptr, 64)) {
        p->variant = 0;
        if (chosen) *chosen = 0;
        return 0;
    }
    // Time each variant storing into a scratch C (the real outputs and RK
    // registers are not touched)
    const size_t esz = p->dtype == FRHIP_F64 ? 8 : 4;
    const long ldc = ncol;
    void *c = nullptr;
    FR_HIP(hipMalloc(&c, (size_t)p->m*ldc*esz));
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    const int cands[] = {0, 1, 2, 4};
    float best = 1e30f;
    int bestv = p->variant, rc = 0;
    for (int v : cands) {
        DensePlan q = *p;
        q.variant = v;
        for (int it = 0; it < reps + 1 && !rc; ++it) {
            if (it == 1) (void)hipEventRecord(e0, nullptr);
            rc = dense_launch(&q, 0, ncol, b0, ld0, b1, ld1, c, ldc, nullptr,
                              0, 5, nullptr, 0, 0, 1, 0, nullptr);
        }
        if (rc) break;
        (void)hipEventRecord(e1, nullptr);
        (void)hipEventSynchronize(e1);
        float ms = 0;
        (void)hipEventElapsedTime(&ms, e0, e1);
        if (ms < best) {
            best = ms;
            bestv = v;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(c);
    if (rc) return rc;
    p->variant = bestv;
    if (chosen) *chosen = bestv;
    return 0;
}

}  // namespace frhip
