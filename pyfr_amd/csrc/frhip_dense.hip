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
//  * the result layouts differ by type: register r of lane group kr holds
//    tile row 4 kr + r for v_mfma_f32_16x16x4f32 and kr + 4 r for
//    v_mfma_f64_16x16x4f64 (checked against NumPy by the GPU tests);
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
#include "frhip_physics.h"

#include <algorithm>
#include <cstring>
#include <utility>
#include <vector>

namespace frhip {

// B rows in flight per wavefront in the k loops
constexpr int FR_PD = 3;

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

// Workgroup index with each XCD given a contiguous range: workgroups are
// dispatched to the 8 XCDs round-robin (block b to XCD b % 8), so
// consecutive indices, whose wavefronts read the same element block's data,
// would otherwise sit in different L2s and fetch it from HBM once each.
__device__ __forceinline__ unsigned xcd_block()
{
    const unsigned b = blockIdx.x, nb = gridDim.x;
    const unsigned per = nb/8, rem = nb % 8, x = b % 8;
    return x*per + min(x, rem) + b/8;
}

// The RK4 step of epilogues 10..13: a device double (xa's third pointer),
// so that captured graphs survive a change of step; read once per kernel,
// as stores through the other pointers could alias it.
template <int EP>
__device__ __forceinline__ double rk4_dt(void *xa, double dta)
{
    if constexpr (EP >= 10 && EP <= 17)
        return *static_cast<const double *>(
            static_cast<void *const *>(xa)[2]);
    else
        return dta;
}

// RK4 stage EP - 9 (EP 10..13) on one value: t the stage's RHS, r0 the
// step's initial state, x1 the register r1 holds (unused in stage 1); n1
// the new r1 and nc the next stage's input (unused in stage 4).  Each
// update is the axnpby of pyfr/integrators/explicit/steppers.py:67-108 it
// replaces, operands in the same order.
template <typename T, int EP>
__device__ __forceinline__ void rk4_vals(T t, T r0, T x1, double dta, T &n1,
                                         T &nc)
{
    if constexpr (EP == 10) {           // r1 = k; r2 = r0 + dt/2 k
        n1 = t;
        T a = T(1)*r0;
        a = a + T(dta/2.0)*t;
        nc = a;
    }
    else if constexpr (EP == 11) {      // r1 = dt/6 r1 + r0 + dt/3 k
        T a = T(dta/6.0)*x1;            // r3 = dt/2 k + r0
        a = a + T(1)*r0;
        a = a + T(dta/3.0)*t;
        n1 = a;
        T b = T(dta/2.0)*t;
        b = b + T(1)*r0;
        nc = b;
    }
    else if constexpr (EP == 12) {      // r1 += dt/3 k; r2 = dt k + r0
        n1 = x1 + T(dta/3.0)*t;
        T b = T(dta)*t;
        b = b + T(1)*r0;
        nc = b;
    }
    else {                              // r1 += dt/6 k
        n1 = x1 + T(dta/6.0)*t;
    }
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
        else if constexpr (EP >= 10) {
            // RK4 stage EP - 9 with four registers: xa points to the device
            // pointers {r1, r0, dt}, dta the step read from dt once per
            // kernel (rk4_dt); c receives the next stage's input.
            T *const *pp = static_cast<T *const *>(xa);
            T *r1 = pp[0] + off;
            const T r0 = pp[1][off];
            T n1, nc;
            rk4_vals<T, EP>(t, r0, EP == 10 ? T(0) : *r1, dta, n1, nc);
            *r1 = n1;
            if constexpr (EP != 13)
                c[off] = nc;
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
__device__ __forceinline__ void dense_tiles(
    int m, int ks0, int ks1, int k0, int k1, const T *__restrict__ af,
    long ncol, const T *__restrict__ b0, long ld0, const T *__restrict__ b1,
    long ld1, T *__restrict__ c, long ldc, const T *__restrict__ rcp,
    long ldr, int nvar, void *__restrict__ xa, double dta, double dtb,
    T alpha, T beta)
{
    dta = rk4_dt<EP>(xa, dta);
    using V = typename V4<T>::type;
    const int lane = threadIdx.x & 63;
    const int kr = lane >> 4, cq = (lane & 15)*4;
    const int KS = ks0 + ks1;
    const int W = blockDim.x >> 6;
    const T *afl = af + lane;
    {
    const long g = (long)blockIdx.x*W + (threadIdx.x >> 6);
    if (g*64 >= ncol) return;
    const long c0 = g*64;

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

    // B rows stream through a ring of FR_PD registers: the row of step
    // ks + FR_PD is requested while step ks computes (past the end the
    // last row is re-read, an L2 hit)
    V bq[FR_PD];
#pragma unroll
    for (int i = 0; i < FR_PD; ++i) bq[i] = *brow(min(i, KS - 1));
    for (int kb = 0; kb < KS; kb += FR_PD) {
#pragma unroll
        for (int i = 0; i < FR_PD; ++i) {
            const int ks = kb + i;
            if (ks >= KS) break;
            T a[MT];
#pragma unroll
            for (int mt = 0; mt < MT; ++mt)
                a[mt] = afl[((long)mt*KS + ks)*64];
#pragma unroll
            for (int mt = 0; mt < MT; ++mt)
#pragma unroll
                for (int t = 0; t < 4; ++t)
                    acc[mt][t] = mfma4(a[mt], bq[i][t], acc[mt][t]);
            bq[i] = *brow(min(ks + FR_PD, KS - 1));
        }
    }

    // Lane (kr, cq), result register r holds row 16 mt + 4 kr + r (f32) or
    // 16 mt + kr + 4 r (f64) of columns c0 + cq + t
    const long e0 = (c0/(64L*nvar))*64 + cq;
#pragma unroll
    for (int mt = 0; mt < MT; ++mt)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int row = 16*mt + (sizeof(T) == 8 ? kr + 4*r : 4*kr + r);
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
}

// The block's copy of the operator fragments in LDS (dynamic shared
// memory), filled cooperatively before any wavefront of the block exits, so
// the operator is read from the L2 once per block instead of once per
// wavefront and k step.  (A persistent grid looping over column groups was
// tried: the loop doubled the VGPR count and halved occupancy.)
template <typename T>
__device__ __forceinline__ const T *stage_frags(const T *__restrict__ af,
                                                int n)
{
    extern __shared__ __attribute__((aligned(16))) char fr_smem[];
    T *sa = reinterpret_cast<T *>(fr_smem);
    for (int i = threadIdx.x; i < n; i += blockDim.x) sa[i] = af[i];
    __syncthreads();
    return sa;
}

template <typename T, int MT, int EP, bool LDSA>
__global__ __launch_bounds__(256) void k_dense_mfma(
    int m, int ks0, int ks1, int k0, int k1, const T *__restrict__ af,
    long ncol, const T *__restrict__ b0, long ld0, const T *__restrict__ b1,
    long ld1, T *__restrict__ c, long ldc, const T *__restrict__ rcp,
    long ldr, int nvar, void *__restrict__ xa, double dta, double dtb,
    T alpha, T beta)
{
    const T *a = LDSA ? stage_frags(af, MT*(ks0 + ks1)*64) : af;
    dense_tiles<T, MT, EP>(m, ks0, ks1, k0, k1, a, ncol, b0, ld0, b1, ld1, c,
                           ldc, rcp, ldr, nvar, xa, dta, dtb, alpha, beta);
}

// ---- flux-fused divergence (affine elements, Euler) ----------------------
//
// tdisf + tdivtpcorf + tdivtconf + negdivconf in one pass: the transformed
// flux F = S f(u) never goes to HBM.  Each lane evaluates the flux of its
// wavefront's variable at one solution point for its four columns and feeds
// it straight into the MFMA B operand.  The operator's F columns are
// regrouped on the host into chunks of 4 points x 3 directions (k step
// 3q + d holds points 4q..4q+3 of direction d), so lane group kr computes
// point 4q + kr once per chunk and uses it for three k steps.  S is one
// matrix per element (straight-sided tets and prisms), held in registers.
// The five variable wavefronts of an element block re-read u from the L2;
// HBM sees u, fn and the output once.
template <typename T, int MT, int EP>
__device__ __forceinline__ void flux_tiles(
    int m, int nupts, int nq, int ks1, int k1, const T *__restrict__ af,
    long ncol, const T *__restrict__ u, long ldu, const T *__restrict__ sm,
    long lds, const T *__restrict__ b1, long ld1, T *__restrict__ c,
    long ldc, const T *__restrict__ rcp, long ldr, void *__restrict__ xa,
    double dta, double dtb, T gm1)
{
    dta = rk4_dt<EP>(xa, dta);
    using V = typename V4<T>::type;
    const int lane = threadIdx.x & 63;
    const int kr = lane >> 4, cq = (lane & 15)*4;
    const int KS = 3*nq + ks1;
    const int W = blockDim.x >> 6;
    const T *afl = af + lane;
    {
    const long g = (long)blockIdx.x*W + (threadIdx.x >> 6);
    if (g*64 >= ncol) return;
    const long c0 = g*64;
    const int v = (int)((c0 >> 6) % 5);
    const long eb = c0/320, ecol = eb*64 + cq;

    V S[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int k = 0; k < 3; ++k)
            S[i][k] = *reinterpret_cast<const V *>(sm + (3*i + k)*lds + ecol);

    V acc[MT][4];
#pragma unroll
    for (int mt = 0; mt < MT; ++mt)
#pragma unroll
        for (int t = 0; t < 4; ++t)
            acc[mt][t] = V{0, 0, 0, 0};

    const T *ub = u + eb*320 + cq;
    // u of chunk q + 1 is requested while chunk q computes
    V un[5];
#pragma unroll
    for (int w = 0; w < 5; ++w)
        un[w] = *reinterpret_cast<const V *>(ub + min(kr, nupts - 1)*ldu +
                                             w*64);
    for (int q = 0; q < nq; ++q) {
        V uw[5];
#pragma unroll
        for (int w = 0; w < 5; ++w) uw[w] = un[w];
        const int jn = min(4*(q + 1) + kr, nupts - 1);
#pragma unroll
        for (int w = 0; w < 5; ++w)
            un[w] = *reinterpret_cast<const V *>(ub + jn*ldu + w*64);

        // the chunk's operator fragments, requested before the flux
        // arithmetic so that their latency overlaps it
        T a[3][MT];
#pragma unroll
        for (int d = 0; d < 3; ++d)
#pragma unroll
            for (int mt = 0; mt < MT; ++mt)
                a[d][mt] = afl[((long)mt*KS + 3*q + d)*64];

        // this variable's flux in the three directions, transformed by S
        // (flux.mako:3-26, tflux.mako:24-28)
        V b[3];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const T invrho = T(1)/uw[0][t], E = uw[4][t];
            const T rv[3] = {uw[1][t], uw[2][t], uw[3][t]};
            const T vel[3] = {invrho*rv[0], invrho*rv[1], invrho*rv[2]};
            const T p = gm1*(E - T(0.5)*invrho*((rv[0])*(rv[0]) +
                                                 (rv[1])*(rv[1]) +
                                                 (rv[2])*(rv[2])));
            T fv[3];
            if (v == 0) {
                fv[0] = rv[0]; fv[1] = rv[1]; fv[2] = rv[2];
            }
            else if (v == 4) {
                fv[0] = (E + p)*vel[0]; fv[1] = (E + p)*vel[1];
                fv[2] = (E + p)*vel[2];
            }
            else {
                const T vj = v == 1 ? vel[0] : v == 2 ? vel[1] : vel[2];
                fv[0] = v == 1 ? rv[0]*vj + p : rv[0]*vj;
                fv[1] = v == 2 ? rv[1]*vj + p : rv[1]*vj;
                fv[2] = v == 3 ? rv[2]*vj + p : rv[2]*vj;
            }
#pragma unroll
            for (int d = 0; d < 3; ++d)
                b[d][t] = S[d][0][t]*fv[0] + S[d][1][t]*fv[1] +
                          S[d][2][t]*fv[2];
        }

#pragma unroll
        for (int d = 0; d < 3; ++d)
#pragma unroll
            for (int mt = 0; mt < MT; ++mt)
#pragma unroll
                for (int t = 0; t < 4; ++t)
                    acc[mt][t] = mfma4(a[d][mt], b[d][t], acc[mt][t]);
    }

    // common normal fluxes at the flux points
    auto brow = [&](int s1) -> const V * {
        const int r = min(4*s1 + kr, k1 - 1);
        return reinterpret_cast<const V *>(b1 + r*ld1 + c0 + cq);
    };
    V bq[FR_PD];
#pragma unroll
    for (int i = 0; i < FR_PD; ++i) bq[i] = *brow(min(i, ks1 - 1));
    for (int kb = 0; kb < ks1; kb += FR_PD) {
#pragma unroll
        for (int i = 0; i < FR_PD; ++i) {
            const int s1 = kb + i;
            if (s1 >= ks1) break;
            const int ks = 3*nq + s1;
            T a[MT];
#pragma unroll
            for (int mt = 0; mt < MT; ++mt)
                a[mt] = afl[((long)mt*KS + ks)*64];
#pragma unroll
            for (int mt = 0; mt < MT; ++mt)
#pragma unroll
                for (int t = 0; t < 4; ++t)
                    acc[mt][t] = mfma4(a[mt], bq[i][t], acc[mt][t]);
            bq[i] = *brow(min(s1 + FR_PD, ks1 - 1));
        }
    }

#pragma unroll
    for (int mt = 0; mt < MT; ++mt)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int row = 16*mt + (sizeof(T) == 8 ? kr + 4*r : 4*kr + r);
            if (row >= m) continue;
            const long off = row*ldc + c0 + cq;
            if constexpr (EP == 0) {
                const V o = {acc[mt][0][r], acc[mt][1][r], acc[mt][2][r],
                             acc[mt][3][r]};
                *reinterpret_cast<V *>(c + off) = o;
            }
            else if constexpr (EP == 1) {
                const V rv = {rcp[row*ldr + ecol], rcp[row*ldr + ecol + 1],
                              rcp[row*ldr + ecol + 2],
                              rcp[row*ldr + ecol + 3]};
                const V o = {-rv[0]*acc[mt][0][r], -rv[1]*acc[mt][1][r],
                             -rv[2]*acc[mt][2][r], -rv[3]*acc[mt][3][r]};
                *reinterpret_cast<V *>(c + off) = o;
            }
            else {
#pragma unroll
                for (int t = 0; t < 4; ++t)
                    epi1<T, EP>(acc[mt][t][r], c, off + t, rcp,
                                row*ldr + ecol + t, xa, dta, dtb, T(1), T(0));
            }
        }
    }
}

template <typename T, int MT, int EP, bool LDSA>
__global__ __launch_bounds__(256) void k_tdivf_flux(
    int m, int nupts, int nq, int ks1, int k1, const T *__restrict__ af,
    long ncol, const T *__restrict__ u, long ldu, const T *__restrict__ sm,
    long lds, const T *__restrict__ b1, long ld1, T *__restrict__ c,
    long ldc, const T *__restrict__ rcp, long ldr, void *__restrict__ xa,
    double dta, double dtb, T gm1)
{
    const T *a = LDSA ? stage_frags(af, MT*(3*nq + ks1)*64) : af;
    flux_tiles<T, MT, EP>(m, nupts, nq, ks1, k1, a, ncol, u, ldu, sm, lds,
                          b1, ld1, c, ldc, rcp, ldr, xa, dta, dtb, gm1);
}

// Block-cooperative form: one block of five wavefronts per 64-element
// block, wavefront v owning variable v.  Per chunk of 4 solution points,
// 256 threads each evaluate the full transformed flux (3 directions x 5
// variables) at one (element, point) and write it to LDS; after one barrier
// every wavefront reads its variable's B operands from there.  u is read
// from HBM once and the flux evaluated once per point (the per-wavefront
// form above reads u and evaluates the flux five times).  Double-buffered:
// the barrier of chunk q also orders chunk q - 1's reads before chunk
// q + 1's writes to the same buffer.
template <typename T, int MT, int EP>
__global__ __launch_bounds__(320) void k_tdivf_flux_blk(
    int m, int nupts, int nq, int ks1, int k1, const T *__restrict__ af,
    long ncol, const T *__restrict__ u, long ldu, const T *__restrict__ sm,
    long lds, const T *__restrict__ b1, long ld1, T *__restrict__ c,
    long ldc, const T *__restrict__ rcp, long ldr, void *__restrict__ xa,
    double dta, double dtb, T gm1)
{
    dta = rk4_dt<EP>(xa, dta);
    using V = typename V4<T>::type;
    extern __shared__ __attribute__((aligned(16))) char fr_smem[];
    T *fbuf = reinterpret_cast<T *>(fr_smem);       // [2][4][3][5][64]
    constexpr int FB = 4*3*5*64;

    const int tid = threadIdx.x, lane = tid & 63, v = tid >> 6;
    const int kr = lane >> 4, cq = (lane & 15)*4;
    const long eb = blockIdx.x;
    const long c0 = (eb*5 + v)*64;
    const int KS = 3*nq + ks1;

    // flux-evaluation role: element e, chunk point jj (threads 0..255)
    const int fe = tid & 63, fj = tid >> 6;
    const long egl = eb*64 + fe;
    T S[3][3];
    if (fj < 4)
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int k = 0; k < 3; ++k) S[i][k] = sm[(3*i + k)*lds + egl];
    const T *ub = u + eb*320 + fe;

    V acc[MT][4];
#pragma unroll
    for (int mt = 0; mt < MT; ++mt)
#pragma unroll
        for (int t = 0; t < 4; ++t)
            acc[mt][t] = V{0, 0, 0, 0};
    const T *afl = af + lane;

    for (int q = 0; q < nq; ++q) {
        T *fb = fbuf + (q & 1)*FB;
        if (fj < 4) {
            const int j = min(4*q + fj, nupts - 1);
            T uu[5];
#pragma unroll
            for (int w = 0; w < 5; ++w) uu[w] = ub[j*ldu + w*64];
            T f[3][5], p, vel[3];
            inviscid_flux(uu, f, p, vel, gm1);
#pragma unroll
            for (int d = 0; d < 3; ++d)
#pragma unroll
                for (int w = 0; w < 5; ++w)
                    fb[((fj*3 + d)*5 + w)*64 + fe] =
                        S[d][0]*f[0][w] + S[d][1]*f[1][w] + S[d][2]*f[2][w];
        }
        T a[3][MT];
#pragma unroll
        for (int d = 0; d < 3; ++d)
#pragma unroll
            for (int mt = 0; mt < MT; ++mt)
                a[d][mt] = afl[((long)mt*KS + 3*q + d)*64];
        __syncthreads();
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const V bd = *reinterpret_cast<const V *>(
                fb + ((kr*3 + d)*5 + v)*64 + cq);
#pragma unroll
            for (int mt = 0; mt < MT; ++mt)
#pragma unroll
                for (int t = 0; t < 4; ++t)
                    acc[mt][t] = mfma4(a[d][mt], bd[t], acc[mt][t]);
        }
    }

    // common normal fluxes at the flux points (this wavefront's variable)
    auto brow = [&](int s1) -> const V * {
        const int r = min(4*s1 + kr, k1 - 1);
        return reinterpret_cast<const V *>(b1 + r*ld1 + c0 + cq);
    };
    V bq[FR_PD];
#pragma unroll
    for (int i = 0; i < FR_PD; ++i) bq[i] = *brow(min(i, ks1 - 1));
    for (int kb = 0; kb < ks1; kb += FR_PD) {
#pragma unroll
        for (int i = 0; i < FR_PD; ++i) {
            const int s1 = kb + i;
            if (s1 >= ks1) break;
            const int ks = 3*nq + s1;
            T a[MT];
#pragma unroll
            for (int mt = 0; mt < MT; ++mt)
                a[mt] = afl[((long)mt*KS + ks)*64];
#pragma unroll
            for (int mt = 0; mt < MT; ++mt)
#pragma unroll
                for (int t = 0; t < 4; ++t)
                    acc[mt][t] = mfma4(a[mt], bq[i][t], acc[mt][t]);
            bq[i] = *brow(min(s1 + FR_PD, ks1 - 1));
        }
    }

    const long ecol = eb*64 + cq;
#pragma unroll
    for (int mt = 0; mt < MT; ++mt)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int row = 16*mt + (sizeof(T) == 8 ? kr + 4*r : 4*kr + r);
            if (row >= m) continue;
            const long off = row*ldc + c0 + cq;
            if constexpr (EP == 0) {
                const V o = {acc[mt][0][r], acc[mt][1][r], acc[mt][2][r],
                             acc[mt][3][r]};
                *reinterpret_cast<V *>(c + off) = o;
            }
            else if constexpr (EP == 1) {
                const V rv = {rcp[row*ldr + ecol], rcp[row*ldr + ecol + 1],
                              rcp[row*ldr + ecol + 2],
                              rcp[row*ldr + ecol + 3]};
                const V o = {-rv[0]*acc[mt][0][r], -rv[1]*acc[mt][1][r],
                             -rv[2]*acc[mt][2][r], -rv[3]*acc[mt][3][r]};
                *reinterpret_cast<V *>(c + off) = o;
            }
            else {
#pragma unroll
                for (int t = 0; t < 4; ++t)
                    epi1<T, EP>(acc[mt][t][r], c, off + t, rcp,
                                row*ldr + ecol + t, xa, dta, dtb, T(1), T(0));
            }
        }
    (void)ncol;
}

// ---- scalar kernel --------------------------------------------------------
template <typename T, int EP>
__global__ __launch_bounds__(256) void k_dense_fma(
    int m, int k0, int k1, const T *__restrict__ a, long ncol,
    const T *__restrict__ b0, long ld0, const T *__restrict__ b1, long ld1,
    T *__restrict__ c, long ldc, const T *__restrict__ rcp, long ldr,
    int nvar, void *__restrict__ xa, double dta, double dtb, T alpha, T beta)
{
    dta = rk4_dt<EP>(xa, dta);
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

// 4x4x1 (16-block) MFMA form, FP32: one lane per element column.
//
// v_mfma_f32_4x4x1f32 computes 16 independent 4x4 outer products; block b
// takes its B value from lane 4b + j and, with the A-broadcast controls
// CBSZ = 4 and ABID = k, every block takes A from lanes 4k..4k+3 (lane
// 4b + j, result register i = A_k[i] * B_b[j]; checked on the device).
// Each lane's B value is its own column's datum, so lane 4b + j owns
// column 4b + j and register i of accumulator g holds output row 4g + i:
// the m-row operator is padded to 4 rows only (36 for the 35-row tet
// operator, against 48 with 16-row tiles).  One operator VGPR holds 16
// (k, 4-row group) entries, selected by ABID: fragment (chunk q, group g)
// has lane 4k + i = A[4g + i][k-th column of chunk q].  Flux chunks cover
// 5 solution points x 3 directions (15 k), evaluated per lane for its own
// element; fn chunks cover 16 rows.
typedef float FrV4f __attribute__((ext_vector_type(4)));

template <int G, int KK>
__device__ __forceinline__ void mfma4x4_k(FrV4f (&acc)[G], const float (&a)[G],
                                          float b)
{
#pragma unroll
    for (int g = 0; g < G; ++g)
        acc[g] = __builtin_amdgcn_mfma_f32_4x4x1f32(a[g], b, acc[g], 4, KK,
                                                    0);
}

template <int G, int N, int... KK>
__device__ __forceinline__ void mfma4x4_chunk(
    FrV4f (&acc)[G], const float (&a)[G], const float (&b)[N],
    std::integer_sequence<int, KK...>)
{
    (mfma4x4_k<G, KK>(acc, a, b[KK]), ...);
}

// One wavefront per (64-element block, variable), in blocks of 4 or 5
// wavefronts (FR_V4 / FR_V4B); no barrier orders the variables' reads of u
// before another's writes, so c must not alias u (an RHS in place runs the
// block-cooperative form instead).  Up to 10 row groups (tets p <= 4) the
// kernel is held to 128 VGPRs, four wavefronts per SIMD: uncapped, the
// epilogues with 1/|J| take 132-133 and run three (a few spilled registers
// cost less than the lost wavefront: p = 4 tets, 0.682 -> 0.648 ms per RHS
// and 6.66 -> 6.09 ms per RK4 step on the 48^3 x 6 box).
template <int G, int EP>
__global__ __launch_bounds__(320)
__attribute__((amdgpu_waves_per_eu(G <= 10 ? 4 : 2))) void k_tdivf_flux4(
    int m, int nupts, int nc0, int nc1, int k1, const float *__restrict__ af,
    long ncol, const float *__restrict__ u, long ldu,
    const float *__restrict__ sm, long lds, const float *__restrict__ b1,
    long ld1, float *__restrict__ c, long ldc, const float *__restrict__ rcp,
    long ldr, void *__restrict__ xa, double dta, double dtb, float gm1)
{
    dta = rk4_dt<EP>(xa, dta);
    const int lane = threadIdx.x & 63;
    const long w = (long)xcd_block()*(blockDim.x >> 6) + (threadIdx.x >> 6);
    if (w*64 >= ncol) return;
    const long c0 = w*64;
    const int v = (int)((c0 >> 6) % 5);
    const long eb = c0/320, ecol = eb*64 + lane;

    float S[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int k = 0; k < 3; ++k) S[i][k] = sm[(3*i + k)*lds + ecol];

    FrV4f acc[G];
#pragma unroll
    for (int g = 0; g < G; ++g) acc[g] = FrV4f{0, 0, 0, 0};

    const float *ub = u + eb*320 + lane;
    const float *afl = af + lane;
    // u of chunk q + 1 is requested while chunk q computes; points past
    // the end re-read the last point, whose operator columns are zero
    float un[5][5];
#pragma unroll
    for (int pp = 0; pp < 5; ++pp)
#pragma unroll
        for (int x = 0; x < 5; ++x)
            un[pp][x] = ub[min(pp, nupts - 1)*ldu + x*64];
    // the operator fragments of chunk q + 1 likewise
    float an[G];
#pragma unroll
    for (int g = 0; g < G; ++g) an[g] = afl[(long)g*64];
    for (int q = 0; q < nc0; ++q) {
        float a[G];
        const long qn = min(q + 1, nc0 + nc1 - 1);
#pragma unroll
        for (int g = 0; g < G; ++g) {
            a[g] = an[g];
            an[g] = afl[(qn*G + g)*64];
        }
        float uc[5][5];
#pragma unroll
        for (int pp = 0; pp < 5; ++pp)
#pragma unroll
            for (int x = 0; x < 5; ++x) {
                uc[pp][x] = un[pp][x];
                un[pp][x] = ub[min(5*(q + 1) + pp, nupts - 1)*ldu + x*64];
            }
        // this variable's transformed flux at the chunk's 5 points
        // (flux.mako:3-26, tflux.mako:24-28)
        float b[15];
#pragma unroll
        for (int pp = 0; pp < 5; ++pp) {
            const float *uw = uc[pp];
            const float invrho = 1.0f/uw[0], E = uw[4];
            const float rv[3] = {uw[1], uw[2], uw[3]};
            const float vel[3] = {invrho*rv[0], invrho*rv[1], invrho*rv[2]};
            const float p = gm1*(E - 0.5f*invrho*((rv[0])*(rv[0]) +
                                                  (rv[1])*(rv[1]) +
                                                  (rv[2])*(rv[2])));
            float fv[3];
            if (v == 0) {
                fv[0] = rv[0]; fv[1] = rv[1]; fv[2] = rv[2];
            }
            else if (v == 4) {
                fv[0] = (E + p)*vel[0]; fv[1] = (E + p)*vel[1];
                fv[2] = (E + p)*vel[2];
            }
            else {
                const float vj = v == 1 ? vel[0] : v == 2 ? vel[1] : vel[2];
                fv[0] = v == 1 ? rv[0]*vj + p : rv[0]*vj;
                fv[1] = v == 2 ? rv[1]*vj + p : rv[1]*vj;
                fv[2] = v == 3 ? rv[2]*vj + p : rv[2]*vj;
            }
#pragma unroll
            for (int d = 0; d < 3; ++d)
                b[3*pp + d] = S[d][0]*fv[0] + S[d][1]*fv[1] + S[d][2]*fv[2];
        }
        mfma4x4_chunk(acc, a, b, std::make_integer_sequence<int, 15>{});
    }

    // common normal fluxes at the flux points, 16 rows per chunk (rows past
    // the end re-read the last row; their operator columns are zero)
    const float *bl = b1 + c0 + lane;
    for (int q = 0; q < nc1; ++q) {
        float a[G], b[16];
        const long qn = min(nc0 + q + 1, nc0 + nc1 - 1);
#pragma unroll
        for (int g = 0; g < G; ++g) {
            a[g] = an[g];
            an[g] = afl[(qn*G + g)*64];
        }
#pragma unroll
        for (int kk = 0; kk < 16; ++kk)
            b[kk] = bl[min(16*q + kk, k1 - 1)*ld1];
        mfma4x4_chunk(acc, a, b, std::make_integer_sequence<int, 16>{});
    }

#pragma unroll
    for (int g = 0; g < G; ++g)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int row = 4*g + i;
            if (row >= m) continue;
            const long off = row*ldc + c0 + lane;
            if constexpr (EP == 0)
                c[off] = acc[g][i];
            else if constexpr (EP == 1)
                c[off] = -rcp[row*ldr + ecol]*acc[g][i];
            else
                epi1<float, EP>(acc[g][i], c, off, rcp, row*ldr + ecol,
                                xa, dta, dtb, 1.0f, 0.0f);
        }
}

// ---- dispatch -------------------------------------------------------------

// Variant codes: 0 scalar kernel; 1, 2, 4 MFMA with that many wavefronts
// per block and the operator read through the L2; 16 + (1, 2, 4) the same
// with the operator staged in LDS once per block.
constexpr int FR_LDSV = 16;
// flux-fused kernels only: the block-cooperative form (five wavefronts per
// element block, flux evaluated once into LDS)
constexpr int FR_BLKV = 32;
// flux-fused kernels only, FP32: the 4x4x1 MFMA form (k_tdivf_flux4)
constexpr int FR_V4 = 64;
// ... with one block of five wavefronts per 64-element block (all its
// variables in one workgroup, so on one XCD and its L2: u is fetched from
// HBM once, not once per XCD the variables land on)
constexpr int FR_V4B = 65;
constexpr size_t FR_LDS_MAX = 65536;

static bool variant_valid(int v, size_t ldsbytes, bool scalar_ok)
{
    const int w = v & ~FR_LDSV;
    if (v == 0) return scalar_ok;
    if (w != 1 && w != 2 && w != 4) return false;
    return !(v & FR_LDSV) || ldsbytes <= FR_LDS_MAX;
}

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
    const int wpb = variant & ~FR_LDSV;      // wavefronts per block
    const bool lds = variant & FR_LDSV;
    const size_t ldsb = lds ? (size_t)p->mt*(p->ks0 + p->ks1)*64*sizeof(T)
                            : 0;
    dim3 block(64*wpb);
#define FR_MT(MTV)                                                          \
    case MTV: {                                                             \
        auto k = lds ? k_dense_mfma<T, MTV, EP, true>                       \
                     : k_dense_mfma<T, MTV, EP, false>;                     \
        dim3 grid(cdiv(ncol/64, wpb));                                      \
        hipLaunchKernelGGL(k, grid, block, ldsb, s, p->m, p->ks0, p->ks1,   \
                           p->k0, p->k1, (const T *)p->afrag, ncol, pb0,    \
                           ld0, pb1, ld1, (T *)c, ldc, prcp, ldr, nvar, xa, \
                           dta, dtb, (T)alpha, (T)beta);                    \
        return 0;                                                           \
    }
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

bool dense_variant_ok(const DensePlan *p, int variant)
{
    const size_t esz = p->dtype == FRHIP_F64 ? 8 : 4;
    return variant_valid(variant, (size_t)p->mt*(p->ks0 + p->ks1)*64*esz,
                         true);
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
    const size_t esz = p->dtype == FRHIP_F64 ? 8 : 4;
    if (!variant_valid(variant, (size_t)p->mt*(p->ks0 + p->ks1)*64*esz,
                       true))
        variant &= ~FR_LDSV;
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
    const int cands[] = {0, 1, 2, 4, FR_LDSV + 1, FR_LDSV + 2, FR_LDSV + 4};
    const size_t ldsb = (size_t)p->mt*(p->ks0 + p->ks1)*64*esz;
    float best = 1e30f;
    int bestv = p->variant, rc = 0;
    for (int v : cands) {
        if (!variant_valid(v, ldsb, true)) continue;
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

// ---- BF16 operands on the matrix cores (precision-form experiment) ------
//
// C = [A0 | A1] @ [B0; B1] with FP32 data in HBM, both MFMA operands rounded
// to bfloat16 (round to nearest even) and FP32 accumulation:
// v_mfma_f32_16x16x16bf16_1k, k steps of 16.  Lane group kr supplies rows
// 16s + 4kr .. 16s + 4kr + 3 of B for its four interleaved column tiles, so
// it loads four B rows of 4 columns and transposes them in registers.

typedef short s4v __attribute__((ext_vector_type(4)));

__device__ __forceinline__ short to_bf16(float f)
{
    unsigned x = __float_as_uint(f);
    x += 0x7fffu + ((x >> 16) & 1u);
    return (short)(x >> 16);
}

static unsigned short host_bf16(float f)
{
    unsigned x;
    std::memcpy(&x, &f, 4);
    x += 0x7fffu + ((x >> 16) & 1u);
    return (unsigned short)(x >> 16);
}

template <int MT, int EP>
__global__ __launch_bounds__(256) void k_dense_bf16(
    int m, int kb0, int kb1, int k0, int k1, const s4v *__restrict__ af,
    long ncol, const float *__restrict__ b0, long ld0,
    const float *__restrict__ b1, long ld1, float *__restrict__ c, long ldc,
    const float *__restrict__ rcp, long ldr, int nvar,
    void *__restrict__ xa, double dta, double dtb, float alpha, float beta)
{
    dta = rk4_dt<EP>(xa, dta);
    using V = V4<float>::type;
    const int lane = threadIdx.x & 63;
    const long c0 = ((long)blockIdx.x*(blockDim.x >> 6) +
                     (threadIdx.x >> 6))*64;
    if (c0 >= ncol) return;
    const int kr = lane >> 4, cq = (lane & 15)*4;
    const int KS = kb0 + kb1;

    V acc[MT][4];
#pragma unroll
    for (int mt = 0; mt < MT; ++mt)
#pragma unroll
        for (int t = 0; t < 4; ++t)
            acc[mt][t] = V{0, 0, 0, 0};

    for (int s = 0; s < KS; ++s) {
        const bool p0 = s < kb0;
        const float *bp = p0 ? b0 : b1;
        const long ld = p0 ? ld0 : ld1;
        const int kk = p0 ? k0 : k1;
        const int r0 = 16*(p0 ? s : s - kb0) + 4*kr;
        V rows[4];
#pragma unroll
        for (int i = 0; i < 4; ++i)
            rows[i] = *reinterpret_cast<const V *>(
                bp + min(r0 + i, kk - 1)*ld + c0 + cq);
        s4v bop[4];
#pragma unroll
        for (int t = 0; t < 4; ++t)
            bop[t] = s4v{to_bf16(rows[0][t]), to_bf16(rows[1][t]),
                         to_bf16(rows[2][t]), to_bf16(rows[3][t])};
#pragma unroll
        for (int mt = 0; mt < MT; ++mt) {
            const s4v a = af[((long)mt*KS + s)*64 + lane];
#pragma unroll
            for (int t = 0; t < 4; ++t)
                acc[mt][t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(
                    a, bop[t], acc[mt][t], 0, 0, 0);
        }
    }

    const long e0 = (c0/(64L*nvar))*64 + cq;
#pragma unroll
    for (int mt = 0; mt < MT; ++mt)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int row = 16*mt + 4*kr + r;
            if (row >= m) continue;
            const long off = row*ldc + c0 + cq;
#pragma unroll
            for (int t = 0; t < 4; ++t)
                epi1<float, EP>(acc[mt][t][r], c, off + t, rcp,
                                row*ldr + e0 + t, xa, dta, dtb, alpha, beta);
        }
}

int bf16_plan_create(Bf16Plan **out, int m, int k0, int k1, const double *a0,
                     const double *a1)
{
    FR_CHECK(m > 0 && k0 > 0 && k1 >= 0, "invalid operator shape");
    const int mt = (m + 15)/16;
    FR_CHECK(mt <= 8 || mt == 12,
             "bf16 operator plans take at most 128 (or exactly 177-192) rows");
    const int kb0 = (k0 + 15)/16, kb1 = (k1 + 15)/16, KS = kb0 + kb1;
    std::vector<unsigned short> frag((size_t)mt*KS*64*4, 0);
    for (int t = 0; t < mt; ++t)
        for (int s = 0; s < KS; ++s)
            for (int l = 0; l < 64; ++l)
                for (int i = 0; i < 4; ++i) {
                    const int r = 16*t + (l & 15);
                    const int kp = 16*(s < kb0 ? s : s - kb0) + 4*(l >> 4) + i;
                    double v = 0.0;
                    if (r < m) {
                        if (s < kb0 && kp < k0) v = a0[(long)r*k0 + kp];
                        else if (s >= kb0 && kp < k1) v = a1[(long)r*k1 + kp];
                    }
                    frag[(((size_t)t*KS + s)*64 + l)*4 + i] =
                        host_bf16((float)v);
                }
    void *d = nullptr;
    FR_HIP(hipMalloc(&d, frag.size()*2));
    hipError_t e = hipMemcpy(d, frag.data(), frag.size()*2,
                             hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        (void)hipFree(d);
        FR_HIP(e);
    }
    *out = new Bf16Plan{m, k0, k1, mt, kb0, kb1, d};
    return 0;
}

void bf16_plan_destroy(Bf16Plan *p)
{
    if (!p) return;
    if (p->afrag) (void)hipFree(p->afrag);
    delete p;
}

template <int EP>
static void launch_bf16_ep(const Bf16Plan *p, long ncol, const void *b0,
                           long ld0, const void *b1, long ld1, void *c,
                           long ldc, const void *rcp, long ldr, int nvar,
                           void *xa, double dta, double dtb, double alpha,
                           double beta, hipStream_t s)
{
    dim3 grid(cdiv(ncol/64, 4)), block(256);
#define FR_MT(MTV)                                                          \
    case MTV:                                                               \
        hipLaunchKernelGGL((k_dense_bf16<MTV, EP>), grid, block, 0, s,      \
                           p->m, p->kb0, p->kb1, p->k0, p->k1,              \
                           (const s4v *)p->afrag, ncol, (const float *)b0,  \
                           ld0, (const float *)b1, ld1, (float *)c, ldc,    \
                           (const float *)rcp, ldr, nvar, xa, dta, dtb,     \
                           (float)alpha, (float)beta);                      \
        break;
    switch (p->mt) {
        FR_MT(1) FR_MT(2) FR_MT(3) FR_MT(4) FR_MT(5) FR_MT(6) FR_MT(7)
        FR_MT(8) FR_MT(12)
    }
#undef FR_MT
}

int bf16_launch(const Bf16Plan *p, int ep, long ncol, const void *b0,
                long ld0, const void *b1, long ld1, void *c, long ldc,
                const void *rcp, long ldr, int nvar, void *xa, double dta,
                double dtb, double alpha, double beta, hipStream_t s)
{
    const size_t va = 16;
    auto al = [va](const void *q) {
        return q == nullptr || (reinterpret_cast<uintptr_t>(q) % va) == 0;
    };
    FR_CHECK(ncol % 64 == 0 && ld0 % 4 == 0 && (p->k1 == 0 || ld1 % 4 == 0)
             && ldc % 4 == 0 && al(b0) && al(b1) && al(c),
             "bf16 plans need 64-column groups and 4-element alignment");
#define FR_EP(E)                                                            \
    case E:                                                                 \
        launch_bf16_ep<E>(p, ncol, b0, ld0, b1, ld1, c, ldc, rcp, ldr,      \
                          nvar, xa, dta, dtb, alpha, beta, s);              \
        break;
    switch (ep) {
        FR_EP(0) FR_EP(1) FR_EP(2) FR_EP(3) FR_EP(4) FR_EP(5) FR_EP(6)
        FR_EP(7) FR_EP(8)
        default:
            set_error("invalid bf16 epilogue");
            return 2;
    }
#undef FR_EP
    FR_LAUNCHED();
    return 0;
}

// ---- flux-fused plans -----------------------------------------------------

struct FluxPlan {
    DensePlan d;        // m, k0 = 3 nupts, k1 = nfpts, fragments
    int nupts, nq;
    void *af4;          // FP32: k_tdivf_flux4 fragments, or null
    int g4, nc0, nc1;   // 4-row groups, flux and fn chunks
};

// Row-group counts with a k_tdivf_flux4 instantiation: tets p = 1..4
// (4, 10, 20, 35 rows), prisms p = 1..4 (6, 18, 40, 75 rows) and pyramids
// p = 1..4 (5, 14, 30, 55 rows)
static bool flux4_ok(int g)
{
    return g == 1 || g == 2 || g == 3 || g == 4 || g == 5 || g == 8 ||
           g == 9 || g == 10 || g == 14 || g == 19;
}

template <typename T, int EP>
static void launch_flux_ep(const FluxPlan *p, int wpb, long ncol,
                           const void *u, long ldu, const void *sm, long lds,
                           const void *fn, long ldf, void *c, long ldc,
                           const void *rcp, long ldr, void *xa, double dta,
                           double dtb, double gamma, hipStream_t s)
{
    if constexpr (sizeof(T) == 4) {
        if (wpb == FR_V4 || wpb == FR_V4B) {
            const int wv = wpb == FR_V4B ? 5 : 4;
            const dim3 grid(cdiv(ncol/64, wv)), block(64*wv);
#define FR_G4(GV)                                                           \
            case GV:                                                        \
                hipLaunchKernelGGL((k_tdivf_flux4<GV, EP>), grid, block, 0, \
                                   s, p->d.m, p->nupts, p->nc0, p->nc1,     \
                                   p->d.k1, (const float *)p->af4, ncol,    \
                                   (const float *)u, ldu, (const float *)sm,\
                                   lds, (const float *)fn, ldf, (float *)c, \
                                   ldc, (const float *)rcp, ldr, xa, dta,   \
                                   dtb, (float)(gamma - 1));                \
                break;
            switch (p->g4) {
                FR_G4(1) FR_G4(2) FR_G4(3) FR_G4(4) FR_G4(5) FR_G4(8)
                FR_G4(9) FR_G4(10) FR_G4(14) FR_G4(19)
            }
#undef FR_G4
            return;
        }
    }
    if (wpb == FR_BLKV) {
        constexpr size_t fbb = 2*4*3*5*64*sizeof(T);
#define FR_MTB(MTV)                                                         \
        case MTV:                                                           \
            hipLaunchKernelGGL((k_tdivf_flux_blk<T, MTV, EP>),              \
                               dim3(cdiv(ncol, 320)), dim3(320), fbb, s,    \
                               p->d.m, p->nupts, p->nq, p->d.ks1, p->d.k1,  \
                               (const T *)p->d.afrag, ncol, (const T *)u,   \
                               ldu, (const T *)sm, lds, (const T *)fn, ldf, \
                               (T *)c, ldc, (const T *)rcp, ldr, xa, dta,   \
                               dtb, (T)(gamma - 1));                        \
            break;
        switch (p->d.mt) {
            FR_MTB(1) FR_MTB(2) FR_MTB(3) FR_MTB(4) FR_MTB(5)
        }
#undef FR_MTB
        return;
    }
    const bool stage = wpb & FR_LDSV;
    const int w = wpb & ~FR_LDSV;
    const size_t ldsb = stage ? (size_t)p->d.mt*(p->d.ks0 + p->d.ks1)*64*
                                    sizeof(T) : 0;
    dim3 block(64*w);
#define FR_MT(MTV)                                                          \
    case MTV: {                                                             \
        auto k = stage ? k_tdivf_flux<T, MTV, EP, true>                     \
                       : k_tdivf_flux<T, MTV, EP, false>;                   \
        dim3 grid(cdiv(ncol/64, w));                                        \
        hipLaunchKernelGGL(k, grid, block, ldsb, s, p->d.m, p->nupts,       \
                           p->nq, p->d.ks1, p->d.k1, (const T *)p->d.afrag, \
                           ncol, (const T *)u, ldu, (const T *)sm, lds,     \
                           (const T *)fn, ldf, (T *)c, ldc,                 \
                           (const T *)rcp, ldr, xa, dta, dtb,               \
                           (T)(gamma - 1));                                 \
        break;                                                              \
    }
    switch (p->d.mt) {
        FR_MT(1) FR_MT(2) FR_MT(3) FR_MT(4) FR_MT(5)
    }
#undef FR_MT
}

template <typename T>
static int launch_flux_t(const FluxPlan *p, int wpb, int ep, long ncol,
                         const void *u, long ldu, const void *sm, long lds,
                         const void *fn, long ldf, void *c, long ldc,
                         const void *rcp, long ldr, void *xa, double dta,
                         double dtb, double gamma, hipStream_t s)
{
#define FR_EP(E)                                                            \
    case E:                                                                 \
        launch_flux_ep<T, E>(p, wpb, ncol, u, ldu, sm, lds, fn, ldf, c,     \
                             ldc, rcp, ldr, xa, dta, dtb, gamma, s);        \
        return 0;
    switch (ep) {
        FR_EP(0) FR_EP(1) FR_EP(2) FR_EP(4)
        FR_EP(10) FR_EP(11) FR_EP(12) FR_EP(13)
        case 3: case 5:
            if constexpr (sizeof(T) == 4) {
                if (ep == 3)
                    launch_flux_ep<T, 3>(p, wpb, ncol, u, ldu, sm, lds, fn,
                                         ldf, c, ldc, rcp, ldr, xa, dta, dtb,
                                         gamma, s);
                else
                    launch_flux_ep<T, 5>(p, wpb, ncol, u, ldu, sm, lds, fn,
                                         ldf, c, ldc, rcp, ldr, xa, dta, dtb,
                                         gamma, s);
                return 0;
            }
            break;
    }
#undef FR_EP
    set_error("invalid flux-fused epilogue");
    return 2;
}

static int flux_launch(const FluxPlan *p, int ep, long ncol, const void *u,
                       long ldu, const void *sm, long lds, const void *fn,
                       long ldf, void *c, long ldc, const void *rcp, long ldr,
                       void *xa, double dta, double dtb, double gamma,
                       hipStream_t s)
{
    const size_t va = 4*(p->d.dtype == FRHIP_F64 ? 8 : 4);
    auto al = [va](const void *q) {
        return (reinterpret_cast<uintptr_t>(q) % va) == 0;
    };
    FR_CHECK(ncol % 320 == 0, "ncol must cover whole 64-element blocks of 5 "
             "variables");
    FR_CHECK(ldu >= ncol && ldf >= ncol && ldc >= ncol && lds*5 >= ncol,
             "invalid leading dimensions");
    FR_CHECK(ldu % 4 == 0 && lds % 4 == 0 && ldf % 4 == 0 && ldc % 4 == 0 &&
             al(u) && al(sm) && al(fn) && al(c),
             "flux-fused operands must be 4-element aligned");
    FR_CHECK(ep == 0 || rcp, "epilogue needs 1/|J|");
    FR_CHECK(ep < 2 || xa, "RK epilogue needs the r1 / accumulator array");
    const size_t esz = p->d.dtype == FRHIP_F64 ? 8 : 4;
    int wpb = p->d.variant > 0 ? p->d.variant : 4;
    if ((wpb == FR_V4 || wpb == FR_V4B) && !p->af4)
        wpb = 4;
    if (wpb != FR_BLKV && wpb != FR_V4 && wpb != FR_V4B &&
        !variant_valid(wpb, (size_t)p->d.mt*(p->d.ks0 + p->d.ks1)*64*esz,
                       false))
        wpb &= ~FR_LDSV;
    // an RHS in place (c == u), or an RK45 stage whose accumulator is the
    // input (xa == u: the first stage, rkvdh2 with r1 the state): only the
    // block-cooperative form keeps an element block's five wavefronts in
    // one block, whose barriers order all reads of u before any write
    if ((c == u && ep != 13) || (xa == u && ep >= 2 && ep <= 5))
        wpb = FR_BLKV;
    const int rc = p->d.dtype == FRHIP_F64
        ? launch_flux_t<double>(p, wpb, ep, ncol, u, ldu, sm, lds, fn, ldf, c,
                                ldc, rcp, ldr, xa, dta, dtb, gamma, s)
        : launch_flux_t<float>(p, wpb, ep, ncol, u, ldu, sm, lds, fn, ldf, c,
                               ldc, rcp, ldr, xa, dta, dtb, gamma, s);
    if (rc) return rc;
    FR_LAUNCHED();
    return 0;
}

}  // namespace frhip

using namespace frhip;

extern "C" {

int frhip_tdivf_plan_create(void **plan, int m, int nupts, int nfpts,
                            const double *a0, const double *a1, int dtype)
{
    FR_API_BEGIN
    FR_CHECK(plan && a0 && a1, "null argument");
    FR_CHECK(m > 0 && m <= 80 && nupts > 0 && nfpts > 0,
             "invalid operator shape (at most 80 rows)");
    FR_CHECK(dtype == FRHIP_F64 || dtype == FRHIP_F32, "invalid dtype");
    const int k0 = 3*nupts, nq = (nupts + 3)/4, ks0 = 3*nq;
    const int ks1 = (nfpts + 3)/4, KS = ks0 + ks1, mt = (m + 15)/16;
    std::vector<double> frag((size_t)mt*KS*64, 0.0);
    for (int t = 0; t < mt; ++t)
        for (int s = 0; s < KS; ++s)
            for (int l = 0; l < 64; ++l) {
                const int r = 16*t + (l & 15), kr = l >> 4;
                double v = 0.0;
                if (r < m) {
                    if (s < ks0) {
                        const int q = s/3, d = s % 3, j = 4*q + kr;
                        if (j < nupts) v = a0[(long)r*k0 + d*nupts + j];
                    }
                    else {
                        const int kp = 4*(s - ks0) + kr;
                        if (kp < nfpts) v = a1[(long)r*nfpts + kp];
                    }
                }
                frag[((size_t)t*KS + s)*64 + l] = v;
            }
    const size_t esz = dtype == FRHIP_F64 ? 8 : 4;
    std::vector<char> buf(frag.size()*esz);
    for (size_t i = 0; i < frag.size(); ++i) {
        if (esz == 8) std::memcpy(&buf[i*8], &frag[i], 8);
        else {
            const float f = (float)frag[i];
            std::memcpy(&buf[i*4], &f, 4);
        }
    }
    void *d = nullptr;
    FR_HIP(hipMalloc(&d, buf.size()));
    hipError_t e = hipMemcpy(d, buf.data(), buf.size(),
                             hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        (void)hipFree(d);
        FR_HIP(e);
    }
    // 4x4x1 fragments (FP32): chunk q < nc0 holds points 5q..5q+4 x 3
    // directions (k = 3 point + direction), chunk nc0 + q fn rows 16q..+15
    void *d4 = nullptr;
    const int g4 = (m + 3)/4, nc0 = (nupts + 4)/5, nc1 = (nfpts + 15)/16;
    if (dtype == FRHIP_F32 && flux4_ok(g4)) {
        std::vector<float> f4((size_t)(nc0 + nc1)*g4*64, 0.0f);
        for (int q = 0; q < nc0 + nc1; ++q)
            for (int g = 0; g < g4; ++g)
                for (int l = 0; l < 64; ++l) {
                    const int r = 4*g + (l & 3), kk = l >> 2;
                    double v = 0.0;
                    if (r < m && q < nc0 && kk < 15) {
                        const int j = 5*q + kk/3, dd = kk % 3;
                        if (j < nupts) v = a0[(long)r*k0 + dd*nupts + j];
                    }
                    else if (r < m && q >= nc0) {
                        const int kp = 16*(q - nc0) + kk;
                        if (kp < nfpts) v = a1[(long)r*nfpts + kp];
                    }
                    f4[((size_t)q*g4 + g)*64 + l] = (float)v;
                }
        hipError_t e4 = hipMalloc(&d4, f4.size()*4);
        if (e4 == hipSuccess)
            e4 = hipMemcpy(d4, f4.data(), f4.size()*4,
                           hipMemcpyHostToDevice);
        if (e4 != hipSuccess) {
            (void)hipFree(d);
            if (d4) (void)hipFree(d4);
            FR_HIP(e4);
        }
    }
    FluxPlan *p = new FluxPlan{{m, k0, nfpts, dtype, mt, ks0, ks1, d, nullptr,
                                4}, nupts, nq, d4, g4, nc0, nc1};
    *plan = p;
    return 0;
    FR_API_END
}

int frhip_tdivf_plan_destroy(void *plan)
{
    FR_API_BEGIN
    FluxPlan *p = static_cast<FluxPlan *>(plan);
    if (p) {
        if (p->d.afrag) (void)hipFree(p->d.afrag);
        if (p->af4) (void)hipFree(p->af4);
        delete p;
    }
    return 0;
    FR_API_END
}

int frhip_tdivf_flux_euler(void *plan, int ncol, const void *u, long ldu,
                           const void *smats, long lds, const void *fn,
                           long ldf, void *c, long ldc, int ep,
                           const void *rcp, long ldr, void *xa, double dta,
                           double dtb, int last, double gamma, void *stream)
{
    FR_API_BEGIN
    FluxPlan *p = static_cast<FluxPlan *>(plan);
    FR_CHECK(p, "null plan");
    FR_CHECK((ep >= 0 && ep <= 3) || (ep >= 10 && ep <= 13),
             "invalid epilogue");
    FR_CHECK(ep != 3 || p->d.dtype == FRHIP_F32,
             "the mixed RK epilogue needs FP32 operands");
    FR_CHECK(ep < 10 || xa, "the RK4 epilogues need the {r1, r0} pair");
    if (ncol == 0) return 0;
    const int epi = (ep >= 2 && ep <= 3 && last) ? ep + 2 : ep;
    return flux_launch(p, epi, ncol, u, ldu, smats, lds, fn, ldf, c, ldc, rcp,
                       ldr, xa, dta, dtb, gamma, as_stream(stream));
    FR_API_END
}

int frhip_tdivf_plan_tune(void *plan, int ncol, const void *u, long ldu,
                          const void *smats, long lds, const void *fn,
                          long ldf, double gamma, int reps, int *variant)
{
    FR_API_BEGIN
    FluxPlan *p = static_cast<FluxPlan *>(plan);
    FR_CHECK(p && reps > 0, "invalid arguments");
    const size_t esz = p->d.dtype == FRHIP_F64 ? 8 : 4;
    void *c = nullptr;
    FR_HIP(hipMalloc(&c, (size_t)p->d.m*ncol*esz));
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    float best = 1e30f;
    int bestv = 4, rc = 0;
    const size_t ldsb = (size_t)p->d.mt*(p->d.ks0 + p->d.ks1)*64*esz;
    for (int v : {1, 2, 4, FR_LDSV + 1, FR_LDSV + 2, FR_LDSV + 4,
                  FR_BLKV, FR_V4, FR_V4B}) {
        if (v == FR_V4 || v == FR_V4B
                ? !p->af4
                : v != FR_BLKV && !variant_valid(v, ldsb, false))
            continue;
        FluxPlan q = *p;
        q.d.variant = v;
        for (int it = 0; it < reps + 1 && !rc; ++it) {
            if (it == 1) (void)hipEventRecord(e0, nullptr);
            rc = flux_launch(&q, 0, ncol, u, ldu, smats, lds, fn, ldf, c,
                             ncol, nullptr, 0, nullptr, 0, 0, gamma, nullptr);
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
    p->d.variant = bestv;
    if (variant) *variant = bestv;
    return 0;
    FR_API_END
}

int frhip_tdivf_plan_set_variant(void *plan, int variant)
{
    FR_API_BEGIN
    FluxPlan *p = static_cast<FluxPlan *>(plan);
    FR_CHECK(p, "null plan");
    const size_t esz = p->d.dtype == FRHIP_F64 ? 8 : 4;
    FR_CHECK(variant == FR_BLKV ||
             ((variant == FR_V4 || variant == FR_V4B) && p->af4) ||
             variant_valid(variant,
                           (size_t)p->d.mt*(p->d.ks0 + p->d.ks1)*64*esz,
                           false), "invalid variant");
    p->d.variant = variant;
    return 0;
    FR_API_END
}

}  // extern "C"
