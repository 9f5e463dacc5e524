// BLAS extensions beyond the plain axnpby/rkvdh2 kernels: the reduction
// provider (expressions of several arrays reduced per variable), axnpby with
// per-variable input/output scales, and rkvdh2 with its embedded error
// estimate.
//
// Reductions accumulate in double whatever the array precision and reduce in
// a fixed two-pass order (per-block partials, then one block per
// (expression, variable)), so a result does not depend on scheduling.  The
// partials live in a caller-supplied workspace, so a call allocates nothing
// and can be captured in a graph.
#include "frhip_common.h"
#include "frhip_expr.h"
#include "frhip_physics.h"

#include <cfloat>
#include <mutex>

namespace frhip {

constexpr int RED_MAXMAT = 8;     // arrays per reduction
constexpr int RED_MAXSLOT = 16;   // expression variables per point
constexpr int RED_MAXSV = 8;      // run-time scalars
constexpr int RED_MAXOUT = 4096;  // nexprs*nvar

struct ExprProgram {
    int nexprs, nslots, ncode, npool;
    int32_t *code, *offs;  // device
    double *pool;          // device
};

template <typename T>
struct RedArgs {
    const T *m[RED_MAXMAT];
    long ld[RED_MAXMAT];
    double sv[RED_MAXSV];
};

// Partials: grid (nblk, nvar), 256 threads.  Point q = r*neles + e of
// variable v = blockIdx.y; the expression variables are, in order, the nmat
// arrays at that point (an array whose bit is set in bcast has one value per
// (row, element), shared by all variables), the npv per-variable constants
// pv[k*nvar + v] and the nsv scalars.
template <typename T, bool MAX>
__global__ __launch_bounds__(256) void k_reduction_part(
    int nrow, int neles, int nvar, int nmat, unsigned bcast,
    RedArgs<T> a, const double *__restrict__ pv, int npv, int nsv,
    int nexprs, const int32_t *__restrict__ code,
    const int32_t *__restrict__ offs, const double *__restrict__ pool,
    double *__restrict__ part)
{
    const int v = blockIdx.y, nblk = gridDim.x;
    double var[RED_MAXSLOT];
    const int nslot = nmat + npv + nsv;
    for (int k = 0; k < npv; ++k) var[nmat + k] = pv[k*nvar + v];
    for (int k = 0; k < nsv; ++k) var[nmat + npv + k] = a.sv[k];

    double acc[RED_MAXSV];
    for (int x = 0; x < nexprs; ++x) acc[x] = MAX ? -DBL_MAX : 0.0;

    const long total = (long)nrow*neles;
    for (long q = (long)blockIdx.x*blockDim.x + threadIdx.x; q < total;
         q += (long)nblk*blockDim.x) {
        const long r = q/neles, e = q - r*neles;
        for (int m = 0; m < nmat; ++m)
            var[m] = (double)a.m[m][r*a.ld[m] + (((bcast >> m) & 1)
                                                 ? e : soa_ix(e, v, nvar))];
        for (int x = 0; x < nexprs; ++x) {
            const double y = expr_eval(code, offs[x], offs[x + 1], pool, var,
                                       nslot);
            acc[x] = MAX ? fmax(acc[x], y) : acc[x] + y;
        }
    }

    __shared__ double sm[256];
    for (int x = 0; x < nexprs; ++x) {
        sm[threadIdx.x] = acc[x];
        __syncthreads();
        for (int s = 128; s > 0; s >>= 1) {
            if ((int)threadIdx.x < s)
                sm[threadIdx.x] = MAX ? fmax(sm[threadIdx.x],
                                             sm[threadIdx.x + s])
                                      : sm[threadIdx.x] + sm[threadIdx.x + s];
            __syncthreads();
        }
        if (threadIdx.x == 0)
            part[((long)x*nvar + v)*nblk + blockIdx.x] = sm[0];
        __syncthreads();
    }
}

// One block per (expression, variable) output: out[x*nvar + v].
template <bool MAX>
__global__ __launch_bounds__(256) void k_reduction_final(
    int nblk, const double *__restrict__ part, double *__restrict__ out)
{
    const int o = blockIdx.x;
    double acc = MAX ? -DBL_MAX : 0.0;
    for (int q = threadIdx.x; q < nblk; q += blockDim.x) {
        const double y = part[(long)o*nblk + q];
        acc = MAX ? fmax(acc, y) : acc + y;
    }
    __shared__ double sm[256];
    sm[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s)
            sm[threadIdx.x] = MAX ? fmax(sm[threadIdx.x], sm[threadIdx.x + s])
                                  : sm[threadIdx.x] + sm[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[o] = sm[0];
}

// Number of partial blocks per variable for nrow*neles points.
inline int red_nblk(long total)
{
    const long b = (total + 255)/256;
    return (int)(b < FRHIP_RED_NBLK ? (b > 0 ? b : 1) : FRHIP_RED_NBLK);
}

// Argument checks shared by the workspace and workspace-free entry points.
inline const char *red_check(const ExprProgram *p, int dtype, int rop,
                             int nrow, int neles, int nvar, int nmat,
                             void *const *mats, const long *lds,
                             unsigned bcast, const double *pv, int npv,
                             const double *sv, int nsv, double *out)
{
    if (!p) return "null program";
    if (dtype != FRHIP_F64 && dtype != FRHIP_F32) return "invalid dtype";
    if (rop != FRHIP_RED_SUM && rop != FRHIP_RED_MAX)
        return "invalid reduction op";
    if (nrow <= 0 || neles < 0 || nvar <= 0) return "invalid extents";
    if (nmat < 1 || nmat > RED_MAXMAT || !mats || !lds)
        return "a reduction takes 1..8 arrays";
    if (npv < 0 || nsv < 0 || nsv > RED_MAXSV || (npv && !pv) ||
        (nsv && !sv))
        return "invalid per-variable constants or scalars";
    if (nmat + npv + nsv != p->nslots)
        return "variable count does not match the program";
    if (p->nexprs > RED_MAXSV || (long)p->nexprs*nvar > RED_MAXOUT)
        return "too many expressions or variables";
    if (bcast >> nmat) return "broadcast mask names a missing array";
    for (int m = 0; m < nmat; ++m) {
        if (!mats[m]) return "null array";
        const long need = ((bcast >> m) & 1) ? neles
                        : (long)((neles + 63)/64)*64*nvar;
        if (lds[m] < need) return "leading dimension too small";
    }
    if (!out) return "null output";
    return nullptr;
}

template <typename T>
void red_launch(const ExprProgram *p, int rop, int nrow, int neles, int nvar,
                int nmat, void *const *mats, const long *lds, unsigned bcast,
                const double *pv, int npv, const double *sv, int nsv,
                double *work, double *out, hipStream_t s)
{
    RedArgs<T> a{};
    for (int m = 0; m < nmat; ++m) {
        a.m[m] = static_cast<const T *>(mats[m]);
        a.ld[m] = lds[m];
    }
    for (int k = 0; k < nsv; ++k) a.sv[k] = sv[k];
    const int nblk = red_nblk((long)nrow*neles);
    const dim3 grid(nblk, nvar), block(256);
    const int nout = p->nexprs*nvar;
    if (rop == FRHIP_RED_MAX) {
        hipLaunchKernelGGL((k_reduction_part<T, true>), grid, block, 0, s,
                           nrow, neles, nvar, nmat, bcast, a, pv, npv, nsv,
                           p->nexprs, p->code, p->offs, p->pool, work);
        hipLaunchKernelGGL((k_reduction_final<true>), dim3(nout), block, 0,
                           s, nblk, work, out);
    }
    else {
        hipLaunchKernelGGL((k_reduction_part<T, false>), grid, block, 0, s,
                           nrow, neles, nvar, nmat, bcast, a, pv, npv, nsv,
                           p->nexprs, p->code, p->offs, p->pool, work);
        hipLaunchKernelGGL((k_reduction_final<false>), dim3(nout), block, 0,
                           s, nblk, work, out);
    }
}

// axnpby with per-variable scales (pyfr/backends/hip/kernels/axnpby.mako:
// 4-44 with pyfr/backends/base/makoutil.py:40-54 axnpby_expr): operand l of
// variable k is multiplied by insc[k] if bit l of inmask is set, and the sum
// by outsc[k] if hasout; the a0 == 0 / a0 == 1 special cases are kept.
constexpr int AX_MAXV = 16;

template <typename T>
struct AxScaled {
    void *p[8];
    T a[8];
    T in[AX_MAXV], out[AX_MAXV];
};

template <typename T>
__global__ __launch_bounds__(256) void k_axnpby_scaled(
    long ncol, long ld, int nv, int nvar, unsigned inmask, int hasout,
    AxScaled<T> c)
{
    const long j = (long)blockIdx.x*blockDim.x + threadIdx.x;
    if (j >= ncol) return;
    const long idx = (long)blockIdx.y*ld + j;
    const int k = (int)((j >> 6) % nvar);
    T *x0 = static_cast<T *>(c.p[0]);
    const int start = (c.a[0] == T(0) || c.a[0] == T(1)) ? 1 : 0;
    T acc = T(0);
    for (int l = start; l < nv; ++l) {
        const T xv = l == 0 ? x0[idx] : static_cast<const T *>(c.p[l])[idx];
        const T t = ((inmask >> l) & 1) ? (c.a[l])*c.in[k]*(xv)
                                        : (c.a[l])*(xv);
        acc = l == start ? t : acc + t;
    }
    if (hasout) acc = c.out[k]*(acc);
    if (start == 1 && c.a[0] == T(1))
        x0[idx] += acc;
    else
        x0[idx] = acc;
}

// rkvdh2 with the embedded error estimate; pyfr/integrators/explicit/
// kernels/rkvdh2.mako:4-31 (errest): stage 0 sets rold = r1 and
// rerr = dt*e*r2; later stages add dt*e*r2 to rerr.  The products are formed
// as the template writes them, (dt*c)*x in the working precision.
template <typename T>
__global__ __launch_bounds__(256) void k_rkvdh2_errest(
    long ncol, long ld, T *__restrict__ r1, T *__restrict__ r2,
    T *__restrict__ rold, T *__restrict__ rerr, T dt, T a, T b, T e,
    int stage, int last)
{
    const long j = (long)blockIdx.x*blockDim.x + threadIdx.x;
    if (j >= ncol) return;
    const long idx = (long)blockIdx.y*ld + j;
    const T t1 = r1[idx], t2 = r2[idx];
    if (stage == 0) {
        rerr[idx] = dt*e*t2;
        rold[idx] = t1;
    }
    else
        rerr[idx] = rerr[idx] + dt*e*t2;
    if (!last) {
        r1[idx] = t1 + dt*a*t2;
        r2[idx] = t1 + dt*b*t2;
    }
    else
        r1[idx] = t1 + dt*b*t2;
}

// RK4 stage update of the precision form, as the fused epilogues 14..17 of
// frhip_mul2 do it: k (FP32) is the stage RHS, r1 and r0 FP64, nxt (FP32)
// receives the next stage's input; the updates of pyfr/integrators/
// explicit/steppers.py:67-108 formed in FP64, operands in the same order.
__global__ __launch_bounds__(256) void k_rk4_mixed(
    long ncol, long ld, int stage, double *__restrict__ r1,
    const double *__restrict__ r0, const float *__restrict__ k,
    float *__restrict__ nxt, double dt)
{
    const long j = (long)blockIdx.x*blockDim.x + threadIdx.x;
    if (j >= ncol) return;
    const long idx = (long)blockIdx.y*ld + j;
    const double td = (double)k[idx], z = r0[idx];
    if (stage == 1) {
        r1[idx] = td;
        double a = 1.0*z;
        a = a + (dt/2.0)*td;
        nxt[idx] = (float)a;
    }
    else if (stage == 2) {
        double a = (dt/6.0)*r1[idx];
        a = a + 1.0*z;
        a = a + (dt/3.0)*td;
        r1[idx] = a;
        double b = (dt/2.0)*td;
        b = b + 1.0*z;
        nxt[idx] = (float)b;
    }
    else if (stage == 3) {
        r1[idx] = r1[idx] + (dt/3.0)*td;
        double b = dt*td;
        b = b + 1.0*z;
        nxt[idx] = (float)b;
    }
    else {
        const double na = r1[idx] + (dt/6.0)*td;
        r1[idx] = na;
        nxt[idx] = (float)na;
    }
}

// Scratch of the workspace-free reduction: the partials of up to
// RED_NOALLOC_MAXOUT (expression, variable) outputs, 256 KiB, held by every
// device the library runs on until the process ends.  Host code never uses
// the array's name: each device's address comes from hipGetSymbolAddress
// with that device current, once (frhip_init, or the first call on the
// device), and is kept in red_scratch_ptr.
constexpr int RED_NOALLOC_MAXOUT = 64;
constexpr int RED_MAXDEV = 64;
static __device__ __align__(256)
    double scratch[RED_NOALLOC_MAXOUT*FRHIP_RED_NBLK];
static double *red_scratch_ptr[RED_MAXDEV];
static std::mutex red_scratch_mutex;

int red_scratch_init(int device)
{
    if (device < 0 || device >= RED_MAXDEV) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(red_scratch_mutex);
    if (red_scratch_ptr[device]) return hipSuccess;
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    red_scratch_ptr[device] = static_cast<double *>(ptr);
    return hipSuccess;
}

}  // namespace frhip

using namespace frhip;

extern "C" {

int frhip_expr_program_create(void **prog, int nexprs, int nslots,
                              const int32_t *code, const int32_t *offs,
                              const double *pool, int npool)
{
    FR_API_BEGIN
    FR_CHECK(prog && code && offs && nexprs > 0 && nexprs <= RED_MAXSV &&
             nslots > 0 && nslots <= RED_MAXSLOT && npool >= 0,
             "invalid program arguments");
    FR_CHECK(npool == 0 || pool, "missing constant pool");
    {
        const char *err = expr_validate(nexprs, code, offs, npool, nslots);
        FR_CHECK(!err, err ? err : "");
    }
    auto *p = new ExprProgram{nexprs, nslots, offs[nexprs], npool, nullptr,
                              nullptr, nullptr};
    const size_t nc = (size_t)2*p->ncode*sizeof(int32_t);
    const size_t no = (size_t)(nexprs + 1)*sizeof(int32_t);
    const size_t np = (size_t)(npool > 0 ? npool : 1)*sizeof(double);
    if (hipMalloc(&p->code, nc) != hipSuccess ||
        hipMalloc(&p->offs, no) != hipSuccess ||
        hipMalloc(&p->pool, np) != hipSuccess ||
        hipMemcpy(p->code, code, nc, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(p->offs, offs, no, hipMemcpyHostToDevice) != hipSuccess ||
        (npool > 0 && hipMemcpy(p->pool, pool, npool*sizeof(double),
                                hipMemcpyHostToDevice) != hipSuccess)) {
        (void)hipFree(p->code);
        (void)hipFree(p->offs);
        (void)hipFree(p->pool);
        delete p;
        FR_CHECK(false, "device allocation for an expression program failed");
    }
    *prog = p;
    return 0;
    FR_API_END
}

int frhip_expr_program_destroy(void *prog)
{
    FR_API_BEGIN
    auto *p = static_cast<ExprProgram *>(prog);
    if (!p) return 0;
    (void)hipFree(p->code);
    (void)hipFree(p->offs);
    (void)hipFree(p->pool);
    delete p;
    return 0;
    FR_API_END
}

int frhip_reduction(void *prog, int dtype, int rop, int nrow, int neles,
                    int nvar, int nmat, void *const *mats, const long *lds,
                    unsigned bcast, const double *pv, int npv,
                    const double *sv, int nsv, double *work, long nwork,
                    double *out, void *stream)
{
    FR_API_BEGIN
    auto *p = static_cast<ExprProgram *>(prog);
    const char *err = red_check(p, dtype, rop, nrow, neles, nvar, nmat, mats,
                                lds, bcast, pv, npv, sv, nsv, out);
    FR_CHECK(!err, err ? err : "");
    const long need = (long)p->nexprs*nvar*red_nblk((long)nrow*neles);
    FR_CHECK(work && nwork >= need, "reduction workspace too small");
    if (dtype == FRHIP_F64)
        red_launch<double>(p, rop, nrow, neles, nvar, nmat, mats, lds, bcast,
                           pv, npv, sv, nsv, work, out, as_stream(stream));
    else
        red_launch<float>(p, rop, nrow, neles, nvar, nmat, mats, lds, bcast,
                          pv, npv, sv, nsv, work, out, as_stream(stream));
    FR_LAUNCHED();
    return 0;
    FR_API_END
}

int frhip_reduction_noalloc(void *prog, int dtype, int rop, int nrow,
                            int neles, int nvar, int nmat, void *const *mats,
                            const long *lds, unsigned bcast, const double *pv,
                            int npv, const double *sv, int nsv, double *out,
                            void *stream)
{
    FR_API_BEGIN
    auto *p = static_cast<ExprProgram *>(prog);
    const char *err = red_check(p, dtype, rop, nrow, neles, nvar, nmat, mats,
                                lds, bcast, pv, npv, sv, nsv, out);
    FR_CHECK(!err, err ? err : "");
    FR_CHECK((long)p->nexprs*nvar <= RED_NOALLOC_MAXOUT,
             "frhip_reduction_noalloc: nexprs*nvar exceeds 64, the size of "
             "its scratch; use frhip_reduction with a workspace");
    int dev = 0;
    FR_HIP(hipGetDevice(&dev));
    FR_CHECK(dev >= 0 && dev < RED_MAXDEV, "device index out of range");
    // set-up on a device frhip_init was not called for
    FR_HIP((hipError_t)red_scratch_init(dev));
    double *work = red_scratch_ptr[dev];
    FR_CHECK(work, "no scratch address for this device");
    if (dtype == FRHIP_F64)
        red_launch<double>(p, rop, nrow, neles, nvar, nmat, mats, lds, bcast,
                           pv, npv, sv, nsv, work, out, as_stream(stream));
    else
        red_launch<float>(p, rop, nrow, neles, nvar, nmat, mats, lds, bcast,
                          pv, npv, sv, nsv, work, out, as_stream(stream));
    FR_LAUNCHED();
    return 0;
    FR_API_END
}

int frhip_reduction_workspace(void *prog, int nrow, int neles, int nvar,
                              long *nwork)
{
    FR_API_BEGIN
    auto *p = static_cast<ExprProgram *>(prog);
    FR_CHECK(p && nwork && nrow > 0 && neles >= 0 && nvar > 0,
             "invalid arguments");
    *nwork = (long)p->nexprs*nvar*red_nblk((long)nrow*neles);
    return 0;
    FR_API_END
}

int frhip_axnpby_scaled(int dtype, int nrow, long ncol, long ld, int nv,
                        int nvar, void *const *xs, const double *a,
                        unsigned inmask, const double *insc,
                        const double *outsc, void *stream)
{
    FR_API_BEGIN
    FR_DTYPE_OK(dtype);
    FR_CHECK(nv >= 1 && nv <= 8 && xs && a, "axnpby takes 1..8 operands");
    FR_CHECK(nvar >= 1 && nvar <= AX_MAXV, "1..16 variables");
    FR_CHECK(nrow > 0 && nrow < 65536, "invalid row count");
    FR_CHECK(ncol % (SOA*nvar) == 0, "ncol must cover whole SoA blocks");
    FR_CHECK((inmask >> nv) == 0, "input-scale mask names a missing operand");
    FR_CHECK(!inmask || insc, "missing input scales");
    if (ncol == 0) return 0;
    dim3 grid(cdiv(ncol, 256), nrow), block(256);
    auto go = [&](auto tag) {
        using T = decltype(tag);
        AxScaled<T> c{};
        for (int l = 0; l < nv; ++l) {
            c.p[l] = xs[l];
            c.a[l] = (T)a[l];
        }
        for (int k = 0; k < nvar; ++k) {
            c.in[k] = insc ? (T)insc[k] : T(1);
            c.out[k] = outsc ? (T)outsc[k] : T(1);
        }
        hipLaunchKernelGGL(k_axnpby_scaled<T>, grid, block, 0,
                           as_stream(stream), ncol, ld, nv, nvar, inmask,
                           outsc ? 1 : 0, c);
    };
    if (dtype == FRHIP_F64) go(double());
    else go(float());
    FR_LAUNCHED();
    return 0;
    FR_API_END
}

int frhip_rk4_mixed(int nrow, long ncol, long ld, int stage, void *r1,
                    const void *r0, const void *k, void *nxt, double dt,
                    void *stream)
{
    FR_API_BEGIN
    FR_CHECK(nrow > 0 && nrow < 65536, "invalid row count");
    FR_CHECK(stage >= 1 && stage <= 4, "RK4 stages are 1..4");
    FR_CHECK(r1 && r0 && k && nxt, "null register");
    if (ncol == 0) return 0;
    dim3 grid(cdiv(ncol, 256), nrow), block(256);
    hipLaunchKernelGGL(k_rk4_mixed, grid, block, 0, as_stream(stream), ncol,
                       ld, stage, (double *)r1, (const double *)r0,
                       (const float *)k, (float *)nxt, dt);
    FR_LAUNCHED();
    return 0;
    FR_API_END
}

int frhip_rkvdh2_errest(int dtype, int nrow, long ncol, long ld, void *r1,
                        void *r2, void *rold, void *rerr, double dt,
                        double a, double b, double e, int stage, int last,
                        void *stream)
{
    FR_API_BEGIN
    FR_DTYPE_OK(dtype);
    FR_CHECK(nrow > 0 && nrow < 65536, "invalid row count");
    FR_CHECK(r1 && r2 && rold && rerr, "null register");
    FR_CHECK(stage >= 0, "invalid stage");
    if (ncol == 0) return 0;
    dim3 grid(cdiv(ncol, 256), nrow), block(256);
    if (dtype == FRHIP_F64)
        hipLaunchKernelGGL(k_rkvdh2_errest<double>, grid, block, 0,
                           as_stream(stream), ncol, ld, (double *)r1,
                           (double *)r2, (double *)rold, (double *)rerr, dt,
                           a, b, e, stage, last);
    else
        hipLaunchKernelGGL(k_rkvdh2_errest<float>, grid, block, 0,
                           as_stream(stream), ncol, ld, (float *)r1,
                           (float *)r2, (float *)rold, (float *)rerr,
                           (float)dt, (float)a, (float)b, (float)e, stage,
                           last);
    FR_LAUNCHED();
    return 0;
    FR_API_END
}

}  // extern "C"
