/*
 * frhip.h -- C ABI of the MI355X (gfx950) flux-reconstruction kernel library.
 *
 * Conventions (mirroring the ctypes LibWrapper contract of the reference,
 * pyfr/ctypesutil.py:8-40 and pyfr/backends/hip/driver.py:74-153):
 *   - every function returns int: 0 on success, nonzero on failure;
 *     frhip_last_error() returns a thread-local message for the last failure;
 *   - no C++ exceptions cross the ABI;
 *   - device buffers are plain pointers owned by the caller; the library
 *     never frees caller memory; operator plans own their constant data;
 *   - streams, events and graphs are opaque void* handles;
 *   - dtype selects the working precision: FRHIP_F64 or FRHIP_F32.
 *
 * Storage layout (the reference's HIP layout, pyfr/backends/base/types.py:40-100
 * with soasz = csubsz = 64 = the gfx950 wavefront): a (nrow, nvar, neles)
 * array is row-major with leading dimension ld = nvar*neles_pad and element
 * e / variable v of row r at  r*ld + ((e/64)*nvar + v)*64 + e%64.
 * A 2-D (nrow, ncol) array is row-major with leading dimension ld.
 */
#ifndef FRHIP_H
#define FRHIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { FRHIP_F64 = 0, FRHIP_F32 = 1 };
/* Operator plans only: FP32 data with both matrix-core operands rounded to
 * bfloat16 and FP32 accumulation (the precision form's BF16 experiment). */
enum { FRHIP_F32_BF16OPS = 2 };
/* Riemann solvers (pyfr/solvers/euler/kernels/rsolvers/<name>.mako). */
enum { FRHIP_RS_RUSANOV = 0, FRHIP_RS_HLLC = 1, FRHIP_RS_HLL = 2,
       FRHIP_RS_ROE = 3, FRHIP_RS_ROEM = 4, FRHIP_RS_EXACT = 5 };

/* ---- runtime: replaces the ctypes HIP driver wrapper
 *      pyfr/backends/hip/driver.py:402-482 (hipSetDevice, hipMalloc, hipFree,
 *      hipMemcpy*, hipMemset*, streams, events) ---- */
const char *frhip_last_error(void);
int frhip_init(int device);
int frhip_device_info(char *name, size_t len, size_t *free_b, size_t *total_b);
int frhip_malloc(void **ptr, size_t nbytes);
int frhip_free(void *ptr);
int frhip_host_malloc(void **ptr, size_t nbytes);
int frhip_host_free(void *ptr);
int frhip_memcpy_htod(void *dst, const void *src, size_t nbytes, void *stream);
int frhip_memcpy_dtoh(void *dst, const void *src, size_t nbytes, void *stream);
/* n doubles at ptr set to v by a kernel on `stream` (no memset node when
 * captured): device-resident run-time arguments such as the RK4 step read
 * by the fused epilogues. */
int frhip_fill_f64(void *ptr, long n, double v, void *stream);
/* Device to page-locked host memory, not waited on (graph-capturable); the
 * data is there once the stream has been synchronised. */
int frhip_memcpy_dtoh_async(void *dst, const void *src, size_t nbytes,
                            void *stream);
int frhip_memcpy_dtod(void *dst, const void *src, size_t nbytes, void *stream);
int frhip_memset(void *ptr, int value, size_t nbytes, void *stream);
/* 2-D copies (hipMemcpy2DAsync): height rows of width bytes between pitched
 * buffers; kind 0 = host to device, 1 = device to host, 2 = device to
 * device.  Host-side copies are synchronised before returning. */
int frhip_memcpy2d(void *dst, size_t dpitch, const void *src, size_t spitch,
                   size_t width, size_t height, int kind, void *stream);
int frhip_stream_create(void **stream);
int frhip_stream_destroy(void *stream);
int frhip_stream_sync(void *stream);
int frhip_device_sync(void);
int frhip_event_create(void **event);
int frhip_event_destroy(void *event);
int frhip_event_record(void *event, void *stream);
/* Work later enqueued on stream waits for event (hipStreamWaitEvent).  On a
 * capturing stream's event this joins the capture: the fork and join edges
 * of a graph's independent kernel chains, in place of the explicit node
 * dependencies of the reference's graph builder
 * (pyfr/backends/hip/driver.py:303-399). */
int frhip_stream_wait_event(void *stream, void *event);
int frhip_event_sync(void *event);
/* *done = 1 if all work before the event has completed, else 0. */
int frhip_event_query(void *event, int *done);
int frhip_event_elapsed(float *ms, void *start, void *end);

/* ---- graphs: replaces HIPGraph (pyfr/backends/hip/types.py:79-116) and the
 *      graph API of pyfr/backends/hip/driver.py:303-399.  Kernels launched on
 *      a capturing stream are recorded, then replayed by frhip_graph_launch. */
int frhip_graph_begin(void *stream);
int frhip_graph_end(void *stream, void **graph_exec);
int frhip_graph_launch(void *graph_exec, void *stream);
int frhip_graph_destroy(void *graph_exec);

/* ---- operator application: replaces the `mul` providers
 *      HIPGiMMiKKernels.mul (pyfr/backends/hip/gimmik.py:24-117) and
 *      HIPRocBLASKernels.mul (pyfr/backends/hip/rocblas.py:93-192):
 *      C(m, ncol) = alpha * A(m, k) @ B(k, ncol) + beta * C.
 * frhip_mul_plan_create takes the dense row-major host operator A.  If its
 * sparsity pattern and values match a compiled structured kernel (hex tensor-product
 * operators), that kernel is used with A's nonzeros uploaded; otherwise the
 * plan is dense (tet/prism operators): matrix-core (MFMA) kernels with the
 * operator in fragment order, or a scalar kernel.  *kind receives 1
 * (structured) or 2 (dense).  dtype FRHIP_F32_BF16OPS always makes a BF16
 * matrix-core plan (*kind = 3) over FP32 data. */
int frhip_mul_plan_create(void **plan, int m, int k, const double *a_host,
                          int dtype, int *kind);
int frhip_mul_plan_destroy(void *plan);
int frhip_mul(void *plan, int ncol, const void *b, long ldb, void *c,
              long ldc, double alpha, double beta, void *stream);

/* Autotune a dense plan for (ncol, ldb): times the scalar kernel and the
 * MFMA kernel at 1, 2 and 4 wavefronts per block, each with the operator
 * read through the L2 or staged in LDS once per block, on the given input,
 * writing only to a scratch output, and keeps the fastest (*variant: 0
 * scalar, 1/2/4 wavefronts per block, +16 for the LDS-staged form; -1 for
 * structured plans, which have no variants).
 * Synchronous (uses the null stream); the input must be ready.  Replaces
 * the reference's rocBLAS solution-index sweep (pyfr/backends/hip/
 * rocblas.py:139-177). */
int frhip_mul_plan_tune(void *plan, int ncol, const void *b, long ldb,
                        int reps, int *variant);

/* ---- fused operator applications: C = [A0 | A1] @ [B0; B1] followed by an
 *      epilogue ep: 0 = store; 1 = negdivconf (c = -rcp*acc, rcp (m, neles)
 *      with leading dimension ldr); 2 = negdivconf then the rkvdh2 stage on
 *      r1 = xa and r2 = c (r1' = r1 + dta*t, r2' = r1 + dtb*t; if last only
 *      r1' = r1 + dtb*t); 3 = as 2 with FP32 operands, an FP64 accumulated
 *      solution xa and the FP32 next-stage input written to c; 10..13 =
 *      negdivconf then RK4 stage ep - 9 with four registers, as for
 *      frhip_tdivf_flux_euler (xa = device pointers {r1, r0, dt}, dt a
 *      device double holding the step, so that a captured graph stays
 *      valid when it changes; dta unused; c the next stage's input;
 *      structured plans only, the updates of
 *      pyfr/integrators/explicit/steppers.py:67-108); 14..17 = 10..13 for
 *      the precision form (structured FP32 plans; no counterpart in the
 *      reference): r1 and r0 FP64, the updates formed in FP64 in the same
 *      order from the FP32 stage RHS, c the FP32 next stage input, and 17
 *      also stores the new solution rounded to FP32 into c; 20..23 = 10..13
 *      with a time step per element (structured FP64 plans; no counterpart
 *      in the reference, which has one dt per mesh): xa points to a device
 *      array of three pointers {r1, r0, dt}, dt holding one value per
 *      element (a pool of independent solves), and the coefficients are
 *      dt[e]/2, dt[e]/3, dt[e]/6 as the pool axnpby forms them.
 * They fuse kernels the reference runs back to back on the same element
 * data: tgradpcoru_upts + tgradcoru_upts (pyfr/solvers/baseadvecdiff/
 * elements.py:47-55) and tdivtpcorf + tdivtconf + negdivconf
 * (pyfr/solvers/baseadvec/elements.py:84-119) + rkvdh2
 * (pyfr/integrators/explicit/kernels/rkvdh2.mako:4-31).  A plan matches a
 * generated kernel by pattern and values (*kind = 1); otherwise it is a
 * dense plan (*kind = 2, as for frhip_mul).  ncol must be a multiple of 320
 * (whole 64-element blocks of 5 variables).  ldr = 0 gives every point of
 * an element the same rcp value (affine elements: one 1/|J| per element). */
int frhip_mul2_plan_create(void **plan, int m, int k0, int k1,
                           const double *a0, const double *a1, int dtype,
                           int *kind);
int frhip_mul2_plan_destroy(void *plan);
int frhip_mul2(void *plan, int ncol, const void *b0, long ld0, const void *b1,
               long ld1, void *c, long ldc, int ep, const void *rcp, long ldr,
               void *xa, double dta, double dtb, int last, void *stream);
/* As frhip_mul_plan_tune for a two-input plan. */
int frhip_mul2_plan_tune(void *plan, int ncol, const void *b0, long ld0,
                         const void *b1, long ld1, int reps, int *variant);
/* Pin a dense plan's variant (a code the tuners report), e.g.
 * from a stored tuning result.  Operands the MFMA kernel cannot take (ncol
 * not a multiple of 64, misaligned pointers) always use the scalar kernel. */
int frhip_mul_plan_set_variant(void *plan, int variant);
int frhip_mul2_plan_set_variant(void *plan, int variant);

/* ---- flux-fused divergence for straight-sided tets and prisms (Euler):
 *      tdisf (pyfr/solvers/euler/kernels/tflux.mako:8-30) + tdivtpcorf +
 *      tdivtconf + negdivconf (pyfr/solvers/baseadvec/elements.py:84-119)
 *      [+ rkvdh2 stage] in one matrix-core pass; the transformed flux is
 *      computed per lane from u and the element's S and never stored.
 * a0 = M1 - M3*M2 (m x 3 nupts), a1 = M3 (m x nfpts), m <= 80.
 * u (nupts, 5, neles) and fn (nfpts, 5, neles) AoSoA; smats (9, neles) with
 * leading dimension lds (row 3i + k = S[i][k]); ep and the RK arguments as
 * for frhip_mul2 (0 store, 1 negdivconf, 2 / 3 RK45 stage); ncol = 5 *
 * padded neles.  The tuner times 1, 2 and 4 wavefronts per block, with
 * and without the LDS-staged operator (variant codes as above, without the
 * scalar kernel), variant 32: one block of five wavefronts per
 * 64-element block, the flux evaluated once per point into LDS, and (FP32)
 * variant 64: the 4x4x1 MFMA form, 65: the same in blocks of the five
 * wavefronts of an element block (one workgroup, so one XCD's L2 serves
 * every variable's reads of u).  c may be u itself (an RHS in place):
 * such a call runs variant 32, whose block barriers order every read of u
 * before any write of c, whatever the plan's variant.
 * ep 10..13: RK4 stages 1..4 on four registers, with the update axnpby of
 * pyfr/integrators/explicit/steppers.py:67-108 in the epilogue (k = the
 * stage RHS; xa points to a device array of three pointers {r1, r0, dt},
 * dt a device double holding the step, read by the kernel; c receives the
 * next stage's input):
 *   10: r1 = k;                        c = r0 + dt/2 k
 *   11: r1 = dt/6 r1 + r0 + dt/3 k;    c = dt/2 k + r0
 *   12: r1 = r1 + dt/3 k;              c = dt k + r0
 *   13: r1 = r1 + dt/6 k               (c unused). */
int frhip_tdivf_plan_create(void **plan, int m, int nupts, int nfpts,
                            const double *a0, const double *a1, int dtype);
int frhip_tdivf_plan_destroy(void *plan);
int frhip_tdivf_plan_tune(void *plan, int ncol, const void *u, long ldu,
                          const void *smats, long lds, const void *fn,
                          long ldf, double gamma, int reps, int *variant);
int frhip_tdivf_plan_set_variant(void *plan, int variant);
int frhip_tdivf_flux_euler(void *plan, int ncol, const void *u, long ldu,
                           const void *smats, long lds, const void *fn,
                           long ldf, void *c, long ldc, int ep,
                           const void *rcp, long ldr, void *xa, double dta,
                           double dtb, int last, double gamma, void *stream);

/* ---- pointwise kernels (element-major 2-D over (npts, neles)) ---- */

/* Geometry modes of the tflux kernels (the `linear` argument). */
#define FRHIP_GEO_CURVED 0      /* stored smats (3, npts, 3, neles) */
#define FRHIP_GEO_HEX_LINEAR 1  /* S from the 8 hex vertices (8, 3, neles) */
#define FRHIP_GEO_AFFINE 2      /* one S per element: (9, neles), row 3i+k */

/* Euler tflux: pyfr/solvers/euler/kernels/tflux.mako:8-30 with flux.mako:3-26.
 * linear = FRHIP_GEO_HEX_LINEAR: S from the 8 vertices (upts: (npts, 3)
 * row-major, smats.mako:13-25); FRHIP_GEO_CURVED: stored smats;
 * FRHIP_GEO_AFFINE: per-element S of straight-sided tets/prisms.
 *
 * base (here and in frhip_tflux_ns, frhip_intcflux_euler,
 * frhip_intcflux_ns): NULL for the reference's kernels, or {rho0, E0} of a
 * constant state u0 = (rho0, 0, 0, 0, E0) at rest for the base-state form
 * (no counterpart in the reference): u then holds u - u0 and the fluxes
 * returned are f(u) - f(u0), f(u0) = (0, p0 I, 0), p0 = (gamma - 1) E0.  On
 * free-stream-preserving elements f(u0) has zero discrete divergence, so the
 * right-hand side is the same while E0 and p0 no longer take FP32 digits
 * (the precision form).  The interface kernels take it with the Rusanov
 * solver only. */
int frhip_tflux_euler(int dtype, int linear, int npts, int neles,
                      const void *u, long ldu, void *f, long ldf,
                      const void *geo, long ldgeo, const void *upts,
                      double gamma, const double *base, void *stream);

/* NS fused tflux: pyfr/solvers/navstokes/kernels/tflux.mako:13-50 with
 * transform_grad.mako:3-10 and navstokes/kernels/flux.mako:48-105.  gradu is
 * transformed in place to physical space; f = S (Fi + Fv).
 * For linear = 0 or 2, rcpdjac is (npts, neles) with leading dimension ldr
 * (ldr = 0: one value per element).
 * consts = {gamma, mu, Pr, beta, tau, sutherland(0/1), cpTref, cpTs}.
 * mu_e: optional per-element viscosity (neles values) replacing consts[1]
 * (used to batch independent solves); NULL for the reference behaviour. */
int frhip_tflux_ns(int dtype, int linear, int npts, int neles,
                   const void *u, long ldu, void *f, long ldf, void *gradu,
                   long ldg, const void *geo, long ldgeo, const void *upts,
                   const void *rcpdjac, long ldr, const double *consts,
                   const void *mu_e, const double *base, void *stream);

/* gradcoru: the corrected reference gradient (3 rows per point) made
 * physical in place, g <- |J|^-1 S^T g;
 * pyfr/solvers/baseadvecdiff/kernels/gradcoru.mako:10-23.  Geometry as for
 * frhip_tflux_ns. */
int frhip_gradcoru(int dtype, int linear, int npts, int neles, void *gradu,
                   long ldg, const void *geo, long ldgeo, const void *upts,
                   const void *rcpdjac, long ldr, void *stream);

/* NS tflux from a physical gradient (the unfused / anti-aliased path):
 * f holds the gradient on entry and the transformed flux on exit;
 * pyfr/solvers/navstokes/kernels/tflux.mako:13-50 without 'fused'.
 * consts and mu_e as for frhip_tflux_ns. */
int frhip_tflux_ns_grad(int dtype, int linear, int npts, int neles,
                        const void *u, long ldu, void *f, long ldf,
                        const void *geo, long ldgeo, const void *upts,
                        const double *consts, const void *mu_e,
                        void *stream);

/* negdivconf: pyfr/solvers/baseadvec/kernels/negdivconf.mako:7-22 (no source
 * terms): t[v] = -rcpdjac * t[v]; rcpdjac (npts, neles) with leading
 * dimension ldr, or ldr = 0 for one value per element. */
int frhip_negdivconf(int dtype, int npts, int neles, void *t, long ldt,
                     const void *rcpdjac, long ldr, void *stream);

/* Interface kernels, 1-D over n flux-point pairs.  Views are (base,
 * mapping) with variable stride 64; vector views add a row stride.
 * Euler intcflux: pyfr/solvers/euler/kernels/intcflux.mako:6-22 with
 * rsolvers/rusanov.mako:4-25 or rsolvers/hllc.mako:4-77. nl: (3, n), ld ldn. */
/* pmin (here and in frhip_intcflux_ns, frhip_bccflux_*): the exact solver's
 * floor for negative Newton iterates of the star pressure, [solver-interfaces]
 * p-min (pyfr/solvers/euler/inters.py:12-26); negative for the default,
 * 5 x machine epsilon of the working precision.  Other solvers ignore it. */
int frhip_intcflux_euler(int dtype, int rsolver, int n, void *base,
                         const int32_t *lmap, const int32_t *rmap,
                         const void *nl, long ldn, double gamma,
                         double pmin, const double *bstate, void *stream);

/* NS intconu: pyfr/solvers/navstokes/kernels/intconu.mako:4-19. */
int frhip_intconu(int dtype, int n, const void *ubase, const int32_t *ulin,
                  const int32_t *urin, void *cbase, const int32_t *ulout,
                  const int32_t *urout, double beta, void *stream);

/* NS intcflux: pyfr/solvers/navstokes/kernels/intcflux.mako:10-54.
 * gbase/glmap/grmap address the physical gradient at the flux points with row
 * stride gstride between dimensions, or, where glstr / grstr are given, the
 * per-point strides of the left / right view (n values each; a view over
 * several element types, whose gradient banks differ in row count: the
 * reference's per-point View.rstrides, pyfr/backends/base/types.py:313-320).
 * consts = {gamma, mu, Pr, beta, tau, sutherland(0/1), cpTref, cpTs}.  mu_p:
 * optional per-pair viscosity (n values), NULL for the reference
 * behaviour. */
int frhip_intcflux_ns(int dtype, int rsolver, int n, void *ubase,
                      const int32_t *lmap, const int32_t *rmap,
                      const void *gbase, const int32_t *glmap,
                      const int32_t *grmap, long gstride,
                      const int32_t *glstr, const int32_t *grstr,
                      const void *nl, long ldn, const double *consts,
                      double pmin, const void *mu_p, const double *bstate,
                      void *stream);

/* ---- boundary interfaces, 1-D over n boundary flux points ----
 * Types: the `type` keys of [soln-bcs-*] (pyfr/solvers/euler/inters.py:57-102,
 * pyfr/solvers/navstokes/inters.py:70-219); the first four exist for Euler and
 * NS, the rest for NS only. */
enum { FRHIP_BC_SUP_IN_FA = 0, FRHIP_BC_SUP_OUT_FN = 1,
       FRHIP_BC_CHAR_RIEM_INV = 2, FRHIP_BC_SLP_ADIA_WALL = 3,
       FRHIP_BC_NO_SLP_ADIA_WALL = 4, FRHIP_BC_NO_SLP_ISOT_WALL = 5,
       FRHIP_BC_SUB_IN_FRV = 6, FRHIP_BC_SUB_IN_FTPTTANG = 7,
       FRHIP_BC_SUB_OUT_FP = 8 };

/* bcv: the boundary's prescribed values per point (rows of n, leading dim
 * ldb, in dtype), in place of the reference's ploc-dependent template
 * expressions: sup-in-fa and char-riem-inv {rho, p, u, v, w}; sub-in-frv
 * {rho, u, v, w}; no-slp-isot-wall {u, v, w}; sub-out-fp {p}; none for the
 * other types (bcv may be NULL).  consts = the 8 constants of
 * frhip_intcflux_ns followed by {cpTw, cpTt, pt, vc0, vc1, vc2} (the
 * isothermal-wall temperature, and sub-in-ftpttang's total enthalpy, total
 * pressure and flow direction); the Euler kernel reads gamma only.
 *
 * Euler bccflux: pyfr/solvers/euler/kernels/bccflux.mako:7-25 with the
 * type's bc_rsolve_state (euler/kernels/bcs/<type>.mako). */
int frhip_bccflux_euler(int dtype, int bctype, int rsolver, int n,
                        void *base, const int32_t *lmap, const void *nl,
                        long ldn, const void *bcv, long ldb,
                        const double *consts, double pmin, void *stream);

/* NS bcconu: pyfr/solvers/navstokes/kernels/bcconu.mako:6-15 (the type's
 * bc_ldg_state, navstokes/kernels/bcs/<type>.mako), ulout = LDG state. */
int frhip_bcconu(int dtype, int bctype, int n, const void *ubase,
                 const int32_t *ulin, void *cbase, const int32_t *ulout,
                 const void *nl, long ldn, const void *bcv, long ldb,
                 const double *consts, void *stream);

/* NS bccflux: pyfr/solvers/navstokes/kernels/bccflux.mako:10-20 with the
 * type's common-flux state (bcs/ghost.mako, bcs/ghost-imperm.mako for the
 * no-slip walls, bcs/slp-adia-wall.mako:21-37).  Gradient view (gstride or
 * per-point glstr) as in frhip_intcflux_ns (unused, may be NULL, for
 * slp-adia-wall). */
int frhip_bccflux_ns(int dtype, int bctype, int rsolver, int n, void *ubase,
                     const int32_t *lmap, const void *gbase,
                     const int32_t *glmap, long gstride,
                     const int32_t *glstr, const void *nl, long ldn,
                     const void *bcv, long ldb, const double *consts,
                     double pmin, void *stream);

/* ---- boundary values that depend on time ----
 * The reference splices each [soln-bcs-*] value expression into the kernel
 * source and binds t at run time (pyfr/solvers/base/inters.py:54-71,
 * pyfr/solvers/euler/inters.py:57-102).  Here each expression is a stack
 * program of (op, arg) int32 pairs, evaluated in double at every boundary
 * point from its coordinates and t:
 *   PUSHC arg: push pool[arg];  PUSHV arg: push x, y, z or t (arg 0..3);
 *   binary ops pop b then a and push a op b; unary ops replace the top.
 * Functions are those the reference's expressions may use (pyfr/nputil.py
 * npeval: pow, min, max, atan2, sin, cos, tan, exp, log, sqrt, abs, tanh,
 * asin, acos, atan). */
enum { FRHIP_BCX_PUSHC = 0, FRHIP_BCX_PUSHV = 1, FRHIP_BCX_ADD = 2,
       FRHIP_BCX_SUB = 3, FRHIP_BCX_MUL = 4, FRHIP_BCX_DIV = 5,
       FRHIP_BCX_POW = 6, FRHIP_BCX_MIN = 7, FRHIP_BCX_MAX = 8,
       FRHIP_BCX_ATAN2 = 9, FRHIP_BCX_NEG = 10, FRHIP_BCX_SIN = 11,
       FRHIP_BCX_COS = 12, FRHIP_BCX_TAN = 13, FRHIP_BCX_EXP = 14,
       FRHIP_BCX_LOG = 15, FRHIP_BCX_SQRT = 16, FRHIP_BCX_ABS = 17,
       FRHIP_BCX_TANH = 18, FRHIP_BCX_ASIN = 19, FRHIP_BCX_ACOS = 20,
       FRHIP_BCX_ATAN = 21 };

/* A validated program of nrows (<= 8) expressions: expression r is the
 * pairs code[2*offs[r] .. 2*offs[r+1]); pool holds its constants.  Checked
 * on the host (opcodes, indices, stack depth <= 16, one result each) and
 * copied to the device with a device time scalar (initially 0). */
int frhip_bc_program_create(void **prog, int nrows, const int32_t *code,
                            const int32_t *offs, const double *pool,
                            int npool);
int frhip_bc_program_destroy(void *prog);

/* Set the program's t on `stream`, ahead of the work that evaluates it
 * (a captured graph keeps reading the same device scalar). */
int frhip_bc_program_set_time(void *prog, double t, void *stream);

/* Set the program's constant pool[idx] on `stream` (run-time parameters of
 * an expression, e.g. the pressure t*im + ic of the controlled
 * char-riem-inv types, pyfr/solvers/euler/kernels/bcs/
 * char-riem-inv-controlled.mako, which the reference binds as ic, im). */
int frhip_bc_program_set_const(void *prog, int idx, double v, void *stream);

/* out[r][i] = expression r at point i (coordinates ploc[0..2][i], row
 * stride ldp) and the program's t, for n points; out has nrows rows of
 * stride ldo, in the working precision. */
int frhip_bc_eval(void *prog, int dtype, int n, const void *ploc, long ldp,
                  void *out, long ldo, void *stream);

/* ---- BLAS extensions: pyfr/backends/hip/blasext.py:8-136 ---- */

/* axnpby: x0 = a0*x0 + sum_i a_i*x_i over (nrow, nvar*neles_pad) arrays
 * with the a0 == 0 / a0 == 1 special cases of hip/kernels/axnpby.mako:4-44.
 * xs[0] is x0; nv <= 8. */
int frhip_axnpby(int dtype, int nrow, long ncol, long ld, int nv,
                 void *const *xs, const double *a, void *stream);

/* axnpby over a pool of independent solves stored side by side (no
 * counterpart in the reference; it batches the RK4 axpys of
 * pyfr/integrators/base.py:401-426 across solves with different dt): the
 * coefficient of operand i at element e is dt[e]/a[i] if bit i of dtmask is
 * set, else a[i].  dt: per-element array (neles_pad values) in dtype.
 * ncol must be a multiple of 64*nvar; nv <= 4. */
int frhip_axnpby_pool(int dtype, int nrow, long ncol, long ld, int nv,
                      void *const *xs, const double *a, unsigned dtmask,
                      const void *dt, int nvar, void *stream);

/* rkvdh2 stage update: pyfr/integrators/explicit/kernels/rkvdh2.mako:4-31
 * (no error estimate): if last: r1 += dt*b*r2 else r1' = r1 + dt*a*r2,
 * r2' = r1 + dt*b*r2. */
int frhip_rkvdh2(int dtype, int nrow, long ncol, long ld, void *r1, void *r2,
                 double dta, double dtb, int last, void *stream);

/* Mixed-precision rkvdh2 stage (no counterpart kernel in the reference, which
 * runs rkvdh2.mako in one precision): FP64 accumulated solution acc, FP32
 * array kin holding the stage RHS on entry and the next stage input on exit.
 * acc' = acc + dtb*k; kin' = (float)(acc + dta*k), or (float)acc' if last.
 * Same (nrow, ld) for both arrays (element counts, not bytes). */
int frhip_rkvdh2_mixed(int nrow, long ncol, long ld, void *acc, void *kin,
                       double dta, double dtb, int last, void *stream);

/* ---- reduction provider: pyfr/backends/base/blasext.py:21-46
 *      (reduction(rop, exprs, vvars, svars, pvars) + retval), replacing
 *      HIPBlasExtKernels._reduction (pyfr/backends/hip/blasext.py:76-136)
 *      and its template pyfr/backends/hip/kernels/reduction.mako ---- */
enum { FRHIP_RED_SUM = 0, FRHIP_RED_MAX = 1 };
/* Partial blocks per variable: a workspace holds nexprs*nvar*FRHIP_RED_NBLK
 * doubles at most (frhip_reduction_workspace gives the exact count). */
enum { FRHIP_RED_NBLK = 512 };

/* Compile-once expression program: nexprs (<= 8) expressions, each a stack
 * program of (op, arg) pairs with the FRHIP_BCX_* opcodes, over nslots
 * (<= 16) variables (FRHIP_BCX_PUSHV arg = variable index).  Validated here
 * (opcodes, operand counts, depth, index ranges). */
int frhip_expr_program_create(void **prog, int nexprs, int nslots,
                              const int32_t *code, const int32_t *offs,
                              const double *pool, int npool);
int frhip_expr_program_destroy(void *prog);

/* out[x*nvar + v] = rop over rows r < nrow and elements e < neles of
 * expression x evaluated (in double) at point (r, v, e) of nmat (nrow, nvar,
 * neles) arrays mats[m] (leading dimension lds[m]).  The expression
 * variables are the arrays' values there (an array whose bit is set in
 * bcast is (nrow, neles) with one value per (row, element), shared by every
 * variable: an extension, e.g. quadrature weights), then pv[k*nvar + v] for
 * k < npv (the reference's pvars, device memory), then the nsv (<= 8)
 * scalars sv (host memory; the reference's bound svars).  Accumulation is in
 * double in a fixed two-pass order; the sum starts at 0, the max at
 * -DBL_MAX.  work: device workspace of nwork doubles, at least what
 * frhip_reduction_workspace reports; out: device, nexprs*nvar doubles.
 * Allocates nothing; graph-capturable.  The reference reduces the result
 * over variables on the host (retval); the caller does the same. */
int frhip_reduction(void *prog, int dtype, int rop, int nrow, int neles,
                    int nvar, int nmat, void *const *mats, const long *lds,
                    unsigned bcast, const double *pv, int npv,
                    const double *sv, int nsv, double *work, long nwork,
                    double *out, void *stream);
int frhip_reduction_workspace(void *prog, int nrow, int neles, int nvar,
                              long *nwork);

/* frhip_reduction without a workspace argument, for callers that cannot
 * allocate device memory (code under graph capture, tight serving loops,
 * harnesses that forbid allocation).  No counterpart in the reference.
 * The partials live in a library-owned device array of 64*FRHIP_RED_NBLK
 * doubles (256 KiB per device, held until the process ends), so
 * nexprs*nvar must be <= 64 (larger calls fail with a message; use
 * frhip_reduction).  Nothing in it has to start at zero: each call
 * overwrites the partials it reads.  Calls that overlap on one device
 * (different streams or threads) race on that array.  Allocates nothing;
 * results equal frhip_reduction's bit for bit. */
int frhip_reduction_noalloc(void *prog, int dtype, int rop, int nrow,
                            int neles, int nvar, int nmat, void *const *mats,
                            const long *lds, unsigned bcast, const double *pv,
                            int npv, const double *sv, int nsv, double *out,
                            void *stream);

/* axnpby with per-variable scales: pyfr/backends/base/blasext.py:9-19
 * (in_scale, in_scale_idxs, out_scale) as rendered by
 * pyfr/backends/hip/kernels/axnpby.mako:4-44 and
 * pyfr/backends/base/makoutil.py:40-54: operand l of variable k is scaled
 * by insc[k] if bit l of inmask is set, the sum by outsc[k] (if outsc is not
 * NULL); the a0 == 0 / a0 == 1 cases as in frhip_axnpby.  insc, outsc: host
 * arrays of nvar (<= 16) values.  ncol = nvar*neles_pad. */
int frhip_axnpby_scaled(int dtype, int nrow, long ncol, long ld, int nv,
                        int nvar, void *const *xs, const double *a,
                        unsigned inmask, const double *insc,
                        const double *outsc, void *stream);

/* RK4 stage `stage` (1..4) of the precision form (no counterpart kernel in
 * the reference, which runs its RK4 axnpby updates in one precision,
 * pyfr/integrators/explicit/steppers.py:67-108): r1, r0 FP64, k the FP32
 * stage RHS, nxt the FP32 next stage input, with the arithmetic of the
 * frhip_mul2 epilogues 14..17:
 *   1: r1 = k;                      nxt = r0 + dt/2 k
 *   2: r1 = dt/6 r1 + r0 + dt/3 k;  nxt = dt/2 k + r0
 *   3: r1 = r1 + dt/3 k;            nxt = dt k + r0
 *   4: r1 = r1 + dt/6 k;            nxt = r1 (the next step's input).
 * Same (nrow, ld) in elements for all four arrays. */
int frhip_rk4_mixed(int nrow, long ncol, long ld, int stage, void *r1,
                    const void *r0, const void *k, void *nxt, double dt,
                    void *stream);

/* rkvdh2 stage with the embedded error estimate:
 * pyfr/integrators/explicit/kernels/rkvdh2.mako:4-31 with errest (rold,
 * rerr registers; e = b - bhat of stage `stage`,
 * pyfr/integrators/explicit/steppers.py:127-128). */
int frhip_rkvdh2_errest(int dtype, int nrow, long ncol, long ld, void *r1,
                        void *r2, void *rold, void *rerr, double dt,
                        double a, double b, double e, int stage, int last,
                        void *stream);

/* Copy/convert between precisions (same layout); n elements. */
int frhip_convert(int dst_dtype, void *dst, int src_dtype, const void *src,
                  long n, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* FRHIP_H */
