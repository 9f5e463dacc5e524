"""Generate the structured operator kernels (GiMMiK-style, ahead of time).

For every hex order and every operator the FR right-hand side applies
(``disu`` M0, ``tdivtpcorf`` M1 - M3*M2, ``tdivtconf`` M3, ``tgradpcoru``
M4 - M6*M0, ``tgradcoru`` M6; pyfr/solvers/baseadvec/elements.py:73-100 and
pyfr/solvers/baseadvecdiff/elements.py:47-55) this writes a fully unrolled
HIP kernel with the operator's sparsity pattern and nonzero values baked in
as literals (exact hex floats; FP32 kernels round the same literals), so the
operator lives in instruction immediates / scalar registers and B streams
through vector registers.  (Reading the values from memory instead makes the
compiler hoist ~nnz scalar loads and spill them to VGPR lanes.)  The run-time
plan uses a kernel only if the requested operator matches its pattern and
values bit for bit.

One thread owns one column of B/C (one (variable, element) pair of the
element-major AoSoA layout), so every B and C access is a fully coalesced
512 B (FP64) wavefront transaction and B is read from HBM exactly once:
  * if k <= m the kernel keeps the k inputs of its column in registers and
    streams the outputs (row-oriented);
  * otherwise it keeps the m accumulators and streams the inputs
    (column-oriented).
Both accumulate each output over the columns of A in ascending order.

Usage: python gen_ops.py OUTDIR
"""
import hashlib
import os
import sys

import numpy as np

sys.path.insert(0, os.path.join(os.path.dirname(__file__), '..', '..'))

from pyfr_amd.shapes import HexShape  # noqa: E402

OPS = ['M0', 'M1 - M3*M2', 'M3', 'M4 - M6*M0', 'M6']
ORDERS = [1, 2, 3, 4]


def pattern_hash(m, k, lin):
    """Hash of the nonzero pattern shared with the runtime matcher."""
    h = hashlib.sha1()
    h.update(np.array([m, k], dtype=np.int64).tobytes())
    h.update(np.asarray(lin, dtype=np.int64).tobytes())
    return int.from_bytes(h.digest()[:8], 'little') & 0x7fffffffffffffff


def lit(x):
    return f'T({float(x).hex()})'


def pipelined(batches, loads, fmas, L):
    """Software-pipelined k loop: the loads of batch q + 1 are issued before
    the FMAs of batch q.  Scheduling barriers keep the compiler from hoisting
    every load of the loop to the top, which spills the accumulators."""
    if not batches:
        return
    loads(batches[0])
    for q, ks in enumerate(batches):
        if q + 1 < len(batches):
            loads(batches[q + 1])
        L.append('    __builtin_amdgcn_sched_barrier(0);')
        fmas(ks)
        L.append('    __builtin_amdgcn_sched_barrier(0);')


def emit_kernel(name, A):
    m, k = A.shape
    rows, cols = np.nonzero(A)          # row-major order of nonzeros
    nnz = len(rows)
    vid = {(r, c): lit(A[r, c]) for r, c in zip(rows, cols)}
    L = []
    L.append('template <typename T, int BM>')
    L.append(f'__global__ __launch_bounds__(256) void {name}('
             'int ncol, const T *__restrict__ b, long ldb, '
             'T *__restrict__ c, long ldc, const T *__restrict__ v, '
             'T alpha, T beta)')
    L.append('{')
    L.append('    const long j = (long)blockIdx.x*blockDim.x + threadIdx.x;')
    L.append('    if (j >= ncol) return;')
    L.append('    b += j; c += j;')
    usedk = sorted(set(cols.tolist()))
    # Loads are issued in batches so that each wavefront keeps many memory
    # requests in flight: all B inputs up front (row-oriented) or GB at a time
    # (column-oriented), and C (read-modify-write modes) GC rows at a time.
    GB, GC = 8, 8

    def store_rows(rs, accs):
        # rs: rows, accs: expressions holding their accumulated values
        for q0 in range(0, len(rs), GC):
            blk = list(zip(rs[q0:q0 + GC], accs[q0:q0 + GC]))
            L.append('    if constexpr (BM == 1 || BM == 2) {')
            for r, _ in blk:
                L.append(f'      const T c{r} = c[{r}*ldc];')
            for r, a in blk:
                if_ = f'BM == 1 ? {a} + c{r} : alpha*{a} + beta*c{r}'
                L.append(f'      c[{r}*ldc] = {if_};')
            L.append('    } else {')
            for r, a in blk:
                L.append(f'      opstore<T, BM>(c + {r}*ldc, {a}, alpha, beta);')
            L.append('    }')

    if k <= m:
        for kk in usedk:
            L.append(f'    const T b{kk} = ldg(b + {kk}*ldb);')
        for q0 in range(0, m, GC):
            rs = list(range(q0, min(m, q0 + GC)))
            for r in rs:
                cs = cols[rows == r]
                if len(cs) == 0:
                    L.append(f'    const T a{r} = T(0);')
                    continue
                terms = f'{vid[(r, cs[0])]}*b{cs[0]}'
                for cc in cs[1:]:
                    terms = f'fma({vid[(r, cc)]}, b{cc}, {terms})'
                L.append(f'    const T a{r} = {terms};')
            store_rows(rs, [f'a{r}' for r in rs])
    else:
        seen = set()
        for r in range(m):
            L.append(f'    T a{r} = T(0);')
        batches = [usedk[q0:q0 + GB] for q0 in range(0, len(usedk), GB)]

        def loads(ks):
            for kk in ks:
                L.append(f'    const T b{kk} = ldg(b + {kk}*ldb);')

        def fmas(ks):
            for kk in ks:
                for r in rows[cols == kk]:
                    if r in seen:
                        L.append(f'    a{r} = fma({vid[(r, kk)]}, b{kk}, a{r});')
                    else:
                        L.append(f'    a{r} = {vid[(r, kk)]}*b{kk};')
                        seen.add(r)
        pipelined(batches, loads, fmas, L)
        store_rows(list(range(m)), [f'a{r}' for r in range(m)])
    L.append('}')
    return '\n'.join(L), rows*k + cols, nnz, A[rows, cols]


# Fused two-input kernels.  C = A0 @ B0 + A1 @ B1, then an epilogue:
#   EP 0: c = acc
#   EP 1: c = -rcp*acc                          (negdivconf)
#   EP 2: t = -rcp*acc; RK45 stage on r1 (xa) and r2 (c) in T
#   EP 3: t = -rcp*acc in float; RK45 stage with FP64 accumulated solution
#         (xa) and FP32 next-stage input (c)
#   EP 4, 5: the last stage of EP 2, 3 (r1 += dtb*t; EP 5 also stores the
#         new solution rounded to float into c)
#   EP 10-13: RK4 stage EP - 9 with four registers: xa points to the device
#         pointers {r1, r0, dt} (dt a device double, so a captured graph
#         stays valid when the step changes), c receives the next stage's
#         input (the axnpby updates of pyfr/integrators/explicit/
#         steppers.py:67-108, operands in the same order)
#   EP 14-17: EP 10-13 for the precision form: r1 and r0 FP64, the stage
#         RHS t and the next stage's input c FP32; the updates are formed in
#         FP64 in the same order, and EP 17 stores the new solution rounded
#         to float into c (the next step's first stage input)
#   EP 20-23: EP 10-13 with a time step per element (pools of independent
#         solves): xa points to {r1, r0, dt} with dt one value per element
# The last stage is a template value, not a run-time flag: a uniform branch
# per output row makes the compiler keep every row's RK operands live and
# spill.  The trailing int argument is unused.
# Used for tgradpcoru + tgradcoru (EP 0: [M4 - M6*M0 | M6] on [u; comm]) and
# tdivtpcorf + tdivtconf + negdivconf (+ rkvdh2) (EP 1-3: [M1 - M3*M2 | M3]
# on [F; fn]).  The fused ops follow the element-local grouping the
# reference's OpenMP backend applies (pyfr/solvers/baseadvec/system.py:126-132,
# pyfr/solvers/baseadvecdiff/system.py:174-203).
FUSED = [('grad', ['M4 - M6*M0', 'M6']), ('div', ['M1 - M3*M2', 'M3'])]
FUSED_EPS = {'double': [0, 1, 2, 4, 10, 11, 12, 13, 20, 21, 22, 23],
             'float': [0, 1, 2, 3, 4, 5, 10, 11, 12, 13, 14, 15, 16, 17]}
# kernel table slot of each epilogue (FusedKernelEntry::fd / ff)
EP_SLOTS = [0, 1, 2, 3, 4, 5, 10, 11, 12, 13, 14, 15, 16, 17, 20, 21, 22,
            23]


def emit_fused(name, A0, A1, gb=8, attrs=''):
    """gb: inputs loaded per batch of the k loop; attrs: function
    attributes (occupancy experiments)."""
    m, k0 = A0.shape
    k1 = A1.shape[1]
    A = np.hstack([A0, A1])
    rows, cols = np.nonzero(A)
    vid = {(r, c): lit(A[r, c]) for r, c in zip(rows, cols)}
    L = []
    L.append('template <typename T, int EP>')
    L.append(f'__global__ __launch_bounds__(256) {attrs} void {name}('
             'int ncol, const T *__restrict__ b0, long ld0, '
             'const T *__restrict__ b1, long ld1, T *__restrict__ c, '
             'long ldc, const T *__restrict__ rcp, long ldr, '
             'void *__restrict__ xa, double dta, double dtb, int last)')
    L.append('{')
    L.append('    const long j = (long)blockIdx.x*blockDim.x + threadIdx.x;')
    L.append('    if (j >= ncol) return;')
    L.append('    b0 += j; b1 += j; c += j;')
    L.append('    const long e = (j/(64*5))*64 + (j & 63);')
    L.append('    if constexpr (EP >= 10 && EP <= 17)')
    L.append('        dta = *static_cast<const double *>(')
    L.append('            static_cast<void *const *>(xa)[2]);')

    def bexpr(kk):
        return (f'ldg(b0 + {kk}*ld0)' if kk < k0 else
                f'ldg(b1 + {kk - k0}*ld1)')

    def epilogue(r, a):
        L.append('    {')
        L.append(f'      const T acc = {a};')
        L.append('      if constexpr (EP == 0) {')
        L.append(f'        __builtin_nontemporal_store(acc, c + {r}*ldc);')
        L.append('      } else {')
        L.append(f'        const T t = -rcp[{r}*ldr + e]*acc;')
        L.append('        if constexpr (EP == 1) {')
        L.append(f'          c[{r}*ldc] = t;')
        L.append('        } else if constexpr (EP == 2) {')
        L.append(f'          T *r1 = static_cast<T *>(xa) + j + {r}*ldc;')
        L.append('          const T t1 = *r1;')
        L.append('          *r1 = t1 + T(dta)*t;')
        L.append(f'          c[{r}*ldc] = t1 + T(dtb)*t;')
        L.append('        } else if constexpr (EP == 4) {')
        L.append(f'          T *r1 = static_cast<T *>(xa) + j + {r}*ldc;')
        L.append('          *r1 = *r1 + T(dtb)*t;')
        L.append('        } else if constexpr (EP >= 20) {')
        L.append('          T *const *pp = static_cast<T *const *>(xa);')
        L.append(f'          T *r1 = pp[0] + j + {r}*ldc;')
        L.append(f'          const T r0 = pp[1][j + {r}*ldc];')
        L.append('          const T dte = pp[2][e];')
        L.append('          if constexpr (EP == 20) {')
        L.append('            *r1 = t;')
        L.append('            T a = T(1)*r0;')
        L.append('            a = a + (dte/T(2))*t;')
        L.append(f'            c[{r}*ldc] = a;')
        L.append('          } else if constexpr (EP == 21) {')
        L.append('            T a = (dte/T(6))*(*r1);')
        L.append('            a = a + T(1)*r0;')
        L.append('            a = a + (dte/T(3))*t;')
        L.append('            *r1 = a;')
        L.append('            T b = (dte/T(2))*t;')
        L.append('            b = b + T(1)*r0;')
        L.append(f'            c[{r}*ldc] = b;')
        L.append('          } else if constexpr (EP == 22) {')
        L.append('            *r1 += (dte/T(3))*t;')
        L.append('            T b = dte*t;')
        L.append('            b = b + T(1)*r0;')
        L.append(f'            c[{r}*ldc] = b;')
        L.append('          } else {')
        L.append('            *r1 += (dte/T(6))*t;')
        L.append('          }')
        L.append('        } else if constexpr (EP >= 14) {')
        L.append('          double *const *pp = '
                 'static_cast<double *const *>(xa);')
        L.append(f'          double *r1 = pp[0] + j + {r}*ldc;')
        L.append(f'          const double r0 = pp[1][j + {r}*ldc];')
        L.append('          const double td = (double)t;')
        L.append('          if constexpr (EP == 14) {')
        L.append('            *r1 = td;')
        L.append('            double a = 1.0*r0;')
        L.append('            a = a + (dta/2.0)*td;')
        L.append(f'            c[{r}*ldc] = (T)a;')
        L.append('          } else if constexpr (EP == 15) {')
        L.append('            double a = (dta/6.0)*(*r1);')
        L.append('            a = a + 1.0*r0;')
        L.append('            a = a + (dta/3.0)*td;')
        L.append('            *r1 = a;')
        L.append('            double b = (dta/2.0)*td;')
        L.append('            b = b + 1.0*r0;')
        L.append(f'            c[{r}*ldc] = (T)b;')
        L.append('          } else if constexpr (EP == 16) {')
        L.append('            *r1 += (dta/3.0)*td;')
        L.append('            double b = dta*td;')
        L.append('            b = b + 1.0*r0;')
        L.append(f'            c[{r}*ldc] = (T)b;')
        L.append('          } else {')
        L.append('            const double na = *r1 + (dta/6.0)*td;')
        L.append('            *r1 = na;')
        L.append(f'            c[{r}*ldc] = (T)na;')
        L.append('          }')
        L.append('        } else if constexpr (EP >= 10) {')
        L.append('          T *const *pp = static_cast<T *const *>(xa);')
        L.append(f'          T *r1 = pp[0] + j + {r}*ldc;')
        L.append(f'          const T r0 = pp[1][j + {r}*ldc];')
        L.append('          if constexpr (EP == 10) {')
        L.append('            *r1 = t;')
        L.append('            T a = T(1)*r0;')
        L.append('            a = a + T(dta/2.0)*t;')
        L.append(f'            c[{r}*ldc] = a;')
        L.append('          } else if constexpr (EP == 11) {')
        L.append('            T a = T(dta/6.0)*(*r1);')
        L.append('            a = a + T(1)*r0;')
        L.append('            a = a + T(dta/3.0)*t;')
        L.append('            *r1 = a;')
        L.append('            T b = T(dta/2.0)*t;')
        L.append('            b = b + T(1)*r0;')
        L.append(f'            c[{r}*ldc] = b;')
        L.append('          } else if constexpr (EP == 12) {')
        L.append('            *r1 += T(dta/3.0)*t;')
        L.append('            T b = T(dta)*t;')
        L.append('            b = b + T(1)*r0;')
        L.append(f'            c[{r}*ldc] = b;')
        L.append('          } else {')
        L.append('            *r1 += T(dta/6.0)*t;')
        L.append('          }')
        L.append('        } else {')
        L.append(f'          double *r1 = static_cast<double *>(xa) + j + '
                 f'{r}*ldc;')
        L.append('          const double t1 = *r1, td = (double)t;')
        L.append('          const double na = t1 + dtb*td;')
        L.append('          *r1 = na;')
        L.append(f'          c[{r}*ldc] = EP == 5 ? (T)na : (T)(t1 + dta*td);')
        L.append('        }')
        L.append('      }')
        L.append('    }')

    usedk = sorted(set(cols.tolist()))
    GB = gb
    if len(usedk) <= m:
        for kk in usedk:
            L.append(f'    const T v{kk} = {bexpr(kk)};')
        for r in range(m):
            cs = cols[rows == r]
            if len(cs) == 0:
                epilogue(r, 'T(0)')
                continue
            terms = f'{vid[(r, cs[0])]}*v{cs[0]}'
            for cc in cs[1:]:
                terms = f'fma({vid[(r, cc)]}, v{cc}, {terms})'
            epilogue(r, terms)
            if r % 8 == 7:
                L.append('    __builtin_amdgcn_sched_barrier(0);')
    else:
        seen = set()
        for r in range(m):
            L.append(f'    T a{r} = T(0);')
        batches = [usedk[q0:q0 + GB] for q0 in range(0, len(usedk), GB)]

        def loads(ks):
            for kk in ks:
                L.append(f'    const T v{kk} = {bexpr(kk)};')

        def fmas(ks):
            for kk in ks:
                for r in rows[cols == kk]:
                    if r in seen:
                        L.append(f'    a{r} = fma({vid[(r, kk)]}, v{kk}, '
                                 f'a{r});')
                    else:
                        L.append(f'    a{r} = {vid[(r, kk)]}*v{kk};')
                        seen.add(r)
        pipelined(batches, loads, fmas, L)
        # the epilogue's own loads (1/|J|, RK registers) go 8 rows at a time
        for r in range(m):
            epilogue(r, f'a{r}')
            if r % 8 == 7:
                L.append('    __builtin_amdgcn_sched_barrier(0);')
    L.append('}')
    return '\n'.join(L), rows*(k0 + k1) + cols, len(rows), A[rows, cols]


def main(outdir):
    """Write one translation unit per order (gen_ops_p<k>.hip, compiled in
    parallel) and the kernel tables (gen_ops_table.hip)."""
    head = ['// Generated by gen_ops.py -- structured FR operator kernels.',
            '#include "frhip_opk.h"', '', 'namespace frhip {', '']
    table, ftable, decls = [], [], []
    files = []
    for p in ORDERS:
        parts = list(head)
        sh = HexShape(p)
        for oi, op in enumerate(OPS):
            A = sh.opmat(op)
            name = f'opk_hex{p}_{oi}'
            src, lin, nnz, vals = emit_kernel(name, A)
            parts.append(f'// hex p={p} {op}: {A.shape[0]}x{A.shape[1]}, '
                         f'nnz={nnz}')
            parts.append(src)
            parts.append(f'extern const int32_t lin_{name}[{nnz}] = {{' +
                         ','.join(str(int(x)) for x in lin) + '};')
            parts.append(f'extern const double val_{name}[{nnz}] = {{' +
                         ','.join(float(x).hex() for x in vals) + '};')
            for bm in range(4):
                for t in ('double', 'float'):
                    parts.append(f'template __global__ void {name}<{t}, {bm}>'
                                 f'(int, const {t} *, long, {t} *, long, '
                                 f'const {t} *, {t}, {t});')
            parts.append('')
            h = pattern_hash(A.shape[0], A.shape[1], lin)
            table.append((name, A.shape[0], A.shape[1], nnz, h,
                          f'hex p={p} {op}'))
            decls.append(f'template <typename T, int BM> __global__ void '
                         f'{name}(int, const T *, long, T *, long, '
                         f'const T *, T, T);')
            decls.append(f'extern const int32_t lin_{name}[];')
            decls.append(f'extern const double val_{name}[];')
        for fname, (op0, op1) in FUSED:
            A0, A1 = sh.opmat(op0), sh.opmat(op1)
            name = f'fk_hex{p}_{fname}'
            # the p = 3 FP64 divergence kernel needs 170 VGPRs, two past the
            # 168 that give three waves per SIMD; capped there it spills two
            # and streams 8 % faster (5.08 against 4.70 TB/s with the RK45
            # stage, 64^3); no effect on the FP32 or smaller kernels
            at = ('__attribute__((amdgpu_waves_per_eu(3)))'
                  if fname == 'div' and A0.shape[0] <= 64 else '')
            src, lin, nnz, vals = emit_fused(name, A0, A1, attrs=at)
            parts.append(f'// hex p={p} fused {fname}: [{op0} | {op1}] '
                         f'{A0.shape[0]}x({A0.shape[1]}+{A1.shape[1]}), '
                         f'nnz={nnz}')
            parts.append(src)
            parts.append(f'extern const int32_t lin_{name}[{nnz}] = {{' +
                         ','.join(str(int(x)) for x in lin) + '};')
            parts.append(f'extern const double val_{name}[{nnz}] = {{' +
                         ','.join(float(x).hex() for x in vals) + '};')
            for t in ('double', 'float'):
                # the gradient kernel only stores
                for ep in FUSED_EPS[t] if fname == 'div' else [0]:
                    parts.append(f'template __global__ void {name}<{t}, {ep}>'
                                 f'(int, const {t} *, long, const {t} *, '
                                 f'long, {t} *, long, const {t} *, long, '
                                 f'void *, double, double, int);')
            parts.append('')
            ftable.append((name, A0.shape[0], A0.shape[1], A1.shape[1], nnz,
                           f'hex p={p} fused {fname}', fname))
            decls.append(f'template <typename T, int EP> __global__ void '
                         f'{name}(int, const T *, long, const T *, long, T *, '
                         f'long, const T *, long, void *, double, double, '
                         f'int);')
            decls.append(f'extern const int32_t lin_{name}[];')
            decls.append(f'extern const double val_{name}[];')
        parts.append('}  // namespace frhip')
        fn = os.path.join(outdir, f'gen_ops_p{p}.hip')
        with open(fn, 'w') as f:
            f.write('\n'.join(parts) + '\n')
        files.append(fn)

    t = list(head) + decls + ['']
    t.append('const FusedKernelEntry g_fk_table[] = {')
    for name, m, k0, k1, nnz, desc, fname in ftable:
        def fns(t):
            eps = FUSED_EPS[t] if fname == 'div' else [0]
            return ', '.join(f'&{name}<{t}, {ep}>' if ep in eps else 'nullptr'
                             for ep in EP_SLOTS)
        fnd, fnf = fns('double'), fns('float')
        t.append(f'    {{{m}, {k0}, {k1}, {nnz}, lin_{name}, val_{name}, '
                 f'"{desc}", {{{fnd}}}, {{{fnf}}}}},')
    t.append('};')
    t.append(f'const int g_fk_count = {len(ftable)};')
    t.append('const OpKernelEntry g_opk_table[] = {')
    for name, m, k, nnz, h, desc in table:
        fnd = ', '.join(f'&{name}<double, {bm}>' for bm in range(4))
        fnf = ', '.join(f'&{name}<float, {bm}>' for bm in range(4))
        t.append(f'    {{{m}, {k}, {nnz}, {h}ULL, lin_{name}, val_{name}, '
                 f'"{desc}", {{{fnd}}}, {{{fnf}}}}},')
    t.append('};')
    t.append(f'const int g_opk_count = {len(table)};')
    t.append('}  // namespace frhip')
    fn = os.path.join(outdir, 'gen_ops_table.hip')
    with open(fn, 'w') as f:
        f.write('\n'.join(t) + '\n')
    files.append(fn)
    return files


if __name__ == '__main__':
    main(sys.argv[1])
