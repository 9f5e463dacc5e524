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

Usage: python gen_ops.py OUTFILE
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
    GB, GC = 16, 8

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
        for q0 in range(0, len(usedk), GB):
            ks = usedk[q0:q0 + GB]
            for kk in ks:
                L.append(f'    const T b{kk} = ldg(b + {kk}*ldb);')
            for kk in ks:
                for r in rows[cols == kk]:
                    if r in seen:
                        L.append(f'    a{r} = fma({vid[(r, kk)]}, b{kk}, a{r});')
                    else:
                        L.append(f'    a{r} = {vid[(r, kk)]}*b{kk};')
                        seen.add(r)
        store_rows(list(range(m)), [f'a{r}' for r in range(m)])
    L.append('}')
    return '\n'.join(L), rows*k + cols, nnz, A[rows, cols]


def main(outfile):
    parts = ['// Generated by gen_ops.py -- structured FR operator kernels.',
             '#include "frhip_opk.h"', '', 'namespace frhip {', '']
    table = []
    for p in ORDERS:
        sh = HexShape(p)
        for oi, op in enumerate(OPS):
            A = sh.opmat(op)
            name = f'opk_hex{p}_{oi}'
            src, lin, nnz, vals = emit_kernel(name, A)
            parts.append(f'// hex p={p} {op}: {A.shape[0]}x{A.shape[1]}, '
                         f'nnz={nnz}')
            parts.append(src)
            linname = f'lin_{name}'
            parts.append(f'static const int32_t {linname}[{nnz}] = {{' +
                         ','.join(str(int(x)) for x in lin) + '};')
            parts.append(f'static const double val_{name}[{nnz}] = {{' +
                         ','.join(float(x).hex() for x in vals) + '};')
            parts.append('')
            h = pattern_hash(A.shape[0], A.shape[1], lin)
            table.append((name, A.shape[0], A.shape[1], nnz, h, linname,
                          f'hex p={p} {op}'))
    parts.append('const OpKernelEntry g_opk_table[] = {')
    for name, m, k, nnz, h, linname, desc in table:
        fnd = ', '.join(f'&{name}<double, {bm}>' for bm in range(4))
        fnf = ', '.join(f'&{name}<float, {bm}>' for bm in range(4))
        parts.append(f'    {{{m}, {k}, {nnz}, {h}ULL, {linname}, '
                     f'val_{name}, "{desc}", {{{fnd}}}, {{{fnf}}}}},')
    parts.append('};')
    parts.append('const int g_opk_count = '
                 f'{len(table)};')
    parts.append('}  // namespace frhip')
    with open(outfile, 'w') as f:
        f.write('\n'.join(parts) + '\n')


if __name__ == '__main__':
    main(sys.argv[1])
