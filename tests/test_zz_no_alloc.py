"""The workspace-free reduction entry point (frhip_reduction_noalloc) through
the C ABI against frhip_reduction with a caller workspace: the same kernels
and the same fixed partial order, so the results agree bit for bit.  Two
calls in a row, the second on a smaller input (the scratch keeps the first
call's partials, which the second must not read), sums and maxima, FP64 and
FP32; and the size limit."""
import ctypes as ct

import numpy as np
import pytest

from pyfr_amd.bcexpr import compile_exprs

pytestmark = pytest.mark.gpu


def _program(lib, exprs, names):
    code, offs, pool = compile_exprs(
        exprs, {}, varmap={n: i for i, n in enumerate(names)})
    prog = ct.c_void_p()
    lib.expr_program_create(ct.byref(prog), len(exprs), len(names),
                            code.ctypes.data, offs.ctypes.data,
                            pool.ctypes.data, len(pool))
    return prog


def _both(be, prog, nexprs, dtc, rop, x, w, nvar):
    """(frhip_reduction, frhip_reduction_noalloc) results for x (nrow,
    nvar, neles) with w (nrow, neles) broadcast."""
    lib = be.lib
    nrow, _, neles = x.ioshape
    mats = (ct.c_void_p*2)(x.data, w.data)
    lds = (ct.c_long*2)(x.leaddim, w.leaddim)
    nw = ct.c_long()
    lib.reduction_workspace(prog, nrow, neles, nvar, ct.byref(nw))
    work = ct.c_void_p()
    lib.malloc(ct.byref(work), 8*nw.value)
    outs = []
    for noalloc in (False, True):
        out = ct.c_void_p()
        lib.malloc(ct.byref(out), 8*nexprs*nvar)
        lib.memset(out, 0xff, 8*nexprs*nvar, be.stream)
        if noalloc:
            lib.reduction_noalloc(prog, dtc, rop, nrow, neles, nvar, 2, mats,
                                  lds, 0b10, None, 0, None, 0, out,
                                  be.stream)
        else:
            lib.reduction(prog, dtc, rop, nrow, neles, nvar, 2, mats, lds,
                          0b10, None, 0, None, 0, work, nw.value, out,
                          be.stream)
        host = np.empty(nexprs*nvar)
        lib.memcpy_dtoh(host.ctypes.data, out, host.nbytes, be.stream)
        lib.free(out)
        outs.append(host)
    lib.free(work)
    return outs


@pytest.mark.parametrize('precision', ['double', 'single'])
def test_noalloc_matches_workspace(backend_factory, precision):
    be = backend_factory(precision)
    lib = be.lib
    rng = np.random.default_rng(9)
    exprs = ['x', 'w*x*x', 'fabs(x) + w']
    prog = _program(lib, exprs, ['x', 'w'])
    dtc = 0 if precision == 'double' else 1
    try:
        # first call large (fills many partials), second smaller
        for nrow, neles in ((64, 20000), (27, 300)):
            x = be.matrix((nrow, 5, neles), rng.uniform(-1, 1, (nrow, 5,
                                                              neles)),
                          tags={'align'})
            w = be.const_matrix(rng.uniform(0.1, 1, (nrow, neles)),
                                tags={'align'})
            xn = x.get().astype(np.float64)
            for rop in (0, 1):
                ref, got = _both(be, prog, len(exprs), dtc, rop, x, w, 5)
                assert np.array_equal(ref, got), (nrow, neles, rop)
                if rop == 0:
                    want = xn.sum(axis=(0, 2))
                    mag = np.abs(xn).sum(axis=(0, 2))
                    assert np.all(np.abs(got[:5] - want) <= 1e-13*mag)
                else:
                    wn = w.get().astype(np.float64)
                    want = (np.abs(xn) + wn[:, None, :]).max(axis=(0, 2))
                    assert np.array_equal(got[10:15], want)
    finally:
        lib.expr_program_destroy(prog)


def test_noalloc_size_limit(backend_factory):
    """nexprs*nvar above 64 is refused before any launch."""
    from pyfr_amd._lib import FrHipError
    be = backend_factory('double')
    lib = be.lib
    prog = _program(lib, ['x'], ['x'])
    try:
        x = be.matrix((4, 65, 64), tags={'align'})
        out = ct.c_void_p()
        lib.malloc(ct.byref(out), 8*65)
        mats = (ct.c_void_p*1)(x.data)
        lds = (ct.c_long*1)(x.leaddim)
        with pytest.raises(FrHipError, match='exceeds 64'):
            lib.reduction_noalloc(prog, 0, 0, 4, 64, 65, 1, mats, lds, 0,
                                  None, 0, None, 0, out, be.stream)
        lib.free(out)
    finally:
        lib.expr_program_destroy(prog)
