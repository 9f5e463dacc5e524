"""Oracle: ctypes driver of the C/OpenMP RHS restatement (``crhs.c``;
test infrastructure and CPU baseline).

``CRhs(orc)`` takes a NumPy ``OracleHexSystem`` (operators, metric terms and
interface pairs) and evaluates the same RHS in C over OpenMP threads, in the
reference OpenMP backend's blocked layout.  Build with ``make -C oracle``.
"""
import ctypes as ct
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, '_build', 'libcrhs.so')

_P, _I, _D = ct.c_void_p, ct.c_int, ct.c_double


def host_isa():
    """'v4' if the host CPU has AVX-512 (F, BW, DQ, VL), else 'v3'."""
    try:
        with open('/proc/cpuinfo') as f:
            flags = next((ln.split(':', 1)[1].split() for ln in f
                          if ln.startswith('flags')), [])
    except OSError:
        flags = []
    need = {'avx512f', 'avx512bw', 'avx512dq', 'avx512vl'}
    return 'v4' if need <= set(flags) else 'v3'


def cpu_model():
    try:
        with open('/proc/cpuinfo') as f:
            for ln in f:
                if ln.startswith('model name'):
                    return ln.split(':', 1)[1].strip()
    except OSError:
        pass
    return 'unknown'


def lib_path(fast=False):
    """The parity build, or (fast) the -ffast-math timing build for the
    host's widest ISA."""
    if not fast:
        return LIB
    return os.path.join(HERE, '_build', f'libcrhs_fast_{host_isa()}.so')


def _load(fast=False):
    path = lib_path(fast)
    if not os.path.exists(path):
        raise RuntimeError(f'{path} is not built (make -C oracle)')
    lib = ct.CDLL(path)
    lib.crhs_create.restype = _P
    lib.crhs_create.argtypes = [_I, _I, _I, _I] + [_P]*7 + [_I] + [_P]*6
    lib.crhs_destroy.argtypes = [_P]
    lib.crhs_set_u.argtypes = [_P, _P]
    lib.crhs_get_out.argtypes = [_P, _P]
    lib.crhs_rhs.argtypes = [_P]
    return lib


class CRhs:
    def __init__(self, orc, fast=False):
        """fast: the timing build (-ffast-math, host ISA) instead of the
        parity build."""
        self.lib = lib = _load(fast)
        self.orc = orc
        ns = orc.system != 'euler'
        c = orc.c
        f64 = lambda a: np.ascontiguousarray(a, dtype=np.float64)
        i64 = lambda a: np.ascontiguousarray(a, dtype=np.int64)
        self._keep = keep = [
            f64(orc.M0), f64(orc.M132), f64(orc.M3), f64(orc.M460),
            f64(orc.M6), f64(orc.smats_u),
            f64(orc.rcpdjac if not ns else orc.rcpdjac_l),
            i64(orc.lf), i64(orc.le), i64(orc.rf), i64(orc.re), f64(orc.nl),
            f64([c['gamma'], c.get('mu', 0.0), c.get('Pr', 1.0),
                 c.get('ldg-beta', 0.0), c.get('ldg-tau', 0.0)])]
        self.h = lib.crhs_create(
            int(ns), orc.neles, orc.nupts, orc.nfpts,
            *[a.ctypes.data for a in keep[:7]], len(orc.lf),
            *[a.ctypes.data for a in keep[7:]])
        self.shape = (orc.nupts, 5, orc.neles)

    def set_u(self, u):
        u = np.ascontiguousarray(u, dtype=np.float64)
        assert u.shape == self.shape
        self.lib.crhs_set_u(self.h, u.ctypes.data)

    def rhs_inplace(self):
        self.lib.crhs_rhs(self.h)

    def rhs(self, u):
        self.set_u(u)
        self.rhs_inplace()
        out = np.empty(self.shape)
        self.lib.crhs_get_out(self.h, out.ctypes.data)
        return out

    def __del__(self):
        try:
            self.lib.crhs_destroy(self.h)
        except Exception:
            pass
