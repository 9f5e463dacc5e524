"""MI355X backend with the reference's backend / kernel-provider interface.

Mirrors ``pyfr/backends/base/backend.py:32-225`` and the types of
``pyfr/backends/base/types.py`` (extents, matrices, slices, views, graphs,
kernels) so that solver-side code written against PyFR's backend API drives
it unchanged, while all device work goes through ``libfrhip.so``.

Kernel names and arguments are the reference's:

* ``mul(a, b, out, alpha=1.0, beta=0.0)``  (``hip/gimmik.py:24``)
* pointwise ``tflux``, ``intcflux``, ``intconu``, ``negdivconf``, ``rkvdh2``
  as ``kernel(name, tplargs=..., dims=[...], **args)``
  (``backends/base/provider.py:188-200``)
* ``axnpby(*arr, in_scale=(), in_scale_idxs=(), out_scale=())`` + ``bind``,
  ``copy(dst, src)``, ``zero(m)`` (``hip/blasext.py:8-74``)

Device layout is the reference HIP layout (AoSoA with 64-wide SoA blocks,
``types.py:40-100``); ``get()``/``set()`` round-trip ``ioshape`` arrays.
"""
from collections import defaultdict
from contextlib import contextmanager
import ctypes as ct
import itertools as it
import math
from weakref import WeakValueDictionary

import numpy as np

from pyfr_amd._lib import get_lib

F64, F32 = 0, 1


class NotSuitableError(Exception):
    """Raised by a provider that declines a kernel (``provider.py:209``)."""


def _rstride(m):
    """Row stride of a per-point (npts, neles) array; 0 for a one-row
    array of per-element values (affine elements)."""
    return 0 if m.nrow == 1 else m.leaddim


def _dtcode(dtype):
    dt = np.dtype(dtype)
    if dt == np.float64:
        return F64
    if dt == np.float32:
        return F32
    raise ValueError(f'Unsupported dtype {dt}')


# ---------------------------------------------------------------------------
# Storage
# ---------------------------------------------------------------------------

class Extent:
    """A named group of allocations committed as one device buffer
    (``types.py:19-37``)."""

    def __init__(self, name=None):
        self.name = name
        self.nbytes = 0
        self.basedata = None
        self._pending = []
        self._storage_root = self

    def reserve(self, obj, nbytes):
        self._pending.append((obj, self.nbytes))
        self.nbytes += nbytes

    def commit(self, alloc_fn):
        self.basedata = alloc_fn(self.nbytes)
        for obj, off in self._pending:
            obj._storage_root = self
            obj.onalloc(self.basedata, off)
        self._pending.clear()


class MatrixBase:
    """Matrix of ``ioshape`` in the AoSoA layout (``types.py:40-150``)."""
    _base_tags = set()

    def __init__(self, backend, dtype, ioshape, initval, extent, tags):
        self.backend = backend
        self.tags = self._base_tags | set(tags)
        self.dtype = np.dtype(dtype).type
        self.itemsize = np.dtype(dtype).itemsize

        shape, ndim = list(ioshape), len(ioshape)
        soasz, csubsz = backend.soasz, backend.csubsz
        if ndim == 2:
            nrow, ncol = shape
            ldmod = csubsz if 'align' in self.tags else 1
            leaddim = ncol - (ncol % -ldmod)
            datashape = [1, nrow, leaddim]
        else:
            nvar, narr = shape[-2], shape[-1]
            nparr = narr - narr % -csubsz
            nrow = shape[0] if ndim == 3 else shape[0]*shape[1]
            ncol = nvar*nparr
            # rows padded by row_pad elements (the leading dimension stays
            # the only stride every kernel and view uses)
            leaddim = ncol + getattr(backend, 'row_pad', 0)
            datashape = [1, *shape[:-2], nparr // soasz, nvar, soasz]

        self.nrow, self.ncol, self.leaddim = nrow, ncol, leaddim
        self.datashape = datashape
        self.ioshape = tuple(ioshape)
        self.nbytes = nrow*leaddim*self.itemsize
        self.traits = (1, nrow, ncol, leaddim, self.dtype)

        if initval is not None:
            initval = np.asarray(initval)
            if initval.shape != self.ioshape:
                raise ValueError('Invalid initial value')
            self._initval = np.asanyarray(initval, dtype=self.dtype)
        else:
            self._initval = None

        backend.malloc(self, extent)

    # -- allocation hooks
    def onalloc(self, basedata, offset):
        self.basedata = int(basedata)
        self.offset = int(offset)
        self.data = self.basedata + self.offset
        if self._initval is not None:
            self._set(self._initval)
        else:
            self.backend.lib.memset(self.data, 0, self.nbytes,
                                    self.backend.stream)
        del self._initval

    @property
    def storage_root(self):
        return self._storage_root

    def same_storage(self, other):
        return self.storage_root is other.storage_root

    # -- AoSoA pack / unpack (host side)
    def _pack(self, ary):
        ary = np.asarray(ary, dtype=self.dtype)
        if ary.ndim == 2:
            out = np.zeros((self.nrow, self.leaddim), dtype=self.dtype)
            out[:, :ary.shape[1]] = ary
            return out
        k = self.backend.soasz
        n = ary.shape[-1]
        ary = np.pad(ary, [(0, 0)]*(ary.ndim - 1) + [(0, -n % k)])
        ary = ary.reshape(ary.shape[:-1] + (-1, k)).swapaxes(-2, -3)
        out = np.zeros((self.nrow, self.leaddim), dtype=self.dtype)
        out[:, :self.ncol] = ary.reshape(self.nrow, self.ncol)
        return out

    def _unpack(self, buf):
        if len(self.ioshape) == 2:
            return buf.reshape(self.nrow, self.leaddim)[:, :self.ioshape[1]]
        buf = buf.reshape(self.nrow, self.leaddim)[:, :self.ncol]
        ary = buf.reshape(self.datashape[1:]).swapaxes(-2, -3)
        ary = ary.reshape(self.ioshape[:-1] + (-1,))
        return ary[..., :self.ioshape[-1]]

    def _get(self):
        buf = np.empty(self.nrow*self.leaddim, dtype=self.dtype)
        self.backend.lib.memcpy_dtoh(buf.ctypes.data, self.data, self.nbytes,
                                     self.backend.stream)
        return self._unpack(buf)

    def _set(self, ary):
        buf = self._pack(ary)
        self.backend.lib.memcpy_htod(self.data, buf.ctypes.data, self.nbytes,
                                     self.backend.stream)

    def get(self):
        if hasattr(self, '_initval'):
            return (self._initval if self._initval is not None
                    else np.zeros(self.ioshape, dtype=self.dtype))
        self.backend.wait()
        return np.array(self._get())

    def slice(self, ra=None, rb=None, ca=None, cb=None):
        ra, rb = ra or 0, rb or self.nrow
        ca, cb = ca or 0, cb or self.ncol
        return self.backend.matrix_slice(self, ra, rb, ca, cb)


class Matrix(MatrixBase):
    def set(self, ary):
        if ary.shape != self.ioshape:
            raise ValueError('Invalid matrix shape')
        if hasattr(self, '_initval'):
            self._initval = np.asanyarray(ary, dtype=self.dtype)
        else:
            self.backend.wait()
            self._set(ary)


class ConstMatrix(MatrixBase):
    _base_tags = {'const'}

    def __init__(self, backend, dtype, initval, tags):
        super().__init__(backend, dtype, np.asarray(initval).shape, initval,
                         None, tags)
        self._host = np.array(initval, dtype=np.float64)

    def get(self):
        return self._host.astype(self.dtype)


class MatrixSlice:
    """Row/column window onto a matrix (``types.py:169-209``)."""

    def __init__(self, backend, mat, ra, rb, ca, cb):
        self.backend = backend
        self.parent = mat
        if not (0 <= ra < rb <= mat.nrow and 0 <= ca < cb <= mat.ncol):
            raise ValueError('Invalid slice region')
        self.ra, self.rb, self.ca, self.cb = ra, rb, ca, cb
        self.nrow, self.ncol = rb - ra, cb - ca
        self.leaddim = mat.leaddim
        self.dtype, self.itemsize = mat.dtype, mat.itemsize
        self.tags = mat.tags | {'slice'}
        self.basedata = mat.basedata
        self.offset = mat.offset + (ra*mat.leaddim + ca)*mat.itemsize
        self.data = self.basedata + self.offset
        self.traits = (1, self.nrow, self.ncol, self.leaddim, self.dtype)
        self._storage_root = mat.storage_root
        # An (nrow, nvar, neles) style slice keeps its parent's ioshape tail
        if len(mat.ioshape) >= 3:
            nv = mat.ioshape[-2]
            self.ioshape = (self.nrow, nv, (self.ncol // nv))
        else:
            self.ioshape = (self.nrow, self.ncol)

    @property
    def storage_root(self):
        return self._storage_root

    def same_storage(self, other):
        return self.storage_root is other.storage_root


class View:
    """Gather descriptor for interface kernels (``types.py:260-320``).

    ``mapping`` holds, per point, the element offset (from ``basedata``) of
    variable 0 in the AoSoA layout; variable v is at +64*v and, for vector
    views, dimension d at +d*rstride.  The row stride is per point
    (``rstrides_np``); ``rstrides_val`` is its value where all points
    share it, else 0 (a view over element types with different point
    counts), and ``rstrides_mat()`` the per-point device array."""

    def __init__(self, backend, matmap, rmap, cmap, rstridemap, vshape, tags):
        self._backend, self._tags = backend, tags
        self.n = len(matmap)
        self.nvrow = vshape[-2] if len(vshape) == 2 else 1
        self.nvcol = vshape[-1] if len(vshape) >= 1 else 1

        mats = [backend.mats[i] for i in np.unique(matmap)]
        self._mats = mats
        self.basedata = mats[0].basedata
        self.refdtype = mats[0].dtype
        if any(not m.same_storage(mats[0]) for m in mats):
            raise TypeError('All viewed matrices must belong to the same '
                            'storage object')
        if any(m.dtype != self.refdtype for m in mats):
            raise TypeError('Mixed data types are not supported')

        k = backend.soasz
        offset = np.empty(self.n, dtype=np.int64)
        leaddim = np.empty(self.n, dtype=np.int64)
        for m in mats:
            ix = np.where(matmap == m.mid)
            offset[ix] = m.offset // m.itemsize
            leaddim[ix] = m.leaddim
        cmap = np.asarray(cmap, dtype=np.int64)
        rowdisp = np.asarray(rmap, dtype=np.int64)*leaddim
        coldisp = (cmap // k)*k*self.nvcol + cmap % k
        mapping = offset + rowdisp + coldisp
        if mapping.size and mapping.max() >= 2**31:
            raise ValueError('View offsets exceed the int32 index range')
        self.mapping_np = mapping.astype(np.int32)
        self.mapping = backend.const_matrix(self.mapping_np[None, :],
                                            dtype=np.int32, tags=tags)
        if self.nvrow > 1:
            rstrides = np.broadcast_to(np.asarray(rstridemap)*leaddim,
                                       (self.n,))
            if rstrides.size and (mapping + (self.nvrow - 1)*rstrides).max() \
                    >= 2**31:
                raise ValueError('View offsets exceed the int32 index range')
            self.rstrides_np = rstrides.astype(np.int32)
            u = np.unique(self.rstrides_np)
            self.rstrides_val = int(u[0]) if len(u) == 1 else 0
        else:
            self.rstrides_np = None
            self.rstrides_val = 0
        self.rstrides = None

    def rstrides_mat(self):
        """The per-point row strides as a device int32 array (made on first
        use)."""
        if self.rstrides is None:
            self.rstrides = self._backend.const_matrix(
                self.rstrides_np[None, :], dtype=np.int32, tags=self._tags)
        return self.rstrides

    @staticmethod
    def stride_args(*views):
        """(uniform stride, None, ...) where the views share one stride at
        every point, else (0, per-point stride arrays...)."""
        vals = {v.rstrides_val for v in views}
        if len(vals) == 1 and 0 not in vals:
            return (vals.pop(), *[None]*len(views))
        return (0, *[v.rstrides_mat().data for v in views])


# ---------------------------------------------------------------------------
# Kernels and graphs
# ---------------------------------------------------------------------------

class Kernel:
    """A launchable unit (``provider.py:10-66``): ``run(stream)`` launches,
    ``bind(**kw)`` updates runtime arguments."""
    dt = float('nan')
    retval = None

    def __init__(self, fn, rtargs=None, mats=(), views=(), name=''):
        self._fn = fn
        self.rtargs = dict(rtargs or {})
        self.mats, self.views, self.name = mats, views, name
        self.version = 0

    @property
    def rtnames(self):
        return set(self.rtargs)

    # runtime arguments the kernel reads from device memory: name ->
    # (device address of a double, backend); binding one updates that
    # double on the stream before the next launch, so a captured graph
    # holding the kernel stays valid
    devargs = {}

    def bind(self, *args, **kwargs):
        if args:
            raise TypeError('positional bind not supported for this kernel')
        for k, v in kwargs.items():
            if k in self.devargs:
                self._devpending = {**getattr(self, '_devpending', {}),
                                    k: float(v)}
                continue
            if k not in self.rtargs:
                raise KeyError(k)
            if self.rtargs[k] != v:
                self.rtargs[k] = v
                self.version += 1

    def flush_devargs(self, stream):
        """Write bound device-resident arguments that changed (a fill
        kernel on the stream, outside any capture)."""
        pend = getattr(self, '_devpending', None)
        if not pend:
            return
        done = getattr(self, '_devdone', {})
        for k, v in pend.items():
            if done.get(k) != v:
                ptr, be = self.devargs[k]
                be.lib.fill_f64(ptr, 1, v, stream)
                done[k] = v
        self._devdone, self._devpending = done, {}

    def launch(self, stream):
        """The launch alone (what a graph capture records); subclasses
        that implement run() themselves launch through it."""
        if self._fn is None:
            self.run(stream)
        else:
            self._fn(stream, **self.rtargs)

    def run(self, stream):
        self.flush_devargs(stream)
        self._fn(stream, **self.rtargs)


class NullKernel(Kernel):
    def __init__(self):
        super().__init__(lambda stream: None, name='null')


class MetaKernel(Kernel):
    def __init__(self, kernels, splits=None):
        self.kernels = list(kernels)
        self.splits = splits
        super().__init__(self._run_all, name='meta')

    def _run_all(self, stream):
        for k in self.kernels:
            k.run(stream)

    def flush_devargs(self, stream):
        for k in self.kernels:
            k.flush_devargs(stream)

    def launch(self, stream):
        for k in self.kernels:
            k.launch(stream)

    @property
    def version(self):
        return sum(getattr(k, 'version', 0) for k in self.kernels)

    @version.setter
    def version(self, v):
        pass


class Concurrent:
    """Kernel chains with no dependencies between them: each chain runs in
    order, all of them after what precedes the group in a graph and before
    what follows (per element type: the disu of tets and of prisms, their
    divergence kernels).  The reference states such independence as node
    dependencies (``pyfr/backends/base/types.py:343-533``)."""

    def __init__(self, chains):
        self.chains = [list(c) for c in chains if c]

    def kernels(self):
        return [k for c in self.chains for k in c]


def flat_kernels(items):
    """The kernels of a list of kernels and Concurrent groups, in order."""
    out = []
    for k in items:
        out += k.kernels() if isinstance(k, Concurrent) else [k]
    return out


class Graph:
    """Kernel DAG replayed as a hipGraph (``types.py:343-533``,
    ``hip/types.py:79-116``).

    Kernels are recorded in insertion order, which respects the declared
    dependencies on a single stream.  On first run (and whenever a member
    kernel's bound arguments change) the sequence is captured into a HIP
    graph; later runs launch that graph.  The chains of a Concurrent group
    are captured on side streams forked from and joined back to the
    capturing stream, so the graph holds them as parallel branches; run
    eagerly (no graphs), everything runs in order on one stream."""

    def __init__(self, backend):
        self.backend = backend
        self.klist = []
        self.knodes = {}
        self.committed = False
        self._exec = None
        self._versions = None
        self._groups = {}           # klist index -> chain lengths
        self._events = []

    def add(self, kern, deps=[], pdeps=[]):
        if self.committed:
            raise RuntimeError('Can not add nodes to a committed graph')
        if isinstance(kern, Concurrent):
            return self.add_concurrent(kern.chains)
        for d in deps:
            if d not in self.knodes:
                raise ValueError('Dependency not in graph')
        self.knodes[kern] = len(self.klist)
        self.klist.append(kern)

    def add_all(self, kerns, deps=[]):
        for k in kerns:
            self.add(k, deps)

    def add_concurrent(self, chains):
        chains = [list(c) for c in chains if c]
        if len(chains) > 1:
            self._groups[len(self.klist)] = [len(c) for c in chains]
        for c in chains:
            for k in c:
                self.add(k)

    def add_mpi_req(self, req, deps=[]):
        raise NotImplementedError('MPI is not supported on this backend')

    def add_mpi_reqs(self, reqs):
        if reqs:
            raise NotImplementedError('MPI is not supported on this backend')

    def commit(self):
        self.committed = True

    def _event(self, i):
        while len(self._events) <= i:
            e = ct.c_void_p()
            self.backend.lib.event_create(ct.byref(e))
            self._events.append(e.value)
        return self._events[i]

    def _capture(self, stream):
        """Launch the kernels on a capturing stream; the chains of each
        concurrent group beyond the first on side streams that wait for an
        event recorded before the group, and that the stream waits for
        after it."""
        be, lib = self.backend, self.backend.lib
        i, nev, n = 0, 0, len(self.klist)
        while i < n:
            sizes = self._groups.get(i)
            if not sizes:
                self.klist[i].launch(stream)
                i += 1
                continue
            side = be.side_streams(len(sizes) - 1)
            fork = self._event(nev)
            nev += 1
            lib.event_record(fork, stream)
            for s in side:
                lib.stream_wait_event(s, fork)
            for j, sz in enumerate(sizes):
                st = stream if j == 0 else side[j - 1]
                for k in self.klist[i:i + sz]:
                    k.launch(st)
                i += sz
            for s in side:
                join = self._event(nev)
                nev += 1
                lib.event_record(join, s)
                lib.stream_wait_event(stream, join)

    def _cur_versions(self):
        return tuple(k.version for k in self.klist)

    def run(self, stream):
        be = self.backend
        if not be.use_graphs:
            for k in self.klist:
                k.run(stream)
            return
        for k in self.klist:
            k.flush_devargs(stream)
        vers = self._cur_versions()
        if self._exec is None or vers != self._versions:
            if self._exec is not None:
                be.lib.graph_destroy(self._exec)
                self._exec = None
            be.ncaptures += 1
            be.lib.graph_begin(stream)
            try:
                self._capture(stream)
            finally:
                ge = ct.c_void_p()
                be.lib.graph_end(stream, ct.byref(ge))
            self._exec, self._versions = ge.value, vers
        be.lib.graph_launch(self._exec, stream)

    def __del__(self):
        try:
            if self._exec is not None:
                self.backend.lib.graph_destroy(self._exec)
            for e in self._events:
                self.backend.lib.event_destroy(e)
        except Exception:
            pass


# ---------------------------------------------------------------------------
# Providers
# ---------------------------------------------------------------------------

def _ptr(m):
    return None if m is None else m.data


# fused epilogue of each stage-update name
_STAGE_EP = {None: 1, 'rk45': 2, 'rk45-mixed': 3, 'rk4-1': 10, 'rk4-2': 11,
             'rk4-3': 12, 'rk4-4': 13, 'rk4m-1': 14, 'rk4m-2': 15,
             'rk4m-3': 16, 'rk4m-4': 17}


def _check_xa(xa, out, ep):
    """The stage-update banks of a fused epilogue index with out's leading
    dimension, so each must match out's shape, leading dimension and dtype
    (FP64 for the mixed-precision epilogues)."""
    if ep < 2:
        return
    mats = (xa,) if ep in (2, 3) else tuple(xa)
    if ep >= 10 and len(mats) != 2:
        raise ValueError('RK4 stage epilogues take the banks (r1, r0)')
    want = np.float64 if ep in (3, 14, 15, 16, 17) else out.dtype
    for m in mats:
        if (m is None or m.nrow != out.nrow or m.ncol != out.ncol or
                m.leaddim != out.leaddim or m.dtype != want):
            raise ValueError('Stage-update banks must match the output '
                             'shape, leading dimension and precision')


class MulProvider:
    """``mul``: structured (sparsity-pattern matched) or dense operator
    kernels; replaces the GiMMiK/rocBLAS providers."""

    def __init__(self, backend):
        self.backend = backend
        self._plans = {}
        self._plans2 = {}
        self._plans3 = {}
        self._tuned = {}

    def _tune(self, fn, plan, ncol, *bargs):
        """Pick a dense plan's kernel variant for this column count by
        timing (once per plan and shape; replaces the reference's rocBLAS
        solution sweep, pyfr/backends/hip/rocblas.py:139-177).  Runs at
        kernel construction, outside any graph capture."""
        be = self.backend
        if not be.autotune:
            return
        key = (plan, ncol, bargs[1::2])
        if key not in self._tuned:
            be.wait()
            v = ct.c_int()
            fn(plan, ncol, *bargs, 5, ct.byref(v))
            self._tuned[key] = v.value

    def _plan(self, a, dtype):
        key = (id(a), dtype)
        if key not in self._plans:
            # the exact FP64 operator, also for FP32 plans (the library
            # rounds it to the plan's precision)
            host = getattr(a, '_host', None)
            if host is None:
                host = a.get()
            host = np.ascontiguousarray(host, dtype=np.float64)
            plan, kind = ct.c_void_p(), ct.c_int()
            self.backend.lib.mul_plan_create(
                ct.byref(plan), host.shape[0], host.shape[1],
                host.ctypes.data, _dtcode(dtype), ct.byref(kind))
            self._plans[key] = (plan.value, kind.value, a)
        return self._plans[key][:2]

    def mul(self, a, b, out, alpha=1.0, beta=0.0):
        if 'const' not in a.tags:
            raise NotSuitableError('mul requires a constant operator')
        if a.ncol != b.nrow or a.nrow != out.nrow:
            raise ValueError('Incompatible matrices for out = a*b')
        if b.ncol != out.ncol or b.dtype != out.dtype:
            raise ValueError('Incompatible matrices for out = a*b')
        lib = self.backend.lib
        plan, kind = self._plan(a, b.dtype)
        ncol, ldb, ldc = b.ncol, b.leaddim, out.leaddim
        bp, cp = b.data, out.data
        if kind == 2:
            self._tune(lib.mul_plan_tune, plan, ncol, bp, ldb)
        k = Kernel(lambda stream, alpha=alpha, beta=beta: lib.mul(
            plan, ncol, bp, ldb, cp, ldc, alpha, beta, stream),
            rtargs={'alpha': alpha, 'beta': beta}, mats=(a, b, out),
            name='mul')
        k.kind = {1: 'structured', 2: 'dense', 3: 'bf16'}[kind]
        return k

    # -- fused two-operator applications (no reference counterpart: they fuse
    #    kernels the reference runs back to back, see include/frhip.h)
    def _plan2(self, a0, a1, dtype):
        key = (id(a0), id(a1), dtype)
        if key not in self._plans2:
            h0 = np.ascontiguousarray(a0._host, dtype=np.float64)
            h1 = np.ascontiguousarray(a1._host, dtype=np.float64)
            if h0.shape[0] != h1.shape[0]:
                raise ValueError('Operators must have the same row count')
            plan, kind = ct.c_void_p(), ct.c_int()
            code = _dtcode(dtype)
            if self.backend.ops_bf16 and code == 1:
                code = 2                       # FRHIP_F32_BF16OPS
            self.backend.lib.mul2_plan_create(
                ct.byref(plan), h0.shape[0], h0.shape[1], h1.shape[1],
                h0.ctypes.data, h1.ctypes.data, code, ct.byref(kind))
            self._plans2[key] = (plan.value, kind.value, a0, a1)
        return self._plans2[key][:2]

    def _fused(self, a0, a1, b0, b1, out, ep, rcpdjac=None, xa=None):
        for a in (a0, a1):
            if 'const' not in a.tags:
                raise NotSuitableError('fused mul requires constant operators')
        if (a0.ncol != b0.nrow or a1.ncol != b1.nrow or
                a0.nrow != out.nrow or b0.ncol != out.ncol or
                b1.ncol != out.ncol or b0.leaddim != b1.leaddim or
                b0.dtype != out.dtype or b1.dtype != out.dtype):
            raise ValueError('Incompatible matrices for the fused mul')
        plan, kind = self._plan2(a0, a1, out.dtype)
        if out.ncol % (64*5):
            raise NotSuitableError('fused kernels need 5-variable AoSoA data')
        if ep >= 10 and kind != 1:
            raise NotSuitableError('RK4 stage epilogues need a structured '
                                   'plan')
        lib = self.backend.lib
        ncol = out.ncol
        if kind == 2:
            self._tune(lib.mul2_plan_tune, plan, ncol, b0.data, b0.leaddim,
                       b1.data, b1.leaddim)
        _check_xa(xa, out, ep)
        if ep >= 10:
            # RK4 stages: xa = (r1, r0), passed with the step's device
            # scalar as the device pointers {r1, r0, dt}
            xap, dtp = self._ptr_triple(*xa)
            xam = tuple(xa)
        else:
            xap, xam, dtp = None if xa is None else xa.data, (), None
        args = (plan, ncol, b0.data, b0.leaddim, b1.data, b1.leaddim,
                out.data, out.leaddim, ep,
                None if rcpdjac is None else rcpdjac.data,
                0 if rcpdjac is None else _rstride(rcpdjac), xap)

        def fn(stream, dta=0.0, dtb=0.0, last=0):
            lib.mul2(*args, dta, dtb, int(last), stream)
        mats = (a0, a1, b0, b1, out) + (() if rcpdjac is None else
                                        (rcpdjac,)) + xam
        if dtp is None:
            k = Kernel(fn, rtargs={'dta': 0.0, 'dtb': 0.0, 'last': 0},
                       mats=mats, name=f'mul2_ep{ep}')
        else:
            # RK4: dt lives in device memory (bind(dta=) writes it)
            k = Kernel(fn, rtargs={'dtb': 0.0, 'last': 0}, mats=mats,
                       name=f'mul2_ep{ep}')
            k.devargs = {'dta': (dtp, self.backend)}
        k.kind = {1: 'structured', 2: 'dense', 3: 'bf16'}[kind]
        return k

    def mul2(self, a0, a1, b0, b1, out):
        """out = a0 @ b0 + a1 @ b1 (tgradpcoru + tgradcoru)."""
        return self._fused(a0, a1, b0, b1, out, 0)

    def tdivf(self, a0, a1, b0, b1, out, rcpdjac, xa=None, stage=None):
        """tdivtpcorf + tdivtconf + negdivconf, optionally followed by the
        rkvdh2 stage update (stage='rk45': xa = r1, out = r2; 'rk45-mixed':
        xa = FP64 accumulated solution, out = FP32 stage input), or RK4
        stage k's updates (stage='rk4-k': xa = (r1, r0), out = the next
        stage's input, dta = dt; structured plans only).  The stage
        coefficients are bound at run time: bind(dta=, dtb=, last=)."""
        return self._fused(a0, a1, b0, b1, out, _STAGE_EP[stage], rcpdjac,
                           xa)

    def tuning_report(self):
        """The variant the tuner chose for each dense plan and shape
        (0: scalar kernel; 1, 2, 4: matrix-core kernel, wavefronts per
        block)."""
        shapes = {}
        for plan, _, a in self._plans.values():
            shapes[plan] = ('mul', f'{a.nrow}x{a.ncol}')
        for plan, _, a0, a1 in self._plans2.values():
            shapes[plan] = ('mul2', f'{a0.nrow}x{a0.ncol}+{a1.ncol}')
        for plan, a0, a1 in self._plans3.values():
            shapes[plan] = ('tdivf_flux', f'{a0.nrow}x{a0.ncol}+{a1.ncol}')
        out = []
        for key, v in self._tuned.items():
            kind, shape = shapes.get(key[0], ('?', '?'))
            out.append({'kernel': kind, 'operator': shape, 'ncol': key[1],
                        'variant': v})
        return out

    def tdivf_flux(self, a0, a1, u, smats, fn, out, rcpdjac, gamma, xa=None,
                   stage=None):
        """tdisf + tdivtpcorf + tdivtconf + negdivconf (+ rkvdh2 stage) for
        straight-sided elements (Euler): the flux of u is formed in the
        kernel from one S per element (smats (9, neles))."""
        ep = _STAGE_EP[stage]
        if ep >= 14:
            raise NotSuitableError('the precision form\'s RK4 epilogues '
                                   'are built for the hex kernels')
        for a in (a0, a1):
            if 'const' not in a.tags:
                raise NotSuitableError('tdivf_flux requires const operators')
        nupts = u.nrow
        if (a0.ncol != 3*nupts or a1.ncol != fn.nrow or
                a0.nrow != out.nrow or out.ncol != u.ncol or
                fn.ncol != u.ncol or smats.nrow != 9):
            raise ValueError('Incompatible matrices for tdivf_flux')
        if out.ncol % (64*5) or a0.nrow > 80:
            raise NotSuitableError('tdivf_flux needs 5-variable AoSoA data '
                                   'and at most 80 operator rows')
        lib, dt = self.backend.lib, out.dtype
        key = (id(a0), id(a1), dt)
        if key not in self._plans3:
            h0 = np.ascontiguousarray(a0._host, dtype=np.float64)
            h1 = np.ascontiguousarray(a1._host, dtype=np.float64)
            plan = ct.c_void_p()
            lib.tdivf_plan_create(ct.byref(plan), h0.shape[0], nupts,
                                  h1.shape[1], h0.ctypes.data,
                                  h1.ctypes.data, _dtcode(dt))
            self._plans3[key] = (plan.value, a0, a1)
        plan = self._plans3[key][0]
        ncol = out.ncol
        if self.backend.autotune and (plan, ncol) not in self._tuned:
            self.backend.wait()
            v = ct.c_int()
            lib.tdivf_plan_tune(plan, ncol, u.data, u.leaddim, smats.data,
                                smats.leaddim, fn.data, fn.leaddim, gamma, 5,
                                ct.byref(v))
            self._tuned[(plan, ncol)] = v.value
        _check_xa(xa, out, ep)
        if ep >= 10:
            # RK4 stages: xa = (r1, r0) with the step's device scalar
            xap, dtp = self._ptr_triple(*xa)
        else:
            xap, dtp = None if xa is None else xa.data, None
        args = (plan, ncol, u.data, u.leaddim, smats.data, smats.leaddim,
                fn.data, fn.leaddim, out.data, out.leaddim, ep, rcpdjac.data,
                _rstride(rcpdjac), xap)

        def fn_(stream, dta=0.0, dtb=0.0, last=0):
            lib.tdivf_flux_euler(*args, dta, dtb, int(last), gamma, stream)
        if dtp is None:
            k = Kernel(fn_, rtargs={'dta': 0.0, 'dtb': 0.0, 'last': 0},
                       mats=(a0, a1, u, smats, fn, out, rcpdjac),
                       name=f'tdivf_flux_ep{ep}')
        else:
            k = Kernel(fn_, rtargs={'dtb': 0.0, 'last': 0},
                       mats=(a0, a1, u, smats, fn, out, rcpdjac),
                       name=f'tdivf_flux_ep{ep}')
            k.devargs = {'dta': (dtp, self.backend)}
        k.kind = 'dense'
        return k

    def _ptr_triple(self, r1, r0):
        """Device pointers {r1, r0, dt} for an RK4 epilogue, dt a device
        double of its own (each kernel binds its step there).  Returns
        (address of the pointer array, address of dt); both are tracked by
        the backend, which frees them at teardown."""
        be = self.backend
        dtp = be._malloc_impl(8)
        be.lib.fill_f64(dtp, 1, 0.0, None)
        h = np.array((r1.data, r0.data, dtp), dtype=np.uint64)
        p = be._malloc_impl(h.nbytes)
        be.lib.memcpy_htod(p, h.ctypes.data, h.nbytes, None)
        be.lib.stream_sync(None)
        return p, dtp

    def __del__(self):
        try:
            for plan, _, _ in self._plans.values():
                self.backend.lib.mul_plan_destroy(plan)
            for plan, _, _, _ in self._plans2.values():
                self.backend.lib.mul2_plan_destroy(plan)
            for plan, _, _ in self._plans3.values():
                self.backend.lib.tdivf_plan_destroy(plan)
        except Exception:
            pass


class BlasExtProvider:
    """``axnpby`` (with the scaled form), ``copy``, ``zero`` and
    ``reduction`` (``pyfr/backends/base/blasext.py:9-46``,
    ``pyfr/backends/hip/blasext.py:8-136``)."""

    def __init__(self, backend):
        self.backend = backend

    def axnpby(self, *arr, in_scale=(), in_scale_idxs=(), out_scale=()):
        if any(arr[0].traits != x.traits for x in arr[1:]):
            raise ValueError('Incompatible matrix types')
        lib, be = self.backend.lib, self.backend
        nv = len(arr)
        nrow, ncol, ld = arr[0].nrow, arr[0].ncol, arr[0].leaddim
        dt = _dtcode(arr[0].dtype)
        ptrs = (ct.c_void_p*nv)(*[x.data for x in arr])
        coefs = (ct.c_double*nv)()
        scaled = bool(in_scale or out_scale or in_scale_idxs)
        if scaled:
            # per-variable scales (axnpby.mako with makoutil.axnpby_expr)
            if len(arr[0].ioshape) != 3:
                raise ValueError('Scaled axnpby needs (nrow, nvar, neles) '
                                 'matrices')
            nvar = arr[0].ioshape[1]
            if in_scale_idxs and not in_scale:
                raise ValueError('in_scale_idxs without in_scale')
            for sc in (in_scale, out_scale):
                if sc and len(sc) != nvar:
                    raise ValueError('One scale per variable is required')
            if any(not 0 <= i < nv for i in in_scale_idxs):
                raise ValueError('Invalid in_scale_idxs')
            inmask = sum(1 << i for i in set(in_scale_idxs))
            insc = (ct.c_double*nvar)(*in_scale) if in_scale else None
            outsc = (ct.c_double*nvar)(*out_scale) if out_scale else None

        class AxnpbyKernel(Kernel):
            def bind(self, *consts):
                if len(consts) != nv:
                    raise ValueError('axnpby expects one constant per array')
                if tuple(coefs) != consts:
                    coefs[:] = consts
                    self.version += 1

            def run(self, stream):
                if scaled:
                    lib.axnpby_scaled(dt, nrow, ncol, ld, nv, nvar, ptrs,
                                      coefs, inmask, insc, outsc, stream)
                else:
                    lib.axnpby(dt, nrow, ncol, ld, nv, ptrs, coefs, stream)

        return AxnpbyKernel(None, mats=arr, name='axnpby')

    def reduction(self, rop, exprs, vvars, svars=[], pvars={}):
        """``reduction(rop, exprs, vvars, svars, pvars)``
        (``pyfr/backends/base/blasext.py:21-46``): each expression of the
        ``vvars`` arrays' values, the per-variable constants ``pvars`` and
        the bound scalars ``svars`` is reduced (``'sum'`` or ``'max'``) over
        every point and variable; ``retval`` after the kernel has run (and
        the stream has been waited on) holds one value per expression.
        Accumulation is in FP64 for every precision.  Extension: an array of
        ioshape (nrow, neles) or (nrow, 1, neles) is broadcast over the
        variables (e.g. quadrature weights), and ``result`` holds the
        (nexprs, nvar) values before the reduction over variables.  All
        device buffers are made here, so ``run`` allocates nothing."""
        from pyfr_amd.bcexpr import compile_exprs
        be, lib = self.backend, self.backend.lib
        if rop not in ('sum', 'max'):
            raise ValueError('Reduction op must be sum or max')
        if not vvars:
            raise ValueError('A reduction needs at least one array')
        names = list(vvars)
        mats = [vvars[n] for n in names]
        full = [m for m in mats if len(m.ioshape) == 3 and m.ioshape[1] > 1]
        ref = full[0] if full else mats[0]
        if len(ref.ioshape) == 3:
            nrow, nvar, neles = ref.ioshape
        else:
            (nrow, neles), nvar = ref.ioshape, 1
        bcast = 0
        for i, m in enumerate(mats):
            if m.dtype != ref.dtype:
                raise ValueError('Incompatible matrix types')
            if m is ref or m.traits == ref.traits:
                continue
            shp = m.ioshape
            if shp in ((nrow, neles), (nrow, 1, neles)) and nvar > 1:
                bcast |= 1 << i
            else:
                raise ValueError('Incompatible matrix types')
        pnames = list(pvars)
        parr = np.array([np.broadcast_to(np.asarray(pvars[n], dtype=float),
                                         (nvar,)) for n in pnames],
                        dtype=np.float64).reshape(len(pnames), nvar)
        slots = names + pnames + list(svars)
        if len(set(slots)) != len(slots):
            raise ValueError('Duplicate variable names')
        code, offs, pool = compile_exprs(
            list(exprs), {}, varmap={n: i for i, n in enumerate(slots)})
        prog = ct.c_void_p()
        lib.expr_program_create(ct.byref(prog), len(exprs), len(slots),
                                code.ctypes.data, offs.ctypes.data,
                                pool.ctypes.data, len(pool))
        be._programs.append(prog.value)
        nw = ct.c_long()
        lib.reduction_workspace(prog, nrow, neles, nvar, ct.byref(nw))
        nout = len(exprs)*nvar
        work = be._malloc_impl(8*max(nw.value, 1))
        out = be._malloc_impl(8*nout)
        pv = None
        if pnames:
            pv = be._malloc_impl(parr.nbytes)
            lib.memcpy_htod(pv, parr.ctypes.data, parr.nbytes, be.stream)
            be.wait()
        host = be._host_array(nout)
        dtc = _dtcode(ref.dtype)
        mptr = (ct.c_void_p*len(mats))(*[m.data for m in mats])
        lds = (ct.c_long*len(mats))(*[m.leaddim for m in mats])
        nsv = len(svars)
        sv = (ct.c_double*max(nsv, 1))()
        ropc = 0 if rop == 'sum' else 1
        nexprs = len(exprs)
        reducer = np.max if rop == 'max' else np.sum

        class ReductionKernel(Kernel):
            @property
            def result(self):
                return host.reshape(nexprs, nvar).copy()

            @property
            def retval(self):
                return reducer(host.reshape(nexprs, nvar), axis=1)

            def bind(self, *consts):
                if len(consts) != nsv:
                    raise ValueError('reduction expects one value per svar')
                if tuple(sv[:nsv]) != tuple(consts):
                    sv[:nsv] = consts
                    self.version += 1

            def run(self, stream):
                lib.reduction(prog, dtc, ropc, nrow, neles, nvar, len(mats),
                              mptr, lds, bcast, pv, len(pnames), sv, nsv,
                              work, nw.value, out, stream)
                lib.memcpy_dtoh_async(host.ctypes.data, out, 8*nout, stream)

        return ReductionKernel(None, mats=tuple(mats), name='reduction')

    def copy(self, dst, src):
        if dst.traits != src.traits:
            raise ValueError('Incompatible matrix types')
        lib, n = self.backend.lib, dst.nrow*dst.leaddim*dst.itemsize
        d, s = dst.data, src.data
        return Kernel(lambda stream: lib.memcpy_dtod(d, s, n, stream),
                      mats=(dst, src), name='copy')

    def zero(self, m):
        lib, n, d = self.backend.lib, m.nrow*m.leaddim*m.itemsize, m.data
        return Kernel(lambda stream: lib.memset(d, 0, n, stream), mats=(m,),
                      name='zero')


class PointwiseProvider:
    """Hand-written pointwise kernels addressed by the reference's kernel
    names.  ``register(mod)`` accepts the reference's template module paths
    and checks that a native kernel exists for them."""

    _known = {
        'pyfr.solvers.euler.kernels.tflux': 'tflux',
        'pyfr.solvers.navstokes.kernels.tflux': 'tflux',
        'pyfr.solvers.euler.kernels.intcflux': 'intcflux',
        'pyfr.solvers.navstokes.kernels.intcflux': 'intcflux',
        'pyfr.solvers.navstokes.kernels.intconu': 'intconu',
        'pyfr.solvers.euler.kernels.bccflux': 'bccflux',
        'pyfr.solvers.navstokes.kernels.bccflux': 'bccflux',
        'pyfr.solvers.navstokes.kernels.bcconu': 'bcconu',
        'pyfr.solvers.baseadvec.kernels.negdivconf': 'negdivconf',
        'pyfr.solvers.baseadvecdiff.kernels.gradcoru': 'gradcoru',
        'pyfr.integrators.explicit.kernels.rkvdh2': 'rkvdh2',
    }

    def __init__(self, backend):
        self.backend = backend
        self.registered = set()

    def register(self, mod):
        if mod not in self._known:
            raise NotSuitableError(f'No native kernel for {mod}')
        self.registered.add(self._known[mod])

    @staticmethod
    def _base(tplargs):
        """The base-state form's {rho0, E0} (tplargs 'base'), or None."""
        b = tplargs.get('base')
        return None if b is None else (ct.c_double*2)(*b)

    # -- tflux (Euler and NS) ------------------------------------------------
    def tflux(self, tplargs, dims, u, f, smats=None, verts=None, upts=None,
              gradu=None, rcpdjac=None, artvisc_vtx=None, **kw):
        lib, be = self.backend.lib, self.backend
        base = self._base(tplargs)
        npts, neles = dims
        ktype = tplargs['ktype']
        # geometry mode: 1 hex vertices, 2 one S per element, 0 stored S
        mode = 1 if 'linear' in ktype else 2 if 'affine' in ktype else 0
        linear = mode == 1
        dt = _dtcode(u.dtype)
        geo = verts if linear else smats
        uptsp = _ptr(upts) if linear else None
        c = tplargs['c']
        ns = 'visc_corr' in tplargs
        if ns:
            consts = (ct.c_double*8)(
                c['gamma'], c.get('mu', 0.0), c.get('Pr', 1.0),
                c.get('ldg-beta', 0.0), c.get('ldg-tau', 0.0),
                1.0 if tplargs.get('visc_corr') == 'sutherland' else 0.0,
                c.get('cpTref', 1.0), c.get('cpTs', 0.0))
        if ns and 'fused' not in ktype:
            if base is not None:
                raise NotSuitableError('the base-state form has no unfused '
                                       'NS flux')
            # f holds the physical gradient on entry (tflux.mako, unfused)
            args = (dt, mode, npts, neles, u.data, u.leaddim, f.data,
                    f.leaddim, geo.data, geo.leaddim, uptsp, consts, None)
            return Kernel(lambda stream: lib.tflux_ns_grad(*args, stream),
                          mats=(u, f, geo), name='tflux_ns_grad')
        if 'fused' in ktype:
            rp = None if linear else _ptr(rcpdjac)
            ldr = 0 if linear else _rstride(rcpdjac)
            args = (dt, mode, npts, neles, u.data, u.leaddim, f.data,
                    f.leaddim, gradu.data, gradu.leaddim, geo.data,
                    geo.leaddim, uptsp, rp, ldr, consts, None, base)
            return Kernel(lambda stream: lib.tflux_ns(*args, stream),
                          mats=(u, f, gradu, geo), name='tflux_ns')
        args = (dt, mode, npts, neles, u.data, u.leaddim, f.data,
                f.leaddim, geo.data, geo.leaddim, uptsp, c['gamma'], base)
        return Kernel(lambda stream: lib.tflux_euler(*args, stream),
                      mats=(u, f, geo), name='tflux')

    # -- gradcoru (baseadvecdiff/kernels/gradcoru.mako) -------------------------
    def gradcoru(self, tplargs, dims, gradu, smats=None, rcpdjac=None,
                 verts=None, upts=None, **kw):
        lib = self.backend.lib
        npts, neles = dims
        ktype = tplargs['ktype']
        mode = 1 if 'linear' in ktype else 2 if 'affine' in ktype else 0
        geo = verts if mode == 1 else smats
        uptsp = _ptr(upts) if mode == 1 else None
        rp = None if mode == 1 else _ptr(rcpdjac)
        ldr = 0 if mode == 1 else _rstride(rcpdjac)
        args = (_dtcode(gradu.dtype), mode, npts, neles, gradu.data,
                gradu.leaddim, geo.data, geo.leaddim, uptsp, rp, ldr)
        return Kernel(lambda stream: lib.gradcoru(*args, stream),
                      mats=(gradu, geo), name='gradcoru')

    # -- negdivconf ----------------------------------------------------------
    def negdivconf(self, tplargs, dims, tdivtconf, rcpdjac, ploc=None, u=None,
                   **kw):
        if tplargs.get('src_macros'):
            raise NotSuitableError('source terms are not supported')
        lib = self.backend.lib
        npts, neles = dims
        args = (_dtcode(tdivtconf.dtype), npts, neles, tdivtconf.data,
                tdivtconf.leaddim, rcpdjac.data, _rstride(rcpdjac))
        return Kernel(lambda stream: lib.negdivconf(*args, stream),
                      mats=(tdivtconf, rcpdjac), name='negdivconf')

    # -- interfaces ----------------------------------------------------------
    @staticmethod
    def _rs(tplargs, dtype):
        rs = tplargs['rsolver']
        ids = {'rusanov': 0, 'hllc': 1, 'hll': 2, 'roe': 3, 'roem': 4,
               'exact': 5}
        if rs not in ids:
            raise NotSuitableError(f'Riemann solver {rs} not available')
        return ids[rs]

    @staticmethod
    def _pmin(tplargs):
        """The exact solver's pressure floor (``p_min`` template argument,
        pyfr/solvers/euler/inters.py:12-26); -1 for the default."""
        pm = tplargs.get('p_min')
        return -1.0 if pm is None else float(pm)

    def intcflux(self, tplargs, dims, ul, ur, nl, gradul=None, gradur=None,
                 artvisc=None, **kw):
        lib = self.backend.lib
        (n,) = dims
        if ul.basedata != ur.basedata:
            raise ValueError('Interface views must share storage')
        dt = _dtcode(ul.refdtype)
        rs = self._rs(tplargs, ul.refdtype)
        c = tplargs['c']
        base = self._base(tplargs)
        if base is not None and rs != 0:
            raise NotSuitableError('the base-state form takes the Rusanov '
                                   'solver')
        if gradul is None:
            args = (dt, rs, n, ul.basedata, ul.mapping.data, ur.mapping.data,
                    nl.data, nl.leaddim, c['gamma'], self._pmin(tplargs),
                    base)
            return Kernel(lambda stream: lib.intcflux_euler(*args, stream),
                          views=(ul, ur), mats=(nl,), name='intcflux')
        if artvisc is not None:
            raise NotSuitableError('artificial viscosity is not supported')
        gs = View.stride_args(gradul, gradur)
        consts = (ct.c_double*8)(
            c['gamma'], c['mu'], c['Pr'], c['ldg-beta'], c['ldg-tau'],
            1.0 if tplargs.get('visc_corr') == 'sutherland' else 0.0,
            c.get('cpTref', 1.0), c.get('cpTs', 0.0))
        args = (dt, rs, n, ul.basedata, ul.mapping.data, ur.mapping.data,
                gradul.basedata, gradul.mapping.data, gradur.mapping.data,
                *gs, nl.data, nl.leaddim, consts, self._pmin(tplargs), None,
                base)
        return Kernel(lambda stream: lib.intcflux_ns(*args, stream),
                      views=(ul, ur, gradul, gradur), mats=(nl,),
                      name='intcflux_ns')

    def intconu(self, tplargs, dims, ulin, urin, ulout, urout, **kw):
        lib = self.backend.lib
        (n,) = dims
        args = (_dtcode(ulin.refdtype), n, ulin.basedata, ulin.mapping.data,
                urin.mapping.data, ulout.basedata, ulout.mapping.data,
                urout.mapping.data, tplargs['c']['ldg-beta'])
        return Kernel(lambda stream: lib.intconu(*args, stream),
                      views=(ulin, urin, ulout, urout), name='intconu')

    # -- boundary interfaces -------------------------------------------------
    BC_TYPES = {'sup-in-fa': 0, 'sup-out-fn': 1, 'char-riem-inv': 2,
                'slp-adia-wall': 3, 'no-slp-adia-wall': 4,
                'no-slp-isot-wall': 5, 'sub-in-frv': 6,
                'sub-in-ftpttang': 7, 'sub-out-fp': 8}

    def _bc(self, tplargs, euler):
        bct = tplargs['bctype']
        if bct not in self.BC_TYPES or (euler and self.BC_TYPES[bct] > 3):
            raise NotSuitableError(f'Boundary type {bct} not available')
        c = tplargs['c']
        vc = list(c.get('vc', (0.0, 0.0, 0.0)))
        consts = (ct.c_double*14)(
            c['gamma'], c.get('mu', 0.0), c.get('Pr', 1.0),
            c.get('ldg-beta', 0.0), c.get('ldg-tau', 0.0),
            1.0 if tplargs.get('visc_corr') == 'sutherland' else 0.0,
            c.get('cpTref', 1.0), c.get('cpTs', 0.0), c.get('cpTw', 0.0),
            c.get('cpTt', 0.0), c.get('pt', 0.0), *vc)
        return self.BC_TYPES[bct], consts

    def bccflux(self, tplargs, dims, ul, nl, gradul=None, artvisc=None,
                bcvals=None, **kw):
        """Boundary common flux (``euler/kernels/bccflux.mako``,
        ``navstokes/kernels/bccflux.mako``).  ``bcvals``: the boundary's
        prescribed values per point (rows as in ``frhip.h``), which the
        reference splices into the kernel as ploc expressions."""
        lib = self.backend.lib
        (n,) = dims
        dt = _dtcode(ul.refdtype)
        rs = self._rs(tplargs, ul.refdtype)
        euler = gradul is None and 'visc_corr' not in tplargs
        bct, consts = self._bc(tplargs, euler)
        bv = None if bcvals is None else bcvals.data
        ldb = 0 if bcvals is None else bcvals.leaddim
        if euler:
            args = (dt, bct, rs, n, ul.basedata, ul.mapping.data, nl.data,
                    nl.leaddim, bv, ldb, consts, self._pmin(tplargs))
            return Kernel(lambda stream: lib.bccflux_euler(*args, stream),
                          views=(ul,), mats=(nl, bcvals), name='bccflux')
        if artvisc is not None:
            raise NotSuitableError('artificial viscosity is not supported')
        g = (None, None, 0, None) if gradul is None else \
            (gradul.basedata, gradul.mapping.data,
             *View.stride_args(gradul))
        args = (dt, bct, rs, n, ul.basedata, ul.mapping.data, *g, nl.data,
                nl.leaddim, bv, ldb, consts, self._pmin(tplargs))
        return Kernel(lambda stream: lib.bccflux_ns(*args, stream),
                      views=(ul, gradul), mats=(nl, bcvals),
                      name='bccflux_ns')

    def bcconu(self, tplargs, dims, ulin, ulout, nlin, bcvals=None, **kw):
        """``navstokes/kernels/bcconu.mako:6-15``."""
        lib = self.backend.lib
        (n,) = dims
        bct, consts = self._bc(tplargs, False)
        bv = None if bcvals is None else bcvals.data
        ldb = 0 if bcvals is None else bcvals.leaddim
        args = (_dtcode(ulin.refdtype), bct, n, ulin.basedata,
                ulin.mapping.data, ulout.basedata, ulout.mapping.data,
                nlin.data, nlin.leaddim, bv, ldb, consts)
        return Kernel(lambda stream: lib.bcconu(*args, stream),
                      views=(ulin, ulout), mats=(nlin, bcvals),
                      name='bcconu')

    # -- rkvdh2 --------------------------------------------------------------
    def rkvdh2(self, tplargs, dims, r1, r2, rold=None, rerr=None, dt=None,
               **kw):
        lib = self.backend.lib
        stage, nst = tplargs['stage'], tplargs['nstages']
        a, b = tplargs['a'], tplargs['b']
        last = stage == nst - 1
        aa = 0.0 if last else a[stage]
        bb = b[stage]
        d = _dtcode(r1.dtype)
        nrow, ncol, ld = r1.nrow, r1.ncol, r1.leaddim
        if tplargs.get('errest'):
            # rkvdh2.mako with its error estimate: rold, rerr registers
            if rold is None or rerr is None:
                raise ValueError('errest needs rold and rerr')
            if any(m.traits != r1.traits for m in (r2, rold, rerr)):
                raise ValueError('Incompatible matrix types')
            ee = tplargs['e'][stage]

            def fne(stream, dt=0.0):
                lib.rkvdh2_errest(d, nrow, ncol, ld, r1.data, r2.data,
                                  rold.data, rerr.data, dt, aa, bb, ee,
                                  stage, int(last), stream)
            return Kernel(fne, rtargs={'dt': 0.0},
                          mats=(r1, r2, rold, rerr), name='rkvdh2')
        if r2.traits != r1.traits:
            raise ValueError('Incompatible matrix types')

        def fn(stream, dt=0.0):
            lib.rkvdh2(d, nrow, ncol, ld, r1.data, r2.data, dt*aa, dt*bb,
                       int(last), stream)
        return Kernel(fn, rtargs={'dt': 0.0}, mats=(r1, r2), name='rkvdh2')


# ---------------------------------------------------------------------------
# Backend
# ---------------------------------------------------------------------------

class AMDBackend:
    """The MI355X backend (``backends/base/backend.py:32-225``,
    ``hip/base.py:9-140``)."""
    name = 'hip-amd'
    blocks = False
    alignb = 256
    soasz = 64
    csubsz = 64

    def __init__(self, cfg):
        self.cfg = cfg
        prec = cfg.get('backend', 'precision', 'double')
        if prec not in {'single', 'double', 'mixed'}:
            raise ValueError('Backend precision must be single, double or '
                             'mixed')
        # 'mixed': FP32 RHS arithmetic (fpdtype) with FP64 time-level state
        # (state_dtype), RK stage updates and reductions
        self.precision = prec
        self.fpdtype = np.float64 if prec == 'double' else np.float32
        self.state_dtype = np.float32 if prec == 'single' else np.float64
        self.fpdtype_eps = float(np.finfo(self.fpdtype).eps)
        self.fpdtype_max = float(np.finfo(self.fpdtype).max)
        if cfg.get('backend', 'memory-model', 'normal') != 'normal':
            raise ValueError('Only the normal memory model is supported')
        self.ixdtype = np.int32
        self.use_graphs = cfg.getbool('backend-hip', 'graphs', True)
        # fuse element-local operator chains where a fused kernel exists
        self.fused = cfg.getbool('backend-hip', 'fused', True)
        # 'bf16': the fused gradient/divergence operator applications run
        # with both matrix-core operands rounded to bfloat16 (FP32 data and
        # accumulation); an accuracy experiment of the precision form
        ops = cfg.get('backend-hip', 'operator-inputs', 'native')
        if ops not in {'native', 'bf16'}:
            raise ValueError('operator-inputs must be native or bf16')
        if ops == 'bf16' and self.fpdtype != np.float32:
            raise ValueError('bf16 operator inputs need FP32 arithmetic')
        self.ops_bf16 = ops == 'bf16'
        # straight-sided Euler elements: fold tdisf into the divergence
        self.fuse_flux = cfg.getbool('backend-hip', 'fuse-flux', True)
        # time the variants of dense operator plans at kernel construction
        self.autotune = cfg.getbool('backend-hip', 'autotune', True)
        # padding of every (nrow, nvar, neles) array's rows, in elements (a
        # multiple of 64): rows of 5*64^3-element banks are multiples of
        # 2 MiB apart, so the rows a wavefront reads together fall on the
        # same HBM channels; 64 elements of padding measured +1.7 points of
        # peak on the NS flux kernel and +0.5 % on the 64^3 RHS
        self.row_pad = cfg.getint('backend-hip', 'row-pad', 64)
        if self.row_pad < 0 or self.row_pad % self.soasz:
            raise ValueError('row-pad must be a non-negative multiple of 64')

        self.lib = get_lib()
        self.lib.init(cfg.getint('backend-hip', 'device-id', 0))
        s = ct.c_void_p()
        self.lib.stream_create(ct.byref(s))
        self.stream = s.value
        # side streams for the concurrent branches of captured graphs
        # (default priority, like the main stream)
        self._side = []

        self.mats = WeakValueDictionary()
        self._mat_counter = it.count()
        self._extents = {}
        self._all_allocs = []
        self._host_allocs = []
        self.ncaptures = 0          # graph captures (re-captures included)
        self._programs = []
        self._mem_current = 0
        self._mem_peak = 0

        self.pointwise = PointwiseProvider(self)
        self._providers = [self.pointwise, BlasExtProvider(self),
                           MulProvider(self)]

    # -- memory
    def _extent_alloc_bytes(self, nbytes):
        return nbytes - (nbytes % -self.alignb)

    def _malloc_impl(self, nbytes):
        p = ct.c_void_p()
        self.lib.malloc(ct.byref(p), nbytes)
        self._all_allocs.append(p.value)
        self._mem_current += nbytes
        self._mem_peak = max(self._mem_peak, self._mem_current)
        return p.value

    def _host_array(self, n, dtype=np.float64):
        """A page-locked host array of n values (freed with the backend),
        the target of asynchronous device-to-host copies."""
        nb = max(int(n)*np.dtype(dtype).itemsize, 8)
        p = ct.c_void_p()
        self.lib.host_malloc(ct.byref(p), nb)
        self._host_allocs.append(p.value)
        buf = (ct.c_char*nb).from_address(p.value)
        return np.frombuffer(buf, dtype=dtype, count=int(n))

    def malloc(self, obj, extent):
        nb = self._extent_alloc_bytes(obj.nbytes)
        if extent is None:
            ext = Extent()
            ext.reserve(obj, nb)
            ext.commit(self._malloc_impl)
        else:
            ext = self._extents.setdefault(extent, Extent(extent))
            ext.reserve(obj, nb)

    def commit(self):
        for ext in self._extents.values():
            ext.commit(self._malloc_impl)
        self._extents.clear()

    def memory_info(self):
        return (self._mem_current, self._mem_peak)

    def _record(self, m):
        m.mid = next(self._mat_counter)
        self.mats[m.mid] = m
        return m

    def const_matrix(self, initval, dtype=None, tags=set()):
        dtype = dtype or self.fpdtype
        initval = np.asarray(initval)
        return self._record(ConstMatrix(self, dtype, initval, tags))

    def matrix(self, ioshape, initval=None, extent=None, tags=set(),
               dtype=None):
        dtype = dtype or self.fpdtype
        return self._record(Matrix(self, dtype, ioshape, initval, extent,
                                   tags))

    def matrix_slice(self, mat, ra, rb, ca, cb):
        return self._record(MatrixSlice(self, mat, ra, rb, ca, cb))

    def view(self, matmap, rmap, cmap, rstridemap=1, vshape=(), tags=set()):
        return View(self, np.asarray(matmap), np.asarray(rmap),
                    np.asarray(cmap), rstridemap, vshape, tags)

    def xchg_matrix(self, *a, **k):
        raise NotImplementedError('MPI exchange is not supported (one GPU)')

    xchg_view = xchg_matrix_for_view = xchg_matrix

    # -- kernels
    def kernel(self, name, *args, **kwargs):
        declined = False
        for prov in self._providers:
            meth = getattr(prov, name, None)
            if meth:
                try:
                    return meth(*args, **kwargs)
                except NotSuitableError:
                    declined = True
        if declined:
            raise NotSuitableError(f'No provider suits kernel {name!r} here')
        raise KeyError(f'Kernel {name!r} has no providers')

    def ordered_meta_kernel(self, kerns):
        return MetaKernel(kerns)

    def unordered_meta_kernel(self, kerns, splits=None):
        return MetaKernel(kerns, splits)

    def graph(self):
        return Graph(self)

    def side_streams(self, n):
        """n streams besides the main one, created on first use."""
        while len(self._side) < n:
            st = ct.c_void_p()
            self.lib.stream_create(ct.byref(st))
            self._side.append(st.value)
        return self._side[:n]

    def run_kernels(self, kernels, wait=False):
        for k in kernels:
            k.run(self.stream)
        if wait:
            self.wait()

    def run_graph(self, graph, wait=False):
        graph.run(self.stream)
        if wait:
            self.wait()

    def wait(self):
        self.lib.stream_sync(self.stream)

    @contextmanager
    def region(self, name):
        yield

    def __del__(self):
        try:
            self.lib.stream_sync(self.stream)
            for p in self._all_allocs:
                self.lib.free(p)
            for p in self._host_allocs:
                self.lib.host_free(p)
            for p in self._programs:
                self.lib.expr_program_destroy(p)
            for st in self._side:
                self.lib.stream_destroy(st)
            self.lib.stream_destroy(self.stream)
        except Exception:
            pass
