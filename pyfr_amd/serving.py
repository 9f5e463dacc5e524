"""Continuous-batching serving of independent periodic-box solves.

The reference has no serving mode (it runs one mesh per process).  This
module batches many independent solves onto one MI355X:

* **Classes and pools.**  A solve is (N, p, Mach, Re-or-inviscid); solves of
  one (system, order) class share a pool: device arrays (register banks,
  scratch, geometry) with room for ``ecap`` elements.  A solve owns a
  64-element-aligned range of the pool, so every element kernel of the class
  advances all of its solves in one launch, and the interface kernels run
  over the concatenated flux-point pairs of all of them.  Per-solve
  parameters (time step, viscosity) are per-element / per-pair arrays.
* **Admission by HBM footprint.**  Pools are preallocated from an HBM budget
  when the engine starts; a request is admitted when a free, contiguous
  range of its size exists (first fit over 64-element chunks), otherwise it
  waits in the queue.  Finished solves free their range immediately.
* **Graphs.**  One RK4 step of a pool (4 RHS, 6 stage axpys) is captured
  into a hipGraph and replayed; it is re-captured only when the pool's
  membership changes.  Each pool has its own stream, so pools overlap.
* **Continuous batching.**  The loop admits arrivals, steps every non-empty
  pool once, retires solves that reached their step count (copying their
  state to the host), and repeats; requests arriving while others run join
  at the next step boundary.

All kernels are the library's hex RHS kernels.  By default a step runs the
fused sequence: ``frhip_mul`` (disu), the interface kernels, ``frhip_mul2``
(gradient, NS), ``frhip_tflux_*``, and the divergence ``frhip_mul2`` whose
epilogue (ep 20-23) does the RK4 stage updates with each element's own dt;
``fused=False`` runs the reference's unfused sequence with
``frhip_axnpby_pool``.  Each element and pair is computed exactly as when
the solve runs alone, so results do not depend on the mix.

Problem definition (cfg0's analytic field at Mach Ma, sound speed 1):
rho = 1, (u, v, w) = Ma (sin x cos y cos z, -cos x sin y cos z, 0),
p = 1/gamma + Ma^2 (cos 2x + cos 2y)(cos 2z + 2)/16, mu = Ma/Re (unit length
and density), RK4 with dt = 0.4 h / ((p+1)^2 (1 + Ma)), h = 2 pi / N.
"""
from dataclasses import dataclass, field
import ctypes as ct
import math
import time

import numpy as np

from pyfr_amd.backend import _dtcode
from pyfr_amd.inifile import Inifile
from pyfr_amd.mesh import BoxMesh
from pyfr_amd.shapes import HexShape
from pyfr_amd.solvers import FluidElements

GAMMA, PR = 1.4, 0.71
SOA = 64


@dataclass
class SolveRequest:
    rid: int
    N: int
    order: int
    mach: float
    re: float = None           # None: inviscid (Euler)
    nsteps: int = 200
    arrival: float = 0.0
    # filled by the engine
    admit_time: float = None
    finish_time: float = None
    result: np.ndarray = field(default=None, repr=False)

    @property
    def system(self):
        return 'euler' if self.re is None else 'navier-stokes'

    @property
    def neles(self):
        return self.N**3

    @property
    def ndofs(self):
        return self.neles*(self.order + 1)**3*5

    @property
    def dt(self):
        h = 2*math.pi/self.N
        return 0.4*h/((self.order + 1)**2*(1 + self.mach))

    @property
    def mu(self):
        return 0.0 if self.re is None else self.mach/self.re


def poisson_requests(seed=0, rate=50.0, count=100, nsteps=200):
    """Seeded Poisson stream: N ~ U{4..24}, p ~ U{1..4}, Mach ~ U[0.05,
    0.3], Re ~ U[100, 1600] or inviscid with probability 0.3."""
    rng = np.random.default_rng(seed)
    t, out = 0.0, []
    for i in range(count):
        t += rng.exponential(1/rate)
        N = int(rng.integers(4, 25))
        p = int(rng.integers(1, 5))
        mach = float(rng.uniform(0.05, 0.3))
        re = None if rng.random() < 0.3 else float(rng.uniform(100, 1600))
        out.append(SolveRequest(i, N, p, mach, re, nsteps, t))
    return out


# ---------------------------------------------------------------------------
# Per-(N, p) host geometry, cached across requests
# ---------------------------------------------------------------------------

class _BoxData:
    """Host-side data of an N^3 box at order p: solution-point coordinates,
    vertices, 1/|J|, stored metric (p < 2), interface pairs and normals."""

    def __init__(self, N, order):
        mesh = BoxMesh(N)
        cfg = Inifile.from_dict({'solver': {'order': order},
                                 'constants': {'gamma': GAMMA}})
        e = FluidElements(HexShape, mesh.spts['hex'], cfg)
        sh = e.basis
        self.N, self.order, self.neles = N, order, e.neles
        self.nupts, self.nfpts = sh.nupts, sh.nfpts
        self.ploc = e.geom.ploc_at(sh.upts)               # (Nu, 3, E)
        self.verts = mesh.spts['hex'].transpose(0, 2, 1)   # (8, 3, E)
        self.rcpdjac = e.rcpdjac_at_np('upts')             # (Nu, E)
        self.smats = (e.geom.smat_at(sh.upts) if order < 2 else None)
        # interface pairs (left fpt/elem, right fpt/elem) ordered by the
        # left side's storage address (row-major in fpt, then element)
        sf = e.srtd_face_fpts
        lf, le, rf, re_ = [], [], [], []
        for (_, el, fl), (_, er, fr) in zip(*mesh.con):
            lf.append(sf[fl][el])
            rf.append(sf[fr][er])
            le.append(np.full(len(sf[fl][el]), el))
            re_.append(np.full(len(sf[fr][er]), er))
        lf, le = np.concatenate(lf), np.concatenate(le)
        rf, re_ = np.concatenate(rf), np.concatenate(re_)
        order_ = np.lexsort((le, lf))
        self.lf, self.le = lf[order_], le[order_]
        self.rf, self.re = rf[order_], re_[order_]
        pn = e._pnorm_fpts                                 # (Nf, E, 3)
        self.nl = pn[self.lf, self.le].T.copy()            # (3, npairs)
        self.npairs = len(self.lf)

    def ics(self, mach):
        x, y, z = self.ploc.transpose(1, 0, 2)
        rho = np.ones_like(x)
        u = mach*np.sin(x)*np.cos(y)*np.cos(z)
        v = -mach*np.cos(x)*np.sin(y)*np.cos(z)
        w = np.zeros_like(x)
        p = 1/GAMMA + mach**2*(np.cos(2*x) + np.cos(2*y))*(np.cos(2*z) + 2)/16
        E = p/(GAMMA - 1) + 0.5*rho*(u*u + v*v + w*w)
        return np.stack([rho, rho*u, rho*v, rho*w, E], axis=1)


# ---------------------------------------------------------------------------
# Pools
# ---------------------------------------------------------------------------

class Pool:
    """Device storage and kernels for one (system, order) class."""

    def __init__(self, backend, system, order, ecap, fused=True):
        self.be = be = backend
        self.fused = fused
        self.lib = backend.lib
        self.system, self.order = system, order
        self.ns = system == 'navier-stokes'
        self.ecap = ecap - ecap % -SOA
        sh = self.shape = HexShape(order)
        self.nupts, self.nfpts = Nu, Nf = sh.nupts, sh.nfpts
        self.linear = order >= 2
        E = self.ecap
        dt_ = be.fpdtype
        m = lambda shape: be.matrix(shape, tags={'align'})
        self.banks = [m((Nu, 5, E)) for _ in range(3)]
        self.scal_fpts = m((Nf, 5, E))
        self.vect_upts = m((3, Nu, 5, E))
        if self.ns:
            self.vect_fpts = m((3, Nf, 5, E))
            self.grad_upts = m((3, Nu, 5, E))
            self.comm = self.vect_fpts.slice(0, Nf)
        if self.linear:
            self.verts = m((8, 3, E))
        else:
            self.smats = m((3, Nu, 3, E))
        self.rcpdjac = m((Nu, E))
        self.dt_e = m((1, E))
        self.mu_e = m((1, E))
        self.upts = be.const_matrix(sh.upts)
        self.ops = {op: be.const_matrix(sh.opmat(op), tags={op, 'align'})
                    for op in ('M0', 'M1 - M3*M2', 'M3', 'M4 - M6*M0', 'M6')}
        self.nbytes = sum(x.nbytes for x in self._all_mats())
        self.nchunks = E//SOA
        self.used = np.zeros(self.nchunks, dtype=bool)
        self.solves = {}            # rid -> [req, off, box, steps done]
        self.version = 0
        self._graphs = {}           # current bank -> (graph exec, next bank)
        self._gversion = -1
        self._maps = None
        # operator plans allocate device memory: create them before any
        # stream capture
        mul = be._providers[2]
        self._plans = {op: mul._plan(a, be.fpdtype)[0]
                       for op, a in self.ops.items()}
        if fused:
            # [M1 - M3*M2 | M3] and [M4 - M6*M0 | M6] as generated fused
            # kernels; the RK4 stage updates with per-element dt in the
            # divergence kernel's epilogue (ep 20-23), whose operands
            # {r1, r0, dt} are device pointer triples, one per bank parity
            self._plans2 = {}
            for k, (a0, a1) in (('div', ('M1 - M3*M2', 'M3')),
                                ('grad', ('M4 - M6*M0', 'M6'))):
                plan, kind = mul._plan2(self.ops[a0], self.ops[a1],
                                        be.fpdtype)
                if kind != 1:
                    raise ValueError(f'no generated fused kernel for hex '
                                     f'p={order}')
                self._plans2[k] = plan
            self._triples = {}
            for r1, r0 in ((1, 0), (0, 1)):
                h = np.array([self.banks[r1].data, self.banks[r0].data,
                              self.dt_e.data], dtype=np.uint64)
                d = be._malloc_impl(h.nbytes)
                self.lib.memcpy_htod(d, h.ctypes.data, h.nbytes, None)
                self._triples[r1, r0] = d
            self.lib.stream_sync(None)
        # interface arrays indexed by element slot: a periodic N^3 box has
        # 3 (p + 1)^2 flux-point pairs per element, so the pairs of a solve
        # at element offset off fill [off*ppe, (off + n)*ppe).  Pairs of
        # free slots are idle (both sides one free element's own point,
        # touching nothing a solve reads), so admission and retirement
        # write one solve's segment and never rebuild the rest
        self.ppe = 3*(order + 1)**2
        self.maxpairs = self.ecap*self.ppe
        isz = np.dtype(be.fpdtype).itemsize
        self._ibufs = {}
        for k, nb in (('l', 4*self.maxpairs), ('r', 4*self.maxpairs),
                      ('nl', 3*isz*self.maxpairs), ('mu', isz*self.maxpairs)):
            p = ct.c_void_p()
            self.lib.malloc(ct.byref(p), nb)
            self._ibufs[k] = p.value
        self.nbytes += (8 + 4*isz)*self.maxpairs
        s = ct.c_void_p()
        self.lib.stream_create(ct.byref(s))
        self.stream = s.value
        self._put_pairs(0, self.ecap, None, None)
        ev = ct.c_void_p()
        self.lib.event_create(ct.byref(ev))
        self.event = ev.value
        self.inflight = 0           # steps enqueued since the last idle
        self.cur = 0                # bank holding the current solution
        self.ncaptures = 0          # graph captures (membership changes)
        self.nsteps_run = 0
        self.host_s = 0.0           # host time spent enqueuing steps
        # NS constants (viscosity per element / pair overrides mu)
        self._consts = (ct.c_double*8)(GAMMA, 0.0, PR, 0.0, 0.0, 0.0, 1.0,
                                       0.0)

    def _all_mats(self):
        out = [*self.banks, self.scal_fpts, self.vect_upts, self.rcpdjac,
               self.dt_e, self.mu_e]
        out += [self.verts] if self.linear else [self.smats]
        if self.ns:
            out += [self.vect_fpts, self.grad_upts]
        return out

    @staticmethod
    def bytes_per_element(system, order):
        Nu, Nf = (order + 1)**3, 6*(order + 1)**2
        n = (3*Nu*5 + Nf*5 + 3*Nu*5 + Nu + 2)*8
        n += 24*8 if order >= 2 else 9*Nu*8
        if system == 'navier-stokes':
            n += (3*Nf*5 + 3*Nu*5)*8
        return n

    # -- slot allocation -----------------------------------------------------
    def find_slot(self, neles):
        need = -(-neles//SOA)
        run = 0
        for c in range(self.nchunks):
            run = run + 1 if not self.used[c] else 0
            if run == need:
                return (c - need + 1)*SOA
        return None

    @property
    def ehi(self):
        idx = np.flatnonzero(self.used)
        return 0 if len(idx) == 0 else (idx[-1] + 1)*SOA

    # -- host <-> device for a column range ---------------------------------
    def _put(self, mat, off, host):
        """Write host (ioshape with neles = host's last dim) into the element
        range [off, off + n) of mat (rows copied as pitched 2-D blocks)."""
        n = host.shape[-1]
        npad = n - n % -SOA
        isz = mat.itemsize
        if len(mat.ioshape) == 2:
            buf = np.zeros((mat.nrow, npad), dtype=mat.dtype)
            buf[:, :n] = host
            w, col0 = npad, off
        else:
            nv = mat.ioshape[-2]
            h = np.zeros(host.shape[:-1] + (npad,), dtype=mat.dtype)
            h[..., :n] = host
            h = h.reshape(h.shape[:-1] + (-1, SOA)).swapaxes(-2, -3)
            buf = np.ascontiguousarray(h.reshape(mat.nrow, -1))
            w, col0 = npad*nv, off*nv
        self.lib.memcpy2d(mat.data + col0*isz, mat.leaddim*isz,
                          buf.ctypes.data, w*isz, w*isz, mat.nrow, 0,
                          self.stream)

    def _get(self, mat, off, n):
        npad = n - n % -SOA
        nv = mat.ioshape[-2]
        isz = mat.itemsize
        w = npad*nv
        buf = np.empty((mat.nrow, w), dtype=mat.dtype)
        self.lib.memcpy2d(buf.ctypes.data, w*isz, mat.data + off*nv*isz,
                          mat.leaddim*isz, w*isz, mat.nrow, 1, self.stream)
        a = buf.reshape(mat.ioshape[:-2] + (-1, nv, SOA)).swapaxes(-2, -3)
        return a.reshape(mat.ioshape[:-1] + (-1,))[..., :n].copy()

    # -- admission / retirement ---------------------------------------------
    def admit(self, req, box, off):
        n = box.neles
        self.used[off//SOA:(off + n - n % -SOA)//SOA] = True
        cur = self.banks[self.cur]
        self._put(cur, off, box.ics(req.mach))
        if self.linear:
            self._put(self.verts, off, box.verts)
        else:
            self._put(self.smats, off, box.smats)
        self._put(self.rcpdjac, off, box.rcpdjac)
        self._put(self.dt_e, off, np.full((1, n), req.dt))
        self._put(self.mu_e, off, np.full((1, n), req.mu))
        self._put_pairs(off, n, box, req.mu)
        self.solves[req.rid] = [req, off, box, 0]
        self.version += 1

    def retire(self, rid):
        req, off, box, _ = self.solves.pop(rid)
        req.result = self._get(self.banks[self.cur], off, box.neles)
        n = box.neles
        self.used[off//SOA:(off + n - n % -SOA)//SOA] = False
        self._put_pairs(off, n, None, None)
        self.version += 1
        return req

    # -- interface maps, one segment per element slot -----------------------
    def _put_pairs(self, off, n, box, mu):
        """Write the pair segment of elements [off, off + n): the solve's
        pairs (box, viscosity mu) or, with box None, idle pairs."""
        ld = self.scal_fpts.leaddim
        k0, k1 = off*self.ppe, (off + n)*self.ppe
        if box is None:
            k = np.arange(k0, k1)
            e, f = k//self.ppe, k % self.ppe
            lm = rm = f*ld + (e//SOA)*SOA*5 + e % SOA
            nl = np.zeros((3, k1 - k0))
            nl[0] = 1.0
            mup = np.zeros(k1 - k0)
        else:
            if box.npairs != k1 - k0:
                raise ValueError('a periodic box has 3 (p+1)^2 pairs per '
                                 'element')
            el, er = box.le + off, box.re + off
            lm = box.lf*ld + (el//SOA)*SOA*5 + el % SOA
            rm = box.rf*ld + (er//SOA)*SOA*5 + er % SOA
            nl, mup = box.nl, np.full(box.npairs, mu)
        if max(lm.max(initial=0), rm.max(initial=0)) >= 2**31:
            raise ValueError('pool too large for int32 interface maps')
        fdt = self.be.fpdtype
        isz = np.dtype(fdt).itemsize
        b, st = self._ibufs, self.stream
        up = self.lib.memcpy_htod
        lm32, rm32 = lm.astype(np.int32), rm.astype(np.int32)
        up(b['l'] + 4*k0, lm32.ctypes.data, lm32.nbytes, st)
        up(b['r'] + 4*k0, rm32.ctypes.data, rm32.nbytes, st)
        mup = np.ascontiguousarray(mup, dtype=fdt)
        up(b['mu'] + isz*k0, mup.ctypes.data, mup.nbytes, st)
        nlc = np.ascontiguousarray(nl, dtype=fdt)
        self.lib.memcpy2d(b['nl'] + isz*k0, isz*self.maxpairs,
                          nlc.ctypes.data, nlc.shape[1]*isz,
                          nlc.shape[1]*isz, 3, 0, st)

    def _build_maps(self):
        """Pairs of the occupied extent [0, ehi) (solves and idle slots)."""
        b = self._ibufs
        self._maps = {'n': self.ehi*self.ppe, 'l': b['l'], 'r': b['r'],
                      'nl': b['nl'], 'ldn': self.maxpairs, 'mu': b['mu']}

    # -- one RK4 step over the pool -----------------------------------------
    def _rhs(self, uin, fout, s):
        lib, be = self.lib, self.be
        E = self.ehi
        V = 5
        ncol = V*E
        d = _dtcode(be.fpdtype)
        u, f = self.banks[uin], self.banks[fout]
        mp = self._maps

        def mm(op, b, c, beta=0.0):
            lib.mul(self._plans[op], ncol, b.data, b.leaddim, c.data,
                    c.leaddim, 1.0, beta, s)

        Nu, Nf = self.nupts, self.nfpts
        geo = self.verts if self.linear else self.smats
        mm('M0', u, self.scal_fpts)
        if not self.ns:
            lib.intcflux_euler(d, 0, mp['n'], self.scal_fpts.data,
                               mp['l'], mp['r'], mp['nl'],
                               mp['ldn'], GAMMA, -1.0, None, s)
            lib.tflux_euler(d, int(self.linear), Nu, E, u.data, u.leaddim,
                            self.vect_upts.data, self.vect_upts.leaddim,
                            geo.data, geo.leaddim, self.upts.data, GAMMA,
                            None, s)
        else:
            consts = self._consts
            lib.intconu(d, mp['n'], self.scal_fpts.data, mp['l'],
                        mp['r'], self.comm.data, mp['l'],
                        mp['r'], 0.0, s)
            mm('M4 - M6*M0', u, self.grad_upts)
            mm('M6', self.comm, self.grad_upts, 1.0)
            lib.tflux_ns(d, int(self.linear), Nu, E, u.data, u.leaddim,
                         self.vect_upts.data, self.vect_upts.leaddim,
                         self.grad_upts.data, self.grad_upts.leaddim,
                         geo.data, geo.leaddim, self.upts.data,
                         self.rcpdjac.data, self.rcpdjac.leaddim, consts,
                         self.mu_e.data, None, s)
            for i in range(3):
                lib.mul(self._plans['M0'], ncol,
                        self.grad_upts.data + i*Nu*u.leaddim*u.itemsize,
                        u.leaddim,
                        self.vect_fpts.data + i*Nf*u.leaddim*u.itemsize,
                        u.leaddim, 1.0, 0.0, s)
            lib.intcflux_ns(d, 0, mp['n'], self.scal_fpts.data,
                            mp['l'], mp['r'], self.vect_fpts.data,
                            mp['l'], mp['r'], Nf*u.leaddim, None, None,
                            mp['nl'], mp['ldn'], consts, -1.0,
                            mp['mu'], None, s)
        mm('M1 - M3*M2', self.vect_upts, f)
        mm('M3', self.scal_fpts, f, 1.0)
        lib.negdivconf(d, Nu, E, f.data, f.leaddim, self.rcpdjac.data,
                       self.rcpdjac.leaddim, s)

    def _rhs_stage(self, uin, stage, r1, r0, s):
        """The fused RHS of RK4 stage `stage` (1..4) from bank uin with the
        stage's updates in the divergence epilogue: r1 the accumulator, r0
        the step's initial state, bank 2 the next stage's input (in place
        for stages 2-4: the divergence kernel reads only the flux and the
        common flux, both complete before it writes)."""
        lib, be = self.lib, self.be
        E = self.ehi
        ncol = 5*E
        d = _dtcode(be.fpdtype)
        u, nxt = self.banks[uin], self.banks[2]
        mp = self._maps
        Nu, Nf = self.nupts, self.nfpts
        geo = self.verts if self.linear else self.smats
        sf, vu = self.scal_fpts, self.vect_upts
        lib.mul(self._plans['M0'], ncol, u.data, u.leaddim, sf.data,
                sf.leaddim, 1.0, 0.0, s)
        if not self.ns:
            lib.intcflux_euler(d, 0, mp['n'], sf.data, mp['l'], mp['r'],
                               mp['nl'], mp['ldn'], GAMMA, -1.0, None, s)
            lib.tflux_euler(d, int(self.linear), Nu, E, u.data, u.leaddim,
                            vu.data, vu.leaddim, geo.data, geo.leaddim,
                            self.upts.data, GAMMA, None, s)
        else:
            consts = self._consts
            gu, vf = self.grad_upts, self.vect_fpts
            lib.intconu(d, mp['n'], sf.data, mp['l'], mp['r'],
                        self.comm.data, mp['l'], mp['r'], 0.0, s)
            lib.mul2(self._plans2['grad'], ncol, u.data, u.leaddim,
                     self.comm.data, self.comm.leaddim, gu.data, gu.leaddim,
                     0, None, 0, None, 0.0, 0.0, 0, s)
            lib.tflux_ns(d, int(self.linear), Nu, E, u.data, u.leaddim,
                         vu.data, vu.leaddim, gu.data, gu.leaddim, geo.data,
                         geo.leaddim, self.upts.data, self.rcpdjac.data,
                         self.rcpdjac.leaddim, consts, self.mu_e.data, None,
                         s)
            for i in range(3):
                lib.mul(self._plans['M0'], ncol,
                        gu.data + i*Nu*gu.leaddim*gu.itemsize, gu.leaddim,
                        vf.data + i*Nf*vf.leaddim*vf.itemsize, vf.leaddim,
                        1.0, 0.0, s)
            lib.intcflux_ns(d, 0, mp['n'], sf.data, mp['l'], mp['r'],
                            vf.data, mp['l'], mp['r'], Nf*vf.leaddim, None,
                            None, mp['nl'], mp['ldn'], consts, -1.0,
                            mp['mu'], None, s)
        lib.mul2(self._plans2['div'], ncol, vu.data, vu.leaddim, sf.data,
                 sf.leaddim, nxt.data, nxt.leaddim, 19 + stage,
                 self.rcpdjac.data, self.rcpdjac.leaddim,
                 self._triples[r1, r0], 0.0, 0.0, 0, s)

    def _add(self, s, a, mask, *ri):
        lib, be = self.lib, self.be
        m = self.banks[0]
        E = self.ehi
        xs = (ct.c_void_p*len(ri))(*[self.banks[r].data for r in ri])
        aa = (ct.c_double*len(ri))(*a)
        lib.axnpby_pool(_dtcode(be.fpdtype), m.nrow, 5*E, m.leaddim, len(ri),
                        xs, aa, mask, self.dt_e.data, 5, s)

    def _step_kernels(self, s):
        """RK4 (``steppers.py:67-108``) with per-solve dt; coefficients
        dt/k are passed as divisors k."""
        r0, r1, r2 = self.cur, 1 - self.cur if self.cur < 2 else 0, 2
        if self.cur == 2:
            raise RuntimeError('invalid register state')
        if self.fused:
            for stage in (1, 2, 3, 4):
                self._rhs_stage(r0 if stage == 1 else r2, stage, r1, r0, s)
            return r1
        add = lambda a, m, *r: self._add(s, a, m, *r)
        self._rhs(r0, r1, s)
        add((0.0, 1.0, 2.0), 0b100, r2, r0, r1)
        self._rhs(r2, r2, s)
        add((6.0, 1.0, 3.0), 0b101, r1, r0, r2)
        add((2.0, 1.0), 0b01, r2, r0)
        self._rhs(r2, r2, s)
        add((1.0, 3.0), 0b10, r1, r2)
        add((1.0, 1.0), 0b01, r2, r0)
        self._rhs(r2, r2, s)
        add((1.0, 6.0), 0b10, r1, r2)
        return r1

    def step(self, use_graph=True):
        """Advance every solve of the pool by one RK4 step (asynchronous on
        the pool's stream)."""
        if not self.solves:
            return
        t0 = time.perf_counter()
        if self._gversion != self.version:
            # the kernels' arguments depend on the occupied extent only
            self._build_maps()
            self._gversion = self.version
        if use_graph:
            key = (self.cur, self.ehi)
            if key not in self._graphs:
                if len(self._graphs) >= 8:
                    g0, _ = self._graphs.pop(next(iter(self._graphs)))
                    self.lib.graph_destroy(g0)
                self.lib.graph_begin(self.stream)
                try:
                    nxt = self._step_kernels(self.stream)
                finally:
                    ge = ct.c_void_p()
                    self.lib.graph_end(self.stream, ct.byref(ge))
                self._graphs[key] = (ge.value, nxt)
                self.ncaptures += 1
            g, nxt = self._graphs[key]
            self.lib.graph_launch(g, self.stream)
            self.cur = nxt
        else:
            self.cur = self._step_kernels(self.stream)
        for v in self.solves.values():
            v[3] += 1
        self.nsteps_run += 1
        self.host_s += time.perf_counter() - t0

    def sync(self):
        self.lib.stream_sync(self.stream)
        self.inflight = 0

    def enqueue(self, nsteps, use_graph=True):
        """Enqueue nsteps RK4 steps and an event; returns immediately."""
        for _ in range(nsteps):
            self.step(use_graph)
        self.lib.event_record(self.event, self.stream)
        self.inflight = nsteps

    def idle(self):
        if not self.inflight:
            return True
        done = ct.c_int()
        self.lib.event_query(self.event, ct.byref(done))
        if done.value:
            self.inflight = 0
        return bool(done.value)

    def due(self):
        """rids of solves that have taken all their steps."""
        return [rid for rid, v in self.solves.items() if v[3] >= v[0].nsteps]

    def steps_to_next_finish(self):
        return min(v[0].nsteps - v[3] for v in self.solves.values())

    def _drop_graphs(self):
        for g, _ in self._graphs.values():
            self.lib.graph_destroy(g)
        self._graphs = {}

    def __del__(self):
        try:
            self._drop_graphs()
            for p in self._ibufs.values():
                self.lib.free(p)
            self.lib.event_destroy(self.event)
            self.lib.stream_destroy(self.stream)
        except Exception:
            pass


# ---------------------------------------------------------------------------
# Engine
# ---------------------------------------------------------------------------

class ServingEngine:
    """Admits, batches and steps solves; see the module docstring."""

    def __init__(self, backend, budget_bytes=48*2**30, max_neles=24**3,
                 classes=None, slots_per_class=6, fused=True):
        self.be = backend
        self.budget = budget_bytes
        self.pools = {}
        self.boxes = {}
        classes = classes or [(s, p) for s in ('euler', 'navier-stokes')
                              for p in (1, 2, 3, 4)]
        # Size each class for slots_per_class solves of the largest mesh,
        # scaled down uniformly if the budget does not allow it
        need = {c: Pool.bytes_per_element(*c)*max_neles*slots_per_class
                for c in classes}
        scale = min(1.0, budget_bytes/sum(need.values()))
        for c in classes:
            ecap = int(max_neles*slots_per_class*scale)
            ecap = max(ecap - ecap % SOA, max_neles - max_neles % -SOA)
            self.pools[c] = Pool(backend, c[0], c[1], ecap, fused)
        self.hbm_bytes = sum(p.nbytes for p in self.pools.values())

    def box(self, N, order):
        key = (N, order)
        if key not in self.boxes:
            self.boxes[key] = _BoxData(N, order)
        return self.boxes[key]

    def try_admit(self, req, now):
        pool = self.pools[(req.system, req.order)]
        off = pool.find_slot(req.neles)
        if off is None:
            return False
        pool.admit(req, self.box(req.N, req.order), off)
        req.admit_time = now
        return True

    def run(self, requests, mode='batched', use_graph=True, clock=None,
            on_step=None, chunk=4):
        """Serve the requests (arrival times relative to the start).

        mode 'batched': continuous batching -- each pool runs independently
        on its own stream, enqueuing up to ``chunk`` steps at a time (fewer
        if a solve finishes sooner); whenever a pool is idle its finished
        solves are retired and queued requests of its class admitted.
        mode 'sequential': one solve at a time in arrival order (the
        baseline).  Returns the finished requests with admit/finish times."""
        clock = clock or time.perf_counter
        for b in {(r.N, r.order) for r in requests}:
            self.box(*b)
        pending = sorted(requests, key=lambda r: r.arrival)
        queue, done = [], []
        t0 = clock()
        qi = 0
        pools = list(self.pools.values())
        while len(done) < len(requests):
            now = clock() - t0
            while qi < len(pending) and pending[qi].arrival <= now:
                queue.append(pending[qi])
                qi += 1
            progressed = False
            for p in pools:
                if not p.idle():
                    continue
                # retire finished solves of this pool
                for rid in p.due():
                    req = p.retire(rid)
                    req.finish_time = clock() - t0
                    done.append(req)
                    progressed = True
                # admission
                if mode == 'batched':
                    rest = []
                    for r in queue:
                        if (r.system, r.order) == (p.system, p.order) and \
                           self.try_admit(r, clock() - t0):
                            progressed = True
                        else:
                            rest.append(r)
                    queue = rest
                elif queue and not any(q.solves for q in pools):
                    head = queue[0]
                    if (head.system, head.order) == (p.system, p.order):
                        if not self.try_admit(head, clock() - t0):
                            raise ValueError(f'request {head.rid} does not '
                                             'fit an empty pool')
                        queue.pop(0)
                        progressed = True
                if p.solves:
                    n = max(1, min(chunk, p.steps_to_next_finish()))
                    p.enqueue(n, use_graph)
                    progressed = True
                    if on_step:
                        p.sync()
                        on_step(self)
            if not progressed:
                if qi < len(pending) and not any(q.solves for q in pools):
                    wait = pending[qi].arrival - (clock() - t0)
                    if wait > 0:
                        time.sleep(min(wait, 0.005))
                else:
                    time.sleep(2e-5)
        return done


def summarize(done):
    lat = np.array([r.finish_time - r.arrival for r in done])
    t_end = max(r.finish_time for r in done)
    t_start = min(r.arrival for r in done)
    work = sum(r.ndofs*4*r.nsteps for r in done)
    return {'n': len(done),
            'p50_s': float(np.percentile(lat, 50)),
            'p99_s': float(np.percentile(lat, 99)),
            'mean_s': float(lat.mean()),
            'makespan_s': float(t_end - t_start),
            'gdof_rhs_per_s': work/(t_end - t_start)/1e9}


def sequential_bound(done):
    """From a one-solve-at-a-time run: each solve's alone time (admission
    to result) and rate; per (system, order) class the best alone rate (its
    largest solves, near the class's saturated rate); the work share of the
    solves that are launch-bound alone (rate below half their class's best);
    and the ratio batched / one-at-a-time that running every solve at its
    class's best rate would give (the bound batching can reach on this
    stream, short of overlapping classes)."""
    cls = {}
    rows = []
    for r in done:
        t = r.finish_time - r.admit_time
        w = r.ndofs*4*r.nsteps
        rows.append(((r.system, r.order), w, t))
        k = (r.system, r.order)
        cls[k] = max(cls.get(k, 0.0), w/t)
    t_seq = sum(t for _, _, t in rows)
    t_ideal = sum(w/cls[k] for k, w, _ in rows)
    work = sum(w for _, w, _ in rows)
    lb = sum(w for k, w, t in rows if w/t < 0.5*cls[k])
    return {'ratio_max': t_seq/t_ideal,
            'launch_bound_work_share': lb/work,
            'class_best_gdof_rhs_per_s': {f'{k[0]}-p{k[1]}': v/1e9
                                          for k, v in sorted(cls.items())}}
