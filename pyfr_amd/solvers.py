"""Solver-side host code: elements, interfaces, systems and explicit
integrators for the Euler and Navier-Stokes FR right-hand side.

This is the build's counterpart of the reference's solver layer, written
against the backend API of ``pyfr_amd.backend`` the way the reference's is
written against ``pyfr/backends/base``:

* elements     -- ``pyfr/solvers/base/elements.py``,
  ``pyfr/solvers/baseadvec/elements.py:55-127``,
  ``pyfr/solvers/baseadvecdiff/elements.py:22-131``,
  ``pyfr/solvers/euler/elements.py:149-206``,
  ``pyfr/solvers/navstokes/elements.py:31-128``;
* interfaces   -- ``pyfr/solvers/base/inters.py``,
  ``pyfr/solvers/baseadvec/inters.py:8-116``,
  ``pyfr/solvers/baseadvecdiff/inters.py:7-76``,
  ``pyfr/solvers/euler/inters.py:28-102``,
  ``pyfr/solvers/navstokes/inters.py:29-219`` (internal and boundary);
* systems      -- ``pyfr/solvers/base/system.py``,
  ``pyfr/solvers/baseadvec/system.py:57-136``,
  ``pyfr/solvers/baseadvecdiff/system.py:53-231``;
* integrators  -- ``pyfr/integrators/explicit/steppers.py:60-243``,
  ``pyfr/integrators/base.py:282-296``, ``:401-426``,
  ``pyfr/integrators/explicit/controllers.py:42-60``.

Scope: meshes of hexahedra, tetrahedra, prisms and pyramids (straight or
curved) on one GPU, periodic or with the reference's boundary conditions (values may
depend on x, y, z and t; the controlled char-riem-inv types included);
anti-aliasing (flux, surf-flux) on every element type; no MPI interfaces,
no source terms.
"""
from functools import cached_property
import math
import re
import time

import numpy as np

from pyfr_amd.geometry import (AffineGeometry, CurvedSimplexGeometry,
                               HexGeometry, hex_corner_index)
from pyfr_amd.mesh import sorted_face_fpts, sorted_face_fpts_structured
from pyfr_amd.shapes import SHAPES as HEX_SHAPES
from pyfr_amd.simplex import SIMPLEX_SHAPES

SHAPES = {**HEX_SHAPES, **SIMPLEX_SHAPES}


def _npeval(expr, env):
    """Evaluate an initial-condition expression (``pyfr/nputil.py:103-117``)."""
    syms = {'__builtins__': {}, 'exp': np.exp, 'log': np.log, 'sin': np.sin,
            'cos': np.cos, 'tan': np.tan, 'tanh': np.tanh, 'sqrt': np.sqrt,
            'abs': np.abs, 'pow': np.power, 'pi': np.pi, 'atan': np.arctan,
            'atan2': np.arctan2, 'asin': np.arcsin, 'acos': np.arccos,
            'max': np.maximum, 'min': np.minimum}
    if '**' in expr or '^' in expr:
        raise ValueError('Direct exponentiation is not supported; use pow')
    return eval(expr, syms, env)


class Interside:
    """One side of a set of interfaces: (etype, eidx, fidx) per interface,
    grouped like ``Connectivity.foreach`` (``pyfr/readers/native.py:76-80``)."""

    def __init__(self, side):
        self.side = side

    def __len__(self):
        return len(self.side)

    def foreach(self):
        groups = {}
        for idx, (et, e, f) in enumerate(self.side):
            groups.setdefault((et, f), ([], []))
            groups[(et, f)][0].append(e)
            groups[(et, f)][1].append(idx)
        for (et, f), (eidxs, idxs) in sorted(groups.items()):
            yield et, f, np.array(eidxs), np.array(idxs)

    def items(self):
        for et, f, eidxs, _ in self.foreach():
            yield et, f, eidxs


# ---------------------------------------------------------------------------
# Elements
# ---------------------------------------------------------------------------

class FluidElements:
    """Base elements with the fluid variable sets
    (``pyfr/solvers/euler/elements.py:8-114``)."""
    convars = ['rho', 'rhou', 'rhov', 'rhow', 'E']
    privars = ['rho', 'u', 'v', 'w', 'p']

    def __init__(self, basiscls, eles, cfg):
        self._be = None
        self.eles = eles
        self.cfg = cfg
        self.nspts, self.neles, self.ndims = eles.shape
        self.nvars = len(self.convars)
        self.kernels = {}
        order = cfg.getint('solver', 'order')
        # anti-aliasing options (pyfr/shapes.py:40-44)
        aa = cfg.get('solver', 'anti-alias', 'none')
        self.antialias = {a.strip() for a in aa.split(',')} - {'none'}
        if self.antialias - {'flux', 'surf-flux'}:
            raise ValueError('Invalid anti-alias options')
        self.fluxaa = 'flux' in self.antialias
        # solution and face point families (pyfr/shapes.py:188-190,
        # :205-218): any tabulated set for tets, prisms and pyramids and
        # their triangular faces; Gauss-Legendre or -Lobatto hex solution
        # points; Gauss-Legendre quad faces
        qfp = cfg.get('solver-interfaces-quad', 'flux-pts', 'gauss-legendre')
        if qfp != 'gauss-legendre':
            raise NotImplementedError(f'quad flux-pts {qfp} is not built')
        hsp = cfg.get('solver-elements-hex', 'soln-pts', None)
        if basiscls is not SHAPES['hex']:
            qrule = (self._simplex_qrule(cfg, basiscls.name, order)
                     if self.fluxaa else None)
            iq = (self._simplex_iqrules(cfg, order)
                  if 'surf-flux' in self.antialias else None)
            self.basis = basis = basiscls(
                order, qrule, iq,
                cfg.get(f'solver-elements-{basiscls.name}', 'soln-pts',
                        None),
                cfg.get('solver-interfaces-tri', 'flux-pts', None))
        elif self.antialias:
            nq1 = (self._gl_n1(cfg, 'solver-elements-hex', 3, order)
                   if self.fluxaa else None)
            nf1 = (self._gl_n1(cfg, 'solver-interfaces-quad', 2, order)
                   if 'surf-flux' in self.antialias else None)
            self.basis = basis = basiscls(order, nq1, nf1, hsp)
        else:
            self.basis = basis = basiscls(order, upts_name=hsp)
        self.name = basis.name
        self.nupts, self.nfpts = basis.nupts, basis.nfpts
        self.nfacefpts = basis.nfacefpts
        # straight-sided tets, right prisms and pyramids with parallelogram
        # bases (vertices only) have one S and |J| per element; others
        # (order-k shape points, or prisms whose map is not affine) take the
        # stored metric at every point, all elements of the type in the
        # curved region
        self.affine = False
        self.curved_simplex = False
        if self.name in SIMPLEX_SHAPES:
            try:
                if self.nspts != basiscls.nverts:
                    raise ValueError('curved')
                self.geom = AffineGeometry(eles, self.name)
                self.affine = True
            except ValueError:
                self.geom = CurvedSimplexGeometry(eles, order, self.name)
                self.curved_simplex = True
        else:
            self.geom = HexGeometry(eles, order)

    @staticmethod
    def _simplex_qrule(cfg, name, order):
        """(rule, npts) of a tet / prism element quadrature rule
        (``pyfr/shapes.py:146-185``): by default the rule the reference's
        selection picks for order 1..4 (``simplex.AA_RULES``), or
        [solver-elements-<name>] quad-pts / quad-npts naming one of the
        tabulated rules."""
        from pyfr_amd.quadtables import TABLES
        from pyfr_amd.simplex import AA_RULES, qdeg_rule
        sect = f'solver-elements-{name}'
        if cfg.hasopt(sect, 'quad-deg'):
            return qdeg_rule(name, cfg.getint(sect, 'quad-deg'),
                             cfg.get(sect, 'quad-pts', None),
                             cfg.getint(sect, 'quad-npts', None))
        if cfg.hasopt(sect, 'quad-pts') or cfg.hasopt(sect, 'quad-npts'):
            key = (name, cfg.get(sect, 'quad-pts'),
                   cfg.getint(sect, 'quad-npts'))
            if key not in TABLES or TABLES[key][1][0] is None:
                raise NotImplementedError(f'quadrature rule {key[1:]} is '
                                          f'not tabulated for {name}')
            return key[1:]
        if order not in AA_RULES[name]:
            raise NotImplementedError(f'no default {name} quadrature rule '
                                      f'at order {order}')
        return AA_RULES[name][order]

    @classmethod
    def _simplex_iqrules(cls, cfg, order):
        """Face rules of surface-flux anti-aliasing on tets and prisms
        (``pyfr/shapes.py:159-185``): triangles by default the rule the
        reference's selection picks (``simplex.SF_TRI_RULES``), or
        [solver-interfaces-tri] quad-pts / quad-npts naming a tabulated
        rule; quads Gauss-Legendre as for hexes."""
        from pyfr_amd.quadtables import TABLES
        from pyfr_amd.simplex import SF_TRI_RULES, qdeg_rule
        sect = 'solver-interfaces-tri'
        if cfg.hasopt(sect, 'quad-deg'):
            tri = qdeg_rule('tri', cfg.getint(sect, 'quad-deg'),
                            cfg.get(sect, 'quad-pts', None),
                            cfg.getint(sect, 'quad-npts', None))
        elif cfg.hasopt(sect, 'quad-pts') or cfg.hasopt(sect, 'quad-npts'):
            key = ('tri', cfg.get(sect, 'quad-pts'),
                   cfg.getint(sect, 'quad-npts'))
            if key not in TABLES or TABLES[key][1][0] is None:
                raise NotImplementedError(f'quadrature rule {key[1:]} is '
                                          'not tabulated for tri')
            tri = key[1:]
        elif order in SF_TRI_RULES:
            tri = SF_TRI_RULES[order]
        else:
            raise NotImplementedError(f'no default tri quadrature rule at '
                                      f'order {order}')
        return {'tri': tri,
                'quad': cls._gl_n1(cfg, 'solver-interfaces-quad', 2, order)}

    @staticmethod
    def _gl_n1(cfg, sect, ndims, order):
        """Points per direction of a Gauss-Legendre anti-aliasing rule
        (``pyfr/shapes.py:147-185``): the element rule of flux
        anti-aliasing ([solver-elements-hex], ndims 3) or the face rule of
        surface-flux anti-aliasing ([solver-interfaces-quad], ndims 2); by
        default (order + 2)^ndims points, or quad-npts, or the rule
        quad-deg selects (which must be a Gauss-Legendre one)."""
        if cfg.hasopt(sect, 'quad-deg'):
            from pyfr_amd.simplex import qdeg_rule
            r, npts = qdeg_rule('hex' if ndims == 3 else 'quad',
                                cfg.getint(sect, 'quad-deg'),
                                cfg.get(sect, 'quad-pts', None),
                                cfg.getint(sect, 'quad-npts', None))
            if r != 'gauss-legendre':
                raise NotImplementedError(
                    f'quad-deg selects the {r} rule of {npts} points; '
                    'only Gauss-Legendre rules are supported here (give '
                    'quad-pts = gauss-legendre)')
            n1 = round(npts ** (1/ndims))
            if n1 < order + 1:
                raise ValueError(f'Invalid quad-deg in [{sect}]')
            return n1
        if cfg.get(sect, 'quad-pts', 'gauss-legendre') != 'gauss-legendre':
            raise NotImplementedError('only Gauss-Legendre quadrature '
                                      'points are supported')
        npts = cfg.getint(sect, 'quad-npts', (order + 2)**ndims)
        n1 = round(npts ** (1/ndims))
        if n1**ndims != npts or n1 < order + 1:
            raise ValueError(f'Invalid quad-npts in [{sect}]')
        return n1

    # -- initial conditions (pyfr/solvers/base/elements.py:86-129)
    @staticmethod
    def pri_to_con(pris, gamma):
        rho, p = pris[0], pris[-1]
        rhovs = [rho*c for c in pris[1:-1]]
        E = p/(gamma - 1) + 0.5*rho*sum(c*c for c in pris[1:-1])
        return [rho, *rhovs, E]

    def set_ics_from_cfg(self):
        """``pyfr/solvers/base/elements.py:86-129``: the [soln-ics]
        expressions at the solution points, or, with [soln-ics]
        quad-deg(-<etype>), at the points of the rule of that degree from
        the family quad-pts(-<etype>) (default: the solution points' family)
        and L2-projected onto the element's space, variable by variable in
        conservative form."""
        env = self.cfg.items_as('constants', float)
        qdeg = (self.cfg.getint('soln-ics', f'quad-deg-{self.name}', 0) or
                self.cfg.getint('soln-ics', 'quad-deg', 0))
        if qdeg:
            qpts, m8 = self._ics_projection(qdeg)
        else:
            qpts, m8 = self.basis.upts, None
        coords = self.geom.ploc_at(qpts)
        env |= dict(zip('xyz', coords.swapaxes(0, 1)))
        ics = [_npeval(self.cfg.getexpr('soln-ics', v), env)
               for v in self.privars]
        ics = [np.broadcast_to(v, (len(qpts), self.neles)) for v in ics]
        con = self.pri_to_con(ics, self.cfg.getfloat('constants', 'gamma'))
        out = np.empty((self.nupts, self.nvars, self.neles))
        for i, v in enumerate(con):
            out[:, i] = v if m8 is None else m8 @ v
        return out

    def _ics_projection(self, qdeg):
        """(points, L2 projection onto the solution points) of the
        quadrature rule of degree >= qdeg (``pyfr/shapes.py:18-19``)."""
        from pyfr_amd.shapes import gauss_legendre, l2_projection_1d
        from pyfr_amd.simplex import qdeg_rule
        if self.name == 'hex':
            name = self.cfg.get('soln-ics', 'quad-pts-hex',
                                self.cfg.get('soln-ics', 'quad-pts',
                                             'gauss-legendre'))
            r, npts = qdeg_rule('hex', qdeg, name)
            if r != 'gauss-legendre':
                raise NotImplementedError('IC projection on hexes takes '
                                          'Gauss-Legendre rules')
            n = round(npts ** (1/3))
            x = gauss_legendre(n)[0]
            P = l2_projection_1d(self.basis.x1d, n)
            pts = np.array([(x[i], x[j], x[k]) for k in range(n)
                            for j in range(n) for i in range(n)])
            return pts, np.kron(P, np.kron(P, P))
        from pyfr_amd.quadtables import TABLES
        name = self.cfg.get('soln-ics', f'quad-pts-{self.name}',
                            self.basis.upts_name)
        key = (self.name, *qdeg_rule(self.name, qdeg, name))
        qp, qw = (np.array(a, dtype=float) for a in TABLES[key])
        ob = self.basis.ubasis.ortho
        return qp, ob.at(self.basis.upts).T @ (qw*ob.at(qp))

    # -- geometry
    @cached_property
    def plocfpts(self):
        p = self.geom.ploc_at(self.basis.fpts)           # (nfpts, 3, E)
        return p.transpose(0, 2, 1)                        # (nfpts, E, 3)

    @cached_property
    def srtd_face_fpts(self):
        if self.name != 'hex':
            return sorted_face_fpts(self.plocfpts, self.basis.facefpts)
        return sorted_face_fpts_structured(self.plocfpts, self.basis.facefpts)

    @cached_property
    def _pnorm_fpts(self):
        return self.geom.pnorm_at(self.basis.fpts, self.basis.norm_fpts)

    def rcpdjac_at_np(self, name):
        return self.geom.rcpdjac_at(getattr(self.basis, name))

    # -- backend set-up (pyfr/solvers/base/elements.py:218-266)
    @property
    def _scratch_bufs(self):
        return {'scal_fpts', 'vect_upts'}

    def set_backend(self, backend, nonce, linoff):
        self.base_tpl = {}
        self._be = backend
        self.grad_fusion = not (backend.blocks or self.fluxaa)
        if self.basis.order >= 2 and not self.curved_simplex:
            self.linoff = linoff - linoff % -backend.csubsz
        else:
            self.linoff = self.neles

        nvars, neles, ndims = self.nvars, self.neles, self.ndims
        sbufs = self._scratch_bufs

        def alloc(ex, shape):
            return backend.matrix(shape, extent=nonce + ex, tags={'align'})

        if 'scal_fpts' in sbufs:
            self._scal_fpts = alloc('scal_fpts', (self.nfpts, nvars, neles))
        if 'vect_upts' in sbufs:
            self._vect_upts = alloc('vect_upts',
                                    (ndims, self.nupts, nvars, neles))
        if 'vect_fpts' in sbufs:
            self._vect_fpts = alloc('vect_fpts',
                                    (ndims, self.nfpts, nvars, neles))
            self._comm_fpts = None   # a slice once allocated (see below)
        if 'grad_upts' in sbufs:
            self._grad_upts = alloc('grad_upts',
                                    (ndims, self.nupts, nvars, neles))
        elif 'vect_upts' in sbufs:
            self._grad_upts = self._vect_upts
        if self.fluxaa:
            nq = self.basis.nqpts
            self._scal_qpts = alloc('scal_qpts', (nq, nvars, neles))
            self._vect_qpts = alloc('vect_qpts', (ndims, nq, nvars, neles))
        self.scal_upts = []

    def post_commit(self):
        if hasattr(self, '_vect_fpts'):
            self._comm_fpts = self._vect_fpts.slice(0, self.nfpts)

    def alloc_bank(self, extent, ic=None, dtype=None):
        m = self._be.matrix((self.nupts, self.nvars, self.neles), ic,
                            extent=extent, tags={'align'}, dtype=dtype)
        self.scal_upts.append(m)
        return m

    _opmats = None

    def opmat(self, expr):
        if self._opmats is None:
            self._opmats = {}
        if expr not in self._opmats:
            self._opmats[expr] = self._be.const_matrix(
                self.basis.opmat(expr), tags={expr, 'align'})
        return self._opmats[expr]

    @property
    def mesh_regions(self):
        if self.affine:
            return {'affine': self.neles}
        off = self.linoff
        if off == 0:
            return {'linear': self.neles}
        elif off >= self.neles:
            return {'curved': self.neles}
        return {'curved': off, 'linear': self.neles - off}

    def _slice_mat(self, mat, region):
        off = self.linoff
        if len(mat.ioshape) >= 3:
            off *= mat.ioshape[-2]
        else:
            off = min(off, mat.ncol)
        if region == 'curved':
            return mat.slice(0, mat.nrow, 0, off)
        elif region == 'linear':
            if off == 0:
                return mat
            return mat.slice(0, mat.nrow, off, mat.ncol)
        raise ValueError('Invalid slice region')

    # -- constant geometric matrices
    @cached_property
    def _verts_linear(self):
        # ploc_at('linspts') = the vertices (8, ndims, neles) of the linear
        # region (the corners of higher-order shape points)
        v = self.eles[hex_corner_index(self.geom.sorder), self.linoff:]
        return self._be.const_matrix(np.ascontiguousarray(
            v.transpose(0, 2, 1)), tags={'align'})

    @cached_property
    def _upts_const(self):
        return self._be.const_matrix(self.basis.upts)

    @cached_property
    def _qpts_const(self):
        return self._be.const_matrix(self.basis.qpts)

    @cached_property
    def _smats_curved_q(self):
        s = self.geom.smat_at(self.basis.qpts)[..., :self.linoff]
        return self._be.const_matrix(s, tags={'align'})

    # the divergence operator on the flux (at quadrature points under flux
    # anti-aliasing, pyfr/solvers/baseadvec/elements.py:84-94)
    @property
    def _divop(self):
        return '(M1 - M3*M2)*M9' if self.fluxaa else 'M1 - M3*M2'

    @property
    def _flux_buf(self):
        return self._vect_qpts if self.fluxaa else self._vect_upts

    def _qptsu(self, uin):
        """qptsu: M7 u at the quadrature points
        (``pyfr/solvers/baseadvec/elements.py:78-82``)."""
        return self._be.kernel('mul', self.opmat('M7'), self.scal_upts[uin],
                               out=self._scal_qpts)

    @cached_property
    def _smats_curved(self):
        s = self.geom.smat_at(self.basis.upts)[..., :self.linoff]
        return self._be.const_matrix(s, tags={'align'})

    @cached_property
    def _rcpdjac_upts(self):
        if self.affine:
            # one value per element (kernels read it with row stride 0)
            return self._be.const_matrix(1/self.geom.djac[None, :],
                                         tags={'align'})
        return self._be.const_matrix(self.rcpdjac_at_np('upts'),
                                     tags={'align'})

    @cached_property
    def _smats_affine(self):
        # (9, neles): row 3i + k holds S[i][k]
        s = self.geom.smats.reshape(self.neles, 9).T
        return self._be.const_matrix(np.ascontiguousarray(s), tags={'align'})

    def _rcpdjac_curved(self):
        r = self.rcpdjac_at_np('upts')[:, :self.linoff]
        return self._be.const_matrix(r, tags={'align'})

    # -- interface helpers (pyfr/solvers/base/elements.py:446-469)
    def get_pnorms_for_inters(self, eidxs, fidx):
        fpts_idx = self.srtd_face_fpts[fidx][eidxs]
        pn = self._pnorm_fpts[fpts_idx, eidxs[:, None]]
        return pn.reshape(-1, self.ndims),

    def get_ploc_for_inters(self, eidxs, fidx):
        fpts_idx = self.srtd_face_fpts[fidx][eidxs]
        return self.plocfpts[fpts_idx, eidxs[:, None]].reshape(-1, self.ndims),

    def _inters_map(self, mid, rmap, eidxs, fidx, lda=None):
        nfp = self.nfacefpts[fidx]
        n = len(eidxs)
        cmap = np.repeat(eidxs, nfp)
        out = [np.full(n*nfp, mid), rmap.ravel(), cmap]
        if lda is not None:
            out.append(np.full(n*nfp, lda))
        return out

    def get_vect_fpts_for_inters(self, eidxs, fidx):
        rmap = self.srtd_face_fpts[fidx][eidxs]
        return self._inters_map(self._vect_fpts.mid, rmap, eidxs, fidx,
                                self.nfpts)

    def get_comm_fpts_for_inters(self, eidxs, fidx):
        # fpts not in upts: the common solution lives in the first nfpts rows
        # of vect_fpts (pyfr/solvers/base/elements.py:70-75, :254-259)
        rmap = self.srtd_face_fpts[fidx][eidxs]
        return self._inters_map(self._vect_fpts.mid, rmap, eidxs, fidx,
                                self.nfpts)


class EulerElements(FluidElements):
    def set_backend(self, backend, nonce, linoff):
        super().set_backend(backend, nonce, linoff)
        be, kernels = backend, self.kernels
        be.pointwise.register('pyfr.solvers.baseadvec.kernels.negdivconf')

        # pyfr/solvers/baseadvec/elements.py:73-119
        kernels['disu'] = lambda uin: be.kernel(
            'mul', self.opmat('M0'), self.scal_upts[uin], out=self._scal_fpts)
        kernels['tdivtpcorf'] = lambda fout: be.kernel(
            'mul', self.opmat(self._divop), self._flux_buf,
            out=self.scal_upts[fout])
        kernels['tdivtconf'] = lambda fout: be.kernel(
            'mul', self.opmat('M3'), self._scal_fpts,
            out=self.scal_upts[fout], beta=float(self.basis.order > 0))
        kernels['negdivconf'] = lambda fout: be.kernel(
            'negdivconf', tplargs={'ndims': self.ndims, 'nvars': self.nvars,
                                   'src_macros': []},
            dims=[self.nupts, self.neles], tdivtconf=self.scal_upts[fout],
            rcpdjac=self._rcpdjac_upts)

        # pyfr/solvers/euler/elements.py:149-206
        be.pointwise.register('pyfr.solvers.euler.kernels.tflux')
        tplargs = {'ndims': self.ndims, 'nvars': self.nvars,
                   'nverts': 8 if self.name == 'hex' else self.nspts,
                   'c': self.cfg.items_as('constants', float)}
        r, s = self.mesh_regions, self._slice_mat

        def tdisf(uin):
            if self.fluxaa:
                return self._tdisf_qpts(uin, tplargs)
            ks = []
            if 'affine' in r:
                ks.append(be.kernel(
                    'tflux', tplargs=tplargs | self.base_tpl | {'ktype': 'affine'},
                    dims=[self.nupts, self.neles], u=self.scal_upts[uin],
                    f=self._vect_upts, smats=self._smats_affine))
            if 'curved' in r:
                ks.append(be.kernel(
                    'tflux', tplargs=tplargs | self.base_tpl | {'ktype': 'curved'},
                    dims=[self.nupts, r['curved']],
                    u=s(self.scal_upts[uin], 'curved'),
                    f=s(self._vect_upts, 'curved'), smats=self._smats_curved))
            if 'linear' in r:
                ks.append(be.kernel(
                    'tflux', tplargs=tplargs | self.base_tpl | {'ktype': 'linear'},
                    dims=[self.nupts, r['linear']],
                    u=s(self.scal_upts[uin], 'linear'),
                    f=s(self._vect_upts, 'linear'), verts=self._verts_linear,
                    upts=self._upts_const))
            return ks[0] if len(ks) == 1 else be.unordered_meta_kernel(ks)

        kernels['tdisf'] = tdisf
        self._add_fused_kernels()

    def _tdisf_qpts(self, uin, tplargs):
        """Flux anti-aliasing: u interpolated to the quadrature points and
        the flux formed there (``baseadvec/elements.py:78-82``,
        ``euler/elements.py:149-206`` with pts = qpts)."""
        be, r, s = self._be, self.mesh_regions, self._slice_mat
        nq = self.basis.nqpts
        ks = [self._qptsu(uin)]
        if 'affine' in r:
            ks.append(be.kernel(
                'tflux', tplargs=tplargs | self.base_tpl | {'ktype': 'affine'},
                dims=[nq, self.neles], u=self._scal_qpts, f=self._vect_qpts,
                smats=self._smats_affine))
        if 'curved' in r:
            ks.append(be.kernel(
                'tflux', tplargs=tplargs | self.base_tpl | {'ktype': 'curved'},
                dims=[nq, r['curved']], u=s(self._scal_qpts, 'curved'),
                f=s(self._vect_qpts, 'curved'), smats=self._smats_curved_q))
        if 'linear' in r:
            ks.append(be.kernel(
                'tflux', tplargs=tplargs | self.base_tpl | {'ktype': 'linear'},
                dims=[nq, r['linear']], u=s(self._scal_qpts, 'linear'),
                f=s(self._vect_qpts, 'linear'), verts=self._verts_linear,
                upts=self._qpts_const))
        return be.ordered_meta_kernel(ks)

    def _xa_mats(self, xa):
        """Stage-update banks of a fused epilogue: none, one bank index, or
        the RK4 pair (r1, r0) of bank indices."""
        if xa is None:
            return None
        if isinstance(xa, tuple):
            return tuple(self.scal_upts[i] for i in xa)
        return self.scal_upts[xa]

    def _add_fused_kernels(self):
        """Fused tdivtpcorf + tdivtconf + negdivconf (+ RK stage update):
        one pass over the element's flux and common normal flux."""
        be = self._be

        def tdivf(fout, xa=None, stage=None):
            return be.kernel(
                'tdivf', self.opmat(self._divop), self.opmat('M3'),
                self._flux_buf, self._scal_fpts, self.scal_upts[fout],
                self._rcpdjac_upts, xa=self._xa_mats(xa), stage=stage)
        self.kernels['tdivf'] = tdivf

        # straight-sided elements, Euler: tdisf folded in as well
        if (self.affine and type(self) is EulerElements and be.fuse_flux and
                not self.fluxaa):
            gamma = self.cfg.getfloat('constants', 'gamma')

            def tdivf_flux(uin, fout, xa=None, stage=None):
                return be.kernel(
                    'tdivf_flux', self.opmat('M1 - M3*M2'),
                    self.opmat('M3'), self.scal_upts[uin],
                    self._smats_affine, self._scal_fpts,
                    self.scal_upts[fout], self._rcpdjac_upts, gamma,
                    xa=self._xa_mats(xa), stage=stage)
            self.kernels['tdivf_flux'] = tdivf_flux


class NavierStokesElements(EulerElements):
    @property
    def _scratch_bufs(self):
        # pyfr/solvers/baseadvecdiff/elements.py:7-20 (gradient fusion, GL)
        return {'scal_fpts', 'vect_fpts', 'vect_upts', 'grad_upts'}

    def set_backend(self, backend, nonce, linoff):
        FluidElements.set_backend(self, backend, nonce, linoff)
        be, kernels = backend, self.kernels
        be.pointwise.register('pyfr.solvers.baseadvec.kernels.negdivconf')
        be.pointwise.register('pyfr.solvers.navstokes.kernels.tflux')

        kernels['disu'] = lambda uin: be.kernel(
            'mul', self.opmat('M0'), self.scal_upts[uin], out=self._scal_fpts)
        kernels['tdivtpcorf'] = lambda fout: be.kernel(
            'mul', self.opmat(self._divop), self._flux_buf,
            out=self.scal_upts[fout])
        kernels['tdivtconf'] = lambda fout: be.kernel(
            'mul', self.opmat('M3'), self._scal_fpts,
            out=self.scal_upts[fout], beta=float(self.basis.order > 0))
        kernels['negdivconf'] = lambda fout: be.kernel(
            'negdivconf', tplargs={'ndims': self.ndims, 'nvars': self.nvars,
                                   'src_macros': []},
            dims=[self.nupts, self.neles], tdivtconf=self.scal_upts[fout],
            rcpdjac=self._rcpdjac_upts)

        # |beta| = 1/2: intconu writes one side only, the other keeps the
        # element's own trace (pyfr/solvers/baseadvecdiff/elements.py:42-45)
        if abs(self.cfg.getfloat('solver-interfaces', 'ldg-beta')) == 0.5:
            kernels['copy_fpts'] = lambda: be.kernel(
                'copy', self._comm_fpts, self._scal_fpts)

        # pyfr/solvers/baseadvecdiff/elements.py:47-55
        kernels['tgradpcoru_upts'] = lambda uin: be.kernel(
            'mul', self.opmat('M4 - M6*M0'), self.scal_upts[uin],
            out=self._grad_upts)
        kernels['tgradcoru_upts'] = lambda: be.kernel(
            'mul', self.opmat('M6'), self._comm_fpts, out=self._grad_upts,
            beta=float(self.basis.order > 0))

        # pyfr/solvers/baseadvecdiff/elements.py:87-101
        def gradcoru_fpts():
            nupts, nfpts = self.nupts, self.nfpts
            vu, vf = self._grad_upts, self._vect_fpts
            muls = [be.kernel('mul', self.opmat('M0'),
                              vu.slice(i*nupts, (i + 1)*nupts),
                              vf.slice(i*nfpts, (i + 1)*nfpts))
                    for i in range(self.ndims)]
            return be.unordered_meta_kernel(muls)
        kernels['gradcoru_fpts'] = gradcoru_fpts

        # pyfr/solvers/navstokes/elements.py:31-128 (tdisf_fused)
        visc_corr = self.cfg.get('solver', 'viscosity-correction', 'none')
        if visc_corr not in {'sutherland', 'none'}:
            raise ValueError('Invalid viscosity-correction option')
        tplargs = {'ndims': self.ndims, 'nvars': self.nvars,
                   'nverts': 8 if self.name == 'hex' else self.nspts,
                   'c': self.cfg.items_as('constants', float),
                   'visc_corr': visc_corr, 'shock_capturing': 'none'}
        r, s = self.mesh_regions, self._slice_mat

        def tdisf_fused(uin):
            ks = []
            if 'affine' in r:
                ks.append(be.kernel(
                    'tflux', tplargs=tplargs | self.base_tpl | {'ktype': 'affine-fused'},
                    dims=[self.nupts, self.neles], u=self.scal_upts[uin],
                    f=self._vect_upts, gradu=self._grad_upts,
                    smats=self._smats_affine, rcpdjac=self._rcpdjac_upts))
            if 'curved' in r:
                ks.append(be.kernel(
                    'tflux', tplargs=tplargs | self.base_tpl | {'ktype': 'curved-fused'},
                    dims=[self.nupts, r['curved']],
                    u=s(self.scal_upts[uin], 'curved'),
                    f=s(self._vect_upts, 'curved'),
                    gradu=s(self._grad_upts, 'curved'),
                    smats=self._smats_curved,
                    rcpdjac=self._rcpdjac_curved(), upts=self._upts_const))
            if 'linear' in r:
                ks.append(be.kernel(
                    'tflux', tplargs=tplargs | self.base_tpl | {'ktype': 'linear-fused'},
                    dims=[self.nupts, r['linear']],
                    u=s(self.scal_upts[uin], 'linear'),
                    f=s(self._vect_upts, 'linear'),
                    gradu=s(self._grad_upts, 'linear'),
                    verts=self._verts_linear, upts=self._upts_const))
            return ks[0] if len(ks) == 1 else be.unordered_meta_kernel(ks)

        if self.fluxaa:
            be.pointwise.register('pyfr.solvers.baseadvecdiff.kernels.gradcoru')

            def tdisf_fused(uin):
                return self._ns_flux_qpts(uin, tplargs)
        kernels['tdisf_fused'] = tdisf_fused

        # Fused tgradpcoru + tgradcoru: grad = (M4 - M6 M0) u + M6 comm
        kernels['tgradf'] = lambda uin: be.kernel(
            'mul2', self.opmat('M4 - M6*M0'), self.opmat('M6'),
            self.scal_upts[uin], self._comm_fpts, self._grad_upts)
        self._add_fused_kernels()

    def _ns_flux_qpts(self, uin, tplargs):
        """Without gradient fusion (flux anti-aliasing): the corrected
        gradient made physical in place at the solution points (gradcoru,
        ``baseadvecdiff/elements.py:68-85``), interpolated to the
        quadrature points (gradcoru_qpts, ``:103-116``) with u (qptsu), and
        the unfused NS flux formed there (``navstokes/elements.py:63-128``).
        Runs between tgradf and gradcoru_fpts, which then interpolates the
        physical gradient."""
        be, r, s = self._be, self.mesh_regions, self._slice_mat
        nu, nq = self.nupts, self.basis.nqpts
        gtpl = {'ndims': self.ndims, 'nvars': self.nvars, 'nverts': 8}
        ks = []
        if 'affine' in r:
            ks.append(be.kernel(
                'gradcoru', tplargs=gtpl | {'ktype': 'affine'},
                dims=[nu, self.neles], gradu=self._grad_upts,
                smats=self._smats_affine, rcpdjac=self._rcpdjac_upts))
        if 'curved' in r:
            ks.append(be.kernel(
                'gradcoru', tplargs=gtpl | {'ktype': 'curved'},
                dims=[nu, r['curved']], gradu=s(self._grad_upts, 'curved'),
                smats=self._smats_curved, rcpdjac=self._rcpdjac_curved()))
        if 'linear' in r:
            ks.append(be.kernel(
                'gradcoru', tplargs=gtpl | {'ktype': 'linear'},
                dims=[nu, r['linear']], gradu=s(self._grad_upts, 'linear'),
                verts=self._verts_linear, upts=self._upts_const))
        gu, gq = self._grad_upts, self._vect_qpts
        ks += [be.kernel('mul', self.opmat('M7'),
                         gu.slice(i*nu, (i + 1)*nu),
                         gq.slice(i*nq, (i + 1)*nq))
               for i in range(self.ndims)]
        ks.append(self._qptsu(uin))
        if 'affine' in r:
            ks.append(be.kernel(
                'tflux', tplargs=tplargs | self.base_tpl | {'ktype': 'affine'},
                dims=[nq, self.neles], u=self._scal_qpts, f=self._vect_qpts,
                smats=self._smats_affine))
        if 'curved' in r:
            ks.append(be.kernel(
                'tflux', tplargs=tplargs | self.base_tpl | {'ktype': 'curved'},
                dims=[nq, r['curved']], u=s(self._scal_qpts, 'curved'),
                f=s(self._vect_qpts, 'curved'), smats=self._smats_curved_q))
        if 'linear' in r:
            ks.append(be.kernel(
                'tflux', tplargs=tplargs | self.base_tpl | {'ktype': 'linear'},
                dims=[nq, r['linear']], u=s(self._scal_qpts, 'linear'),
                f=s(self._vect_qpts, 'linear'), verts=self._verts_linear,
                upts=self._qpts_const))
        return be.ordered_meta_kernel(ks)


# ---------------------------------------------------------------------------
# Interfaces
# ---------------------------------------------------------------------------

def _p_min(be, cfg):
    """The exact solver's pressure floor: ``[solver-interfaces] p-min``,
    default 5 x machine epsilon of the working precision
    (``pyfr/solvers/euler/inters.py:12-26``)."""
    return cfg.getfloat('solver-interfaces', 'p-min',
                        5*float(np.finfo(be.fpdtype).eps))


def _get_inter_arrays(interside, meth, elemap, perm=Ellipsis):
    """``pyfr/solvers/base/inters.py:6-17``."""
    parts, reorder = [], []
    for etype, fidx, eidxs, idx in interside.foreach():
        parts.append(getattr(elemap[etype], meth)(eidxs, fidx))
        reorder.append(np.repeat(idx, elemap[etype].nfacefpts[fidx]))
    if not parts:
        return []
    ro = np.argsort(np.concatenate(reorder), kind='stable')[perm]
    return [np.concatenate(a)[ro] for a in zip(*parts)]


class EulerIntInters:
    """Internal interfaces (``pyfr/solvers/base/inters.py:20-119``,
    ``baseadvec/inters.py:8-29``, ``euler/inters.py:28-37``)."""

    def __init__(self, be, lhs, rhs, elemap, cfg):
        self._be = be
        self.elemap = elemap
        self.cfg = cfg
        first = next(iter(elemap.values()))
        self.ndims, self.nvars = first.ndims, first.nvars
        self.lhs, self.rhs = Interside(lhs), Interside(rhs)
        self.ninters = len(lhs)
        self.ninterfpts = sum(elemap[et].nfacefpts[f]*len(e)
                              for et, f, e in self.lhs.items())
        self.c = cfg.items_as('constants', float)
        self.kernels = {}
        self._perm = Ellipsis
        scal = {t: e._scal_fpts for t, e in elemap.items()}
        self._gen_perm(scal)
        self._pnorm_lhs = self._const_mat(self.lhs, 'get_pnorms_for_inters')

        rsolver = cfg.get('solver-interfaces', 'riemann-solver')
        self._tplargs = dict(ndims=self.ndims, nvars=self.nvars,
                             rsolver=rsolver, c=self.c,
                             p_min=_p_min(be, cfg))
        be.pointwise.register('pyfr.solvers.euler.kernels.intcflux')
        self.kernels['comm_flux'] = lambda: be.kernel(
            'intcflux', tplargs=self._tplargs, dims=[self.ninterfpts],
            ul=self.scal_lhs, ur=self.scal_rhs, nl=self._pnorm_lhs)

    def _gen_perm(self, scal):
        self._perm = self._get_perm_for_field(self.lhs, scal)

    def _const_mat(self, inter, meth):
        m = _get_inter_arrays(inter, meth, self.elemap, self._perm)
        m = m[0] if m else np.empty((0, self.ndims))
        return self._be.const_matrix(np.atleast_2d(m.T))

    def _field_rows(self, inter):
        matmap, rowmap, colmap, reorder = [], [], [], []
        for etype, fidx, eidxs, idx in inter.foreach():
            eles = self.elemap[etype]
            fpts = eles.srtd_face_fpts[fidx][eidxs]
            nfp = fpts.shape[1]
            rowmap.append(fpts.ravel())
            colmap.append(np.repeat(eidxs, nfp))
            matmap.append(np.full(len(eidxs)*nfp, etype))
            reorder.append(np.repeat(idx, nfp))
        ro = np.argsort(np.concatenate(reorder), kind='stable')
        return (np.concatenate(matmap)[ro], np.concatenate(rowmap)[ro],
                np.concatenate(colmap)[ro])

    def _get_perm_for_field(self, inter, field):
        """Order the points by their LHS storage address
        (``pyfr/solvers/base/inters.py:73-90``)."""
        et, r, c = self._field_rows(inter)
        mids = np.array([field[t].mid for t in et])
        v = self._be.view(mids, r, c, vshape=())
        return np.argsort(v.mapping_np)

    def _view(self, inter, meth, vshape=()):
        vm = _get_inter_arrays(inter, meth, self.elemap, self._perm)
        return self._be.view(*vm, vshape=vshape)


class NavierStokesIntInters(EulerIntInters):
    """``baseadvecdiff/inters.py:7-35``, ``navstokes/inters.py:29-47``."""

    def __init__(self, be, lhs, rhs, elemap, cfg):
        self._beta = cfg.getfloat('solver-interfaces', 'ldg-beta')
        super().__init__(be, lhs, rhs, elemap, cfg)
        self.c |= cfg.items_as('solver-interfaces', float)
        visc_corr = cfg.get('solver', 'viscosity-correction', 'none')
        self._tplargs = dict(ndims=self.ndims, nvars=self.nvars,
                             rsolver=self._tplargs['rsolver'], c=self.c,
                             visc_corr=visc_corr, p_min=_p_min(be, cfg))
        self._vect_lhs = self._view(self.lhs, 'get_vect_fpts_for_inters',
                                    (self.ndims, self.nvars))
        self._vect_rhs = self._view(self.rhs, 'get_vect_fpts_for_inters',
                                    (self.ndims, self.nvars))
        self._comm_lhs = self._view(self.lhs, 'get_comm_fpts_for_inters',
                                    (self.nvars,))
        self._comm_rhs = self._view(self.rhs, 'get_comm_fpts_for_inters',
                                    (self.nvars,))
        be.pointwise.register('pyfr.solvers.navstokes.kernels.intconu')
        be.pointwise.register('pyfr.solvers.navstokes.kernels.intcflux')
        self.kernels['con_u'] = lambda: be.kernel(
            'intconu', tplargs=self._tplargs, dims=[self.ninterfpts],
            ulin=self.scal_lhs, urin=self.scal_rhs, ulout=self._comm_lhs,
            urout=self._comm_rhs)
        self.kernels['comm_flux'] = lambda: be.kernel(
            'intcflux', tplargs=self._tplargs, dims=[self.ninterfpts],
            ul=self.scal_lhs, ur=self.scal_rhs, gradul=self._vect_lhs,
            gradur=self._vect_rhs, nl=self._pnorm_lhs)

    def _gen_perm(self, scal):
        side = self.lhs if self._beta != -0.5 else self.rhs
        self._perm = self._get_perm_for_field(side, scal)


# Boundary values each type reads from its [soln-bcs-*] section, as
# expressions of the constants and x, y, z (pyfr/solvers/euler/inters.py:
# 57-102, pyfr/solvers/navstokes/inters.py:111-219), and their defaults
BC_VALUES = {'sup-in-fa': ('rho', 'p', 'u', 'v', 'w'),
             'char-riem-inv': ('rho', 'p', 'u', 'v', 'w'),
             'sub-in-frv': ('rho', 'u', 'v', 'w'),
             'no-slp-isot-wall': ('u', 'v', 'w'),
             'sub-out-fp': ('p',)}
BC_VALUES['char-riem-inv-mass-flow'] = BC_VALUES['char-riem-inv']
BC_VALUES['char-riem-inv-pressure'] = BC_VALUES['char-riem-inv']
BC_DEFAULTS = {'sub-in-frv': {'u': 0, 'v': 0, 'w': 0},
               'no-slp-isot-wall': {'u': 0, 'v': 0, 'w': 0}}
# controlled char-riem-inv types: their target option
CONTROLLED = {'char-riem-inv-mass-flow': 'mass-flow-rate',
              'char-riem-inv-pressure': 'pressure'}


def surface_rule(cfg, sect, kind):
    """(points, weights) of a face quadrature rule of a boundary's surface
    integrals (``pyfr/quadrules/surface.py:40-55``): the rule named by
    [solver-interfaces-<kind>] flux-pts (or the section's
    quad-pts-<kind>) with the fewest points of degree >= quad-deg(-<kind>).
    Quads: Gauss-Legendre; triangles: the tabulated Williams-Shunn rules."""
    from pyfr_amd.quadtables import TABLES
    from pyfr_amd.shapes import gauss_legendre
    rname = cfg.get(f'solver-interfaces-{kind}', 'flux-pts',
                    'gauss-legendre' if kind == 'quad' else 'williams-shunn')
    rname = cfg.get(sect, f'quad-pts-{kind}', rname)
    qdeg = (cfg.getint(sect, f'quad-deg-{kind}')
            if cfg.hasopt(sect, f'quad-deg-{kind}')
            else cfg.getint(sect, 'quad-deg'))
    if kind == 'quad' and rname == 'gauss-legendre':
        n = max(1, -(-(qdeg + 1)//2))
        g, w = gauss_legendre(n)
        return (np.array([(g[i], g[j]) for j in range(n) for i in range(n)]),
                np.array([w[i]*w[j] for j in range(n) for i in range(n)]))
    # stated degrees of the tabulated Williams-Shunn triangle rules, and
    # the rules of that family that are not tabulated
    ws = {3: 2, 6: 4, 10: 5, 15: 7, 36: 12}
    missing = {21: 8, 28: 10, 45: 14}
    if kind == 'tri' and rname == 'williams-shunn':
        cands = sorted((n, d) for n, d in (ws | missing).items() if d >= qdeg)
        if cands and cands[0][0] in ws:
            pts, wts = TABLES[('tri', 'williams-shunn', cands[0][0])]
            return np.array(pts), np.array(wts)
    raise NotImplementedError(f'no tabulated {kind} rule {rname} of degree '
                              f'{qdeg}')


class ControlledBC:
    """The controller of the char-riem-inv-mass-flow / -pressure types
    (``pyfr/solvers/euler/inters.py:108-294``): every ``nsteps`` RHS
    evaluations (from ``tstart``) it measures the boundary's mass flow
    (or mean pressure) by a surface integral of the input state
    (``pyfr/quadrules/surface.py:8-55``), averages it exponentially
    (alpha), corrects the boundary pressure (eta) and interpolates it
    linearly in t; the kernel reads p = t*im + ic."""

    def __init__(self, bc, target_opt):
        cfg, sect = bc.cfg, bc.cfgsect
        self.bc = bc
        self.kind = target_opt
        self.target, self.alpha, self.eta = (
            bc._eval_opt(k) for k in (target_opt, 'alpha', 'eta'))
        self.tstart = cfg.getfloat(sect, 'tstart', 0.0)
        self.nsteps = cfg.getint(sect, 'nsteps', 100)
        self.interp_c = (bc._eval_opt('p') if target_opt == 'mass-flow-rate'
                         else self.target)
        self.interp_m = 0.0
        self.meas_avg = 0.0
        self.tprev = None
        self.counter = 0
        self.pool_idx = {}
        self._bound = None
        self.gamma = cfg.getfloat('constants', 'gamma')
        # surface quadrature per (etype, face): points, weights, m0, normals
        self.groups = []
        area = 0.0
        for etype, fidx, eidxs in bc.lhs.items():
            eles = bc.elemap[etype]
            kind, proj = _face_kind_proj(eles.basis, fidx)
            fp, qw = surface_rule(cfg, sect, kind)
            ppts = np.array([proj(s, t) for s, t in fp], dtype=float)
            m0 = eles.basis.nodal_basis_at(ppts)           # (nq, nupts)
            norm = eles.basis.norm_fpts[eles.basis.facefpts[fidx][0]]
            pn = eles.geom.pnorm_at(ppts, np.array([norm]*len(ppts)))
            pn = pn[:, np.asarray(eidxs)].transpose(2, 0, 1)  # (3, nq, ne)
            self.groups.append((etype, np.asarray(eidxs), m0, qw, pn))
            area += (np.sqrt((pn**2).sum(axis=0))*qw[:, None]).sum()
        self.area = area

    def params(self):
        return {'ic': self.interp_c, 'im': self.interp_m}

    def _measure(self, solns):
        tot = 0.0
        for etype, eidxs, m0, qw, pn in self.groups:
            u = solns[etype][:, :, eidxs]                  # (nupts, 5, ne)
            uf = np.einsum('qu,uve->vqe', m0, u)            # (5, nq, ne)
            if self.kind == 'mass-flow-rate':
                tot += np.einsum('hij,hij', uf[1:-1], pn*qw[None, :, None])
            else:
                rho, E = uf[0], uf[-1]
                p = (self.gamma - 1)*(E - 0.5*(uf[1:-1]**2).sum(axis=0)/rho)
                nmag = np.sqrt((pn**2).sum(axis=0))
                tot += np.einsum('ij,ij', p, nmag*qw[:, None])
        return tot if self.kind == 'mass-flow-rate' else tot/self.area

    def _correction(self, p0, dt):
        if self.kind == 'mass-flow-rate':
            return p0 + dt*self.eta*(1 - self.target/self.meas_avg)
        return p0 + dt*self.eta*(self.target - self.meas_avg)

    def prepare(self, system, uin, t):
        update = self.counter % self.nsteps == 0
        if (update or not self.tprev) and t >= self.tstart:
            solns = dict(zip(system.ele_types, system.ele_scal_upts(uin)))
            meas = self._measure(solns)
            if not self.tprev:
                self.meas_avg = meas
                self.tprev = t
            else:
                a = self.alpha
                self.meas_avg = a*meas + (1 - a)*self.meas_avg
                dt = t - self.tprev
                self.tprev = t
                p0 = self.interp_m*t + self.interp_c
                p1 = self._correction(p0, dt)
                self.interp_m = (p1 - p0)/dt
                self.interp_c = p0 - self.interp_m*t
        vals = (self.interp_c, self.interp_m)
        if vals != self._bound:
            lib, prog = self.bc._be.lib, self.bc._bcprog
            for name, v in zip(('ic', 'im'), vals):
                for i in self.pool_idx.get(name, []):
                    lib.bc_program_set_const(prog, int(i), float(v),
                                             self.bc._be.stream)
            self._bound = vals
        self.counter += 1


def _face_kind_proj(basis, fidx):
    """(face kind, projection (s, t) -> reference point) of face fidx."""
    if basis.name == 'hex':
        from pyfr_amd.shapes import HEX_FACES
        nax, side, sax, tax = HEX_FACES[fidx]

        def proj(s, t):
            p = [0.0, 0.0, 0.0]
            p[nax], p[sax], p[tax] = float(side), s, t
            return p
        return 'quad', proj
    kind, proj, _ = basis.faces[fidx]
    return kind, proj


class EulerBCInters(EulerIntInters):
    """Boundary interfaces (``pyfr/solvers/baseadvec/inters.py:52-116``,
    ``euler/inters.py:40-102``).  The reference evaluates the boundary
    values inside the kernel from expressions of ``ploc`` (and ``t``).
    Here values that do not depend on t are evaluated once on the host at
    the boundary flux points and passed as per-point rows; if any depends
    on t, the expressions are compiled to stack programs
    (``pyfr_amd.bcexpr``) that a kernel evaluates into those rows at each
    RHS, from a device time scalar set before the RHS (``set_time``)."""
    _bcprog = None
    _ctrl = None
    bctypes = ('sup-in-fa', 'sup-out-fn', 'char-riem-inv', 'slp-adia-wall',
               'char-riem-inv-mass-flow', 'char-riem-inv-pressure')
    bccflux_mod = 'pyfr.solvers.euler.kernels.bccflux'

    def __init__(self, be, lhs, elemap, cfgsect, cfg):
        from pyfr_amd.backend import NotSuitableError
        self._be = be
        self.elemap = elemap
        self.cfg, self.cfgsect = cfg, cfgsect
        self.name = cfgsect.removeprefix('soln-bcs-')
        self.type = cfg.get(cfgsect, 'type')
        if self.type not in self.bctypes:
            raise NotSuitableError(f'Boundary type {self.type} is not '
                                   'available')
        first = next(iter(elemap.values()))
        self.ndims, self.nvars = first.ndims, first.nvars
        self.lhs = Interside(lhs)
        self.ninters = len(lhs)
        self.ninterfpts = sum(elemap[et].nfacefpts[f]*len(e)
                              for et, f, e in self.lhs.items())
        self.c = cfg.items_as('constants', float)
        self.kernels = {}
        scal = {t: e._scal_fpts for t, e in elemap.items()}
        self._perm = self._get_perm_for_field(self.lhs, scal)
        self._pnorm_lhs = self._const_mat(self.lhs, 'get_pnorms_for_inters')
        self._bc_consts()
        if self.type in CONTROLLED:
            self._ctrl = ControlledBC(self, CONTROLLED[self.type])
        self._bcvals = self._eval_values()
        rsolver = cfg.get('solver-interfaces', 'riemann-solver')
        # the controlled types are char-riem-inv with a run-time pressure
        # (pyfr/solvers/euler/kernels/bcs/char-riem-inv-controlled.mako)
        ktype = 'char-riem-inv' if self.type in CONTROLLED else self.type
        self._tplargs = dict(ndims=self.ndims, nvars=self.nvars,
                             rsolver=rsolver, c=self.c, bctype=ktype,
                             p_min=_p_min(be, cfg))
        self._kernels()

    def _kernels(self):
        be = self._be
        be.pointwise.register(self.bccflux_mod)
        self.kernels['comm_flux'] = lambda: self._with_values(be.kernel(
            'bccflux', tplargs=self._tplargs, dims=[self.ninterfpts],
            ul=self.scal_lhs, nl=self._pnorm_lhs, bcvals=self._bcvals))

    def _with_values(self, kern):
        """kern preceded, for time-dependent values, by their evaluation."""
        if self._bcprog is None:
            return kern
        from pyfr_amd.backend import Kernel, _dtcode
        lib, prog, n = self._be.lib, self._bcprog, self.ninterfpts
        pl, bv = self._ploc_mat, self._bcvals
        args = (prog, _dtcode(bv.dtype), n, pl.data, pl.leaddim, bv.data,
                bv.leaddim)
        ev = Kernel(lambda stream: lib.bc_eval(*args, stream), mats=(pl, bv),
                    name='bc_eval')
        return self._be.ordered_meta_kernel([ev, kern])

    @property
    def time_dependent(self):
        return self._bcprog is not None

    def set_time(self, t):
        """Set t for the next RHS, on the solver stream."""
        if self._bcprog is not None:
            self._be.lib.bc_program_set_time(self._bcprog, float(t),
                                             self._be.stream)

    def prepare(self, system, uin, t):
        """Before each RHS: the controlled types' update from the input
        bank (``pyfr/solvers/euler/inters.py:161-196``)."""
        if self._ctrl is not None:
            self._ctrl.prepare(system, uin, t)

    def __del__(self):
        try:
            if self._bcprog is not None:
                self._be.lib.bc_program_destroy(self._bcprog)
        except Exception:
            pass

    def _eval_opt(self, k):
        """A scalar option (the reference's ``_eval_opts``)."""
        env = self.cfg.items_as('constants', float)
        return float(_npeval(self.cfg.getexpr(self.cfgsect, k), env))

    def _bc_consts(self):
        pass

    def _eval_values(self):
        import ctypes as ct
        from pyfr_amd.bcexpr import compile_exprs, uses_time
        keys = BC_VALUES.get(self.type)
        if not keys:
            return None
        m = _get_inter_arrays(self.lhs, 'get_ploc_for_inters', self.elemap,
                              self._perm)
        ploc = m[0] if m else np.empty((0, self.ndims))
        consts = self.cfg.items_as('constants', float)
        dflt = BC_DEFAULTS.get(self.type, {})
        exprs = [str(self.cfg.getexpr(self.cfgsect, k, dflt[k]))
                 if k in dflt else
                 't*im + ic' if k == 'p' and self._ctrl else
                 self.cfg.getexpr(self.cfgsect, k) for k in keys]
        if any(uses_time(e) for e in exprs):
            params = self._ctrl.params() if self._ctrl else None
            res = compile_exprs(exprs, consts, params)
            code, offs, pool = res[:3]
            if self._ctrl:
                self._ctrl.pool_idx = res[3]
            prog = ct.c_void_p()
            self._be.lib.bc_program_create(
                ct.byref(prog), len(exprs), code.ctypes.data,
                offs.ctypes.data, pool.ctypes.data if len(pool) else None,
                len(pool))
            self._bcprog = prog.value
            self._ploc_mat = self._be.const_matrix(
                np.ascontiguousarray(ploc.T), tags={'align'})
            return self._be.matrix((len(exprs), self.ninterfpts),
                                   tags={'align'})
        env = consts | dict(zip('xyz', ploc.T))
        rows = [np.broadcast_to(_npeval(e, env), (self.ninterfpts,))
                for e in exprs]
        return self._be.const_matrix(np.array(rows, dtype=float))


class NavierStokesBCInters(EulerBCInters):
    """``baseadvecdiff/inters.py:60-76``, ``navstokes/inters.py:70-219``."""
    bctypes = ('char-riem-inv-mass-flow', 'char-riem-inv-pressure',
               'sup-in-fa', 'sup-out-fn', 'char-riem-inv', 'slp-adia-wall',
               'no-slp-adia-wall', 'no-slp-isot-wall', 'sub-in-frv',
               'sub-in-ftpttang', 'sub-out-fp')
    bccflux_mod = 'pyfr.solvers.navstokes.kernels.bccflux'

    def _bc_consts(self):
        self.c |= self.cfg.items_as('solver-interfaces', float)
        if self.type == 'no-slp-isot-wall':
            self.c['cpTw'] = self._eval_opt('cpTw')
        elif self.type == 'sub-in-ftpttang':
            # navstokes/inters.py:177-203
            gamma = self.cfg.getfloat('constants', 'gamma')
            self.c['cpTt'] = self._eval_opt('cpTt')
            self.c['pt'] = self._eval_opt('pt')
            self.c['Rdcp'] = (gamma - 1.0)/gamma
            theta = self._eval_opt('theta')*np.pi/180.0
            phi = self._eval_opt('phi')*np.pi/180.0
            vc = np.array([np.cos(theta), np.sin(theta), 1.0])
            vc[:2] *= np.sin(phi)
            vc[2] *= np.cos(phi)
            self.c['vc'] = tuple(float(v) for v in vc)

    def _kernels(self):
        be = self._be
        self._tplargs['visc_corr'] = self.cfg.get(
            'solver', 'viscosity-correction', 'none')
        self._vect_lhs = self._view(self.lhs, 'get_vect_fpts_for_inters',
                                    (self.ndims, self.nvars))
        self._comm_lhs = self._view(self.lhs, 'get_comm_fpts_for_inters',
                                    (self.nvars,))
        be.pointwise.register('pyfr.solvers.navstokes.kernels.bcconu')
        be.pointwise.register(self.bccflux_mod)
        # time-dependent values are evaluated ahead of con_u, which runs
        # before comm_flux in every RHS
        self.kernels['con_u'] = lambda: self._with_values(be.kernel(
            'bcconu', tplargs=self._tplargs, dims=[self.ninterfpts],
            ulin=self.scal_lhs, ulout=self._comm_lhs, nlin=self._pnorm_lhs,
            bcvals=self._bcvals))
        self.kernels['comm_flux'] = lambda: be.kernel(
            'bccflux', tplargs=self._tplargs, dims=[self.ninterfpts],
            ul=self.scal_lhs, gradul=self._vect_lhs, nl=self._pnorm_lhs,
            bcvals=self._bcvals)


# ---------------------------------------------------------------------------
# Systems
# ---------------------------------------------------------------------------

class BaseSystem:
    elementscls = None
    intinterscls = None
    name = None

    def __init__(self, backend, mesh, cfg, nregs, nwork=0):
        """nregs register banks in the state precision (bank 0 holds the
        initial condition), then nwork banks in the working precision."""
        self.backend = backend
        self.mesh = mesh
        self.cfg = cfg
        self.ndims = mesh.ndims
        self.nvars = 5
        nonce = 'sys0'

        # Elements, ICs and register banks (system.py:28-105)
        elemap = {}
        for etype, spts in mesh.spts.items():
            elemap[etype] = self.elementscls(SHAPES[etype], spts, cfg)
        ics = {et: e.set_ics_from_cfg() for et, e in elemap.items()}
        self.base = self._base_state(backend, mesh, cfg, elemap, ics)
        if self.base is not None:
            ics = {et: u - self.base[None, :, None] for et, u in ics.items()}
        # curved elements come first; the linear region starts after the
        # last one (pyfr/solvers/base/system.py:241-246)
        curved = getattr(mesh, 'spts_curved', {})
        for et, e in elemap.items():
            c = curved.get(et)
            linoff = 0 if c is None else int(np.max(np.nonzero(c)[0],
                                                    initial=-1) + 1)
            e.set_backend(backend, nonce, linoff=linoff)
        self.ele_map = elemap
        self.ele_types = list(elemap)
        self.ele_ndofs = [e.neles*e.nupts*e.nvars for e in elemap.values()]
        self.ele_shapes = {et: (e.nupts, e.nvars, e.neles)
                           for et, e in elemap.items()}
        self.ele_banks = []
        self.nregs, self.nwork = nregs, nwork
        for et, e in elemap.items():
            banks = [e.alloc_bank(f'bank_{et}', ic=ics[et] if i == 0 else None,
                                  dtype=backend.state_dtype)
                     for i in range(nregs)]
            banks += [e.alloc_bank(f'work_{et}', dtype=backend.fpdtype)
                      for i in range(nwork)]
            self.ele_banks.append(banks)
        backend.commit()
        for e in elemap.values():
            e.post_commit()

        # Interfaces; boundaries from the mesh's bcon_<name> face lists and
        # the [soln-bcs-<name>] sections (pyfr/solvers/base/system.py:
        # 235-262)
        lhs, rhs = mesh.con
        self._int_inters = [self.intinterscls(backend, lhs, rhs, elemap, cfg)]
        self._bc_inters = []
        for name, faces in sorted(getattr(mesh, 'bcon', {}).items()):
            sect = f'soln-bcs-{name}'
            if not cfg.hasopt(sect, 'type'):
                raise ValueError(f'No [{sect}] type for boundary {name}')
            self._bc_inters.append(
                self.bcinterscls(backend, faces, elemap, sect, cfg))
        self._commit()
        self._kcache = {}
        if self.base is not None:
            b = (float(self.base[0]), float(self.base[4]))
            for e in elemap.values():
                e.base_tpl = {'base': b}
            for i in self._int_inters:
                i._tplargs['base'] = b

    @staticmethod
    def _base_state(backend, mesh, cfg, elemap, ics):
        """The constant state u0 = (rho0, 0, 0, 0, E0) of the base-state
        form, or None.  ``[backend-hip] base-state``: 'auto' (default) uses
        it for the precision form ('mixed') where it applies, 'on' in any
        precision (or raises), 'off' never.  It applies to hex meshes
        without boundaries, the Rusanov solver and no flux anti-aliasing.
        rho0 and p0 = (gamma - 1) E0 are the means of the initial density
        and pressure over the solution points.  No counterpart in the
        reference."""
        opt = cfg.get('backend-hip', 'base-state', 'auto')
        if opt not in {'auto', 'on', 'off'}:
            raise ValueError('base-state must be auto, on or off')
        if opt == 'off' or (opt == 'auto' and
                            getattr(backend, 'precision', '') != 'mixed'):
            return None
        why = None
        if any(e.name != 'hex' for e in elemap.values()):
            why = 'hex elements only'
        elif getattr(mesh, 'bcon', {}):
            why = 'meshes without boundaries only'
        elif cfg.get('solver-interfaces', 'riemann-solver') != 'rusanov':
            why = 'the Rusanov solver only'
        elif any(e.fluxaa for e in elemap.values()):
            why = 'no flux anti-aliasing'
        if why:
            if opt == 'on':
                raise ValueError(f'The base-state form applies to {why}')
            return None
        gamma = cfg.getfloat('constants', 'gamma')
        u = np.concatenate([v.transpose(1, 0, 2).reshape(5, -1)
                            for v in ics.values()], axis=1)
        rho0 = float(np.mean(u[0]))
        p = (gamma - 1)*(u[4] - 0.5*(u[1]**2 + u[2]**2 + u[3]**2)/u[0])
        return np.array([rho0, 0.0, 0.0, 0.0, float(np.mean(p))/(gamma - 1)])

    def _commit(self):
        # make_field_views for scal_fpts (baseadvec/system.py:20-55)
        for i in self._int_inters:
            i.scal_lhs = i._view(i.lhs, '_scal_fpts_for_inters', (self.nvars,))
            i.scal_rhs = i._view(i.rhs, '_scal_fpts_for_inters', (self.nvars,))
        for b in self._bc_inters:
            b.scal_lhs = b._view(b.lhs, '_scal_fpts_for_inters', (self.nvars,))

    @property
    def _inters(self):
        """Internal then boundary interfaces (the reference runs their
        kernels of one kind together in one graph stage)."""
        return self._int_inters + self._bc_inters

    @property
    def gndofs(self):
        return sum(self.ele_ndofs)

    def rhs(self, t, uinbank, foutbank):
        """``pyfr/solvers/base/system.py:420-430``."""
        self._set_time(t, uinbank)
        for g in self._rhs_graphs(uinbank, foutbank):
            self.backend.run_graph(g)

    def _set_time(self, t, uin=None):
        """Before each RHS: the controlled boundaries' update from the input
        bank uin, then t for boundary values that depend on it (the
        reference's preparefns and kernel binding,
        pyfr/solvers/base/system.py:408-418)."""
        for b in self._bc_inters:
            if uin is not None:
                b.prepare(self, uin, t)
            b.set_time(t)

    def _rhs_graphs(self, uin, fout):
        key = (uin, fout)
        if key not in self._kcache:
            g = None
            if self.backend.fused:
                g = self._try_fused(uin, fout)
            self._kcache[key] = g or self._build_graphs(uin, fout)
        return self._kcache[key]

    def _try_fused(self, uin, fout, xa=None, stage=None):
        from pyfr_amd.backend import NotSuitableError, flat_kernels
        try:
            ks = self._fused_kernels(uin, fout, xa, stage)
        except NotSuitableError:
            return None
        g = self.backend.graph()
        g.add_all(ks)
        g.commit()
        # every element type's stage kernel takes the stage arguments
        g.stage_kernels = [k for k in flat_kernels(ks) if 'dtb' in k.rtargs]
        return [g]

    @property
    def fused_stages(self):
        """True if RK45 stage updates can be fused into the RHS."""
        if not self.backend.fused:
            return False
        if not hasattr(self, '_fused_ok'):
            from pyfr_amd.backend import NotSuitableError
            w = len(self.ele_banks[0]) - 1     # a working-precision bank
            try:
                self._fused_kernels(w, w, None, None)
                self._fused_ok = True
            except NotSuitableError:
                self._fused_ok = False
        return self._fused_ok

    def rhs_stage(self, t, uin, fout, xa, stage, dta, dtb, last, mixed=False):
        """RHS into fout fused with the rkvdh2 stage update
        (``rkvdh2.mako:4-31``): xa' = xa + dta*k, fout' = xa + dtb*k (if last,
        only xa' = xa + dtb*k).  mixed: xa is the FP64 accumulated solution
        and fout the FP32 stage input."""
        key = (uin, fout, xa, stage, mixed)
        if key not in self._kcache:
            self._kcache[key] = self._try_fused(
                uin, fout, xa, 'rk45-mixed' if mixed else 'rk45')
        self._set_time(t, uin)
        for g in self._kcache[key]:
            for k in g.stage_kernels:
                k.bind(dta=dta, dtb=dtb, last=int(last))
            self.backend.run_graph(g)

    def _fused_rk4_ok(self, mixed):
        """True if every element type's RHS can end in a fused divergence
        kernel with the RK4 stage epilogues (the generated hex kernels, the
        flux-fused tet/prism kernel; mixed: the precision form's, hexes
        only); tried once on banks 2, 3 (stage input, next input) and
        (1, 0) (accumulator, initial state)."""
        key = ('_rk4ok', mixed)
        if key not in self._kcache:
            from pyfr_amd.backend import NotSuitableError
            ok = self.backend.fused and len(self.ele_banks[0]) >= 4
            try:
                if ok:
                    self._fused_kernels(2, 3, (1, 0),
                                        'rk4m-1' if mixed else 'rk4-1')
            except NotSuitableError:
                ok = False
            self._kcache[key] = ok
        return self._kcache[key]

    @property
    def fused_rk4(self):
        return self._fused_rk4_ok(False)

    @property
    def fused_rk4_mixed(self):
        return self._fused_rk4_ok(True)

    def rhs_rk4_stage(self, t, stage, uin, nxt, r1, r0, dt, mixed=False):
        """RHS of RK4 stage `stage` at t with the stage's register updates
        in the divergence kernel's epilogue (``steppers.py:67-108``; mixed:
        the precision form's FP64 updates).  The epilogues read dt from
        device memory (bind writes it on the stream), so each stage's graph
        is captured once, whatever the steps before an output time."""
        from pyfr_amd.backend import flat_kernels
        key = ('rk4', stage, uin, nxt, r1, r0, mixed)
        if key not in self._kcache:
            ks = self._fused_kernels(uin, nxt, (r1, r0),
                                     f'rk4{"m" if mixed else ""}-{stage}')
            g = self.backend.graph()
            g.add_all(ks)
            g.commit()
            g.stage_kernels = [k for k in flat_kernels(ks)
                               if 'dta' in k.devargs]
            self._kcache[key] = g
        g = self._kcache[key]
        for k in g.stage_kernels:
            k.bind(dta=dt)
        self._set_time(t, uin)
        self.backend.run_graph(g)

    def ele_scal_upts(self, idx):
        """The bank's arrays as stored (u - u0 for a solution bank of the
        base-state form)."""
        return [b[idx].get() for b in self.ele_banks]

    def soln_upts(self, idx):
        """The solution in register idx (u0 added back in the base-state
        form)."""
        out = self.ele_scal_upts(idx)
        if self.base is not None:
            out = [a.astype(np.float64) + self.base[None, :, None]
                   for a in out]
        return out

    def set_soln_upts(self, idx, arrs):
        if self.base is not None:
            arrs = [a - self.base[None, :, None] for a in arrs]
        self.set_ele_scal_upts(idx, arrs)

    # -- FP64 reductions over register banks, through the backend's
    #    reduction provider (the reference's residual/integrate plugins use
    #    the same kernel, pyfr/plugins/residual.py, nancheck.py:19-39)
    def _reduce(self, idx, kind):
        """kind 'sum', 'sumsq' (quadrature-weighted) or 'max' (of |u|)
        over solution register idx; the base state is added back (pvars
        b), so the results are those of u."""
        be = self.backend
        if not hasattr(self, '_red_kerns'):
            self._red_kerns, self._red_w = {}, {}
        key = (idx, kind)
        if key not in self._red_kerns:
            ks = []
            pv = {} if self.base is None else {'b': list(self.base)}
            x = 'x' if self.base is None else '(x + b)'
            for et, banks in zip(self.ele_map, self.ele_banks):
                m, e = banks[idx], self.ele_map[et]
                if kind == 'max':
                    ks.append(be.kernel('reduction', 'max', [f'fabs({x})'],
                                        {'x': m}, pvars=pv))
                    continue
                if (et, m.dtype) not in self._red_w:
                    # quadrature weights times |J| at the solution points
                    wts = (e.basis.upts_wts[:, None] /
                           e.rcpdjac_at_np('upts'))
                    self._red_w[(et, m.dtype)] = be.const_matrix(
                        wts, dtype=m.dtype, tags={'align'})
                expr = f'w*{x}' if kind == 'sum' else f'w*{x}*{x}'
                ks.append(be.kernel('reduction', 'sum', [expr],
                                    {'x': m, 'w': self._red_w[(et, m.dtype)]},
                                    pvars=pv))
            self._red_kerns[key] = ks
        ks = self._red_kerns[key]
        be.run_kernels(ks, wait=True)
        return [k.result[0] for k in ks]

    def conservation_sums(self, idx):
        """Per-variable integral of the bank over the domain (sum of
        w_u |J| u), FP64 accumulation."""
        return np.sum(self._reduce(idx, 'sum'), axis=0)

    def l2_norms(self, idx):
        """Per-variable quadrature L2 norm of the bank."""
        return np.sqrt(np.sum(self._reduce(idx, 'sumsq'), axis=0))

    def max_abs(self, idx):
        return np.max(self._reduce(idx, 'max'), axis=0)

    def set_ele_scal_upts(self, idx, arrs):
        for b, a in zip(self.ele_banks, arrs):
            b[idx].set(a)


def _scal_fpts_for_inters(self, eidxs, fidx):
    rmap = self.srtd_face_fpts[fidx][eidxs]
    return self._inters_map(self._scal_fpts.mid, rmap, eidxs, fidx)


FluidElements._scal_fpts_for_inters = _scal_fpts_for_inters


class EulerSystem(BaseSystem):
    name = 'euler'
    elementscls = EulerElements
    intinterscls = EulerIntInters
    bcinterscls = EulerBCInters

    def _fused_kernels(self, uin, fout, xa, stage):
        """The fused RHS as kernels and Concurrent groups: the element
        types' kernels of one phase touch only their own arrays, so a
        captured graph runs them side by side."""
        from pyfr_amd.backend import Concurrent, NotSuitableError
        eles = list(self.ele_map.values())
        ks = [Concurrent([[e.kernels['disu'](uin)] for e in eles])]
        ks += [i.kernels['comm_flux']() for i in self._inters]
        chains = []
        for e in eles:
            if 'tdivf_flux' in e.kernels:
                try:
                    chains.append([e.kernels['tdivf_flux'](uin, fout, xa,
                                                           stage)])
                    continue
                except NotSuitableError:
                    pass
            chains.append([e.kernels['tdisf'](uin),
                           e.kernels['tdivf'](fout, xa, stage)])
        ks.append(Concurrent(chains))
        return ks

    def _build_graphs(self, uin, fout):
        """One graph holding ``g_intf`` then ``g_flux_div`` of
        ``pyfr/solvers/baseadvec/system.py:57-136`` (no MPI to overlap)."""
        from pyfr_amd.backend import Concurrent
        be = self.backend
        g = be.graph()
        eles = list(self.ele_map.values())
        ks = [Concurrent([[e.kernels['disu'](uin)] for e in eles])]
        ks += [i.kernels['comm_flux']() for i in self._inters]
        ks.append(Concurrent([[e.kernels['tdisf'](uin),
                               e.kernels['tdivtpcorf'](fout),
                               e.kernels['tdivtconf'](fout),
                               e.kernels['negdivconf'](fout)]
                              for e in eles]))
        g.add_all(ks)
        g.commit()
        return [g]


class NavierStokesSystem(BaseSystem):
    name = 'navier-stokes'
    elementscls = NavierStokesElements
    intinterscls = NavierStokesIntInters
    bcinterscls = NavierStokesBCInters

    @staticmethod
    def _copy_fpts(eles):
        return [e.kernels['copy_fpts']() for e in eles
                if 'copy_fpts' in e.kernels]

    def _fused_kernels(self, uin, fout, xa, stage):
        from pyfr_amd.backend import Concurrent
        eles = list(self.ele_map.values())
        ks = [Concurrent([[e.kernels['disu'](uin)] + self._copy_fpts([e])
                          for e in eles])]
        ks += [i.kernels['con_u']() for i in self._inters]
        ks.append(Concurrent([[e.kernels['tgradf'](uin),
                               e.kernels['tdisf_fused'](uin),
                               e.kernels['gradcoru_fpts']()] for e in eles]))
        ks += [i.kernels['comm_flux']() for i in self._inters]
        ks.append(Concurrent([[e.kernels['tdivf'](fout, xa, stage)]
                              for e in eles]))
        return ks

    def _build_graphs(self, uin, fout):
        """``pyfr/solvers/baseadvecdiff/system.py:53-231`` with gradient
        fusion, as one graph."""
        from pyfr_amd.backend import Concurrent
        be = self.backend
        g = be.graph()
        eles = list(self.ele_map.values())
        ks = [Concurrent([[e.kernels['disu'](uin)] + self._copy_fpts([e])
                          for e in eles])]
        ks += [i.kernels['con_u']() for i in self._inters]
        ks.append(Concurrent([[e.kernels['tgradpcoru_upts'](uin),
                               e.kernels['tgradcoru_upts'](),
                               e.kernels['tdisf_fused'](uin),
                               e.kernels['gradcoru_fpts']()] for e in eles]))
        ks += [i.kernels['comm_flux']() for i in self._inters]
        ks.append(Concurrent([[e.kernels['tdivtpcorf'](fout),
                               e.kernels['tdivtconf'](fout),
                               e.kernels['negdivconf'](fout)]
                              for e in eles]))
        g.add_all(ks)
        g.commit()
        return [g]


SYSTEMS = {'euler': EulerSystem, 'navier-stokes': NavierStokesSystem}


# ---------------------------------------------------------------------------
# Explicit integrators with the fixed-dt ("none") controller
# ---------------------------------------------------------------------------

class BaseIntegrator:
    nregs = None
    nrhs_per_step = None

    nwork = 0

    def __init__(self, backend, mesh, cfg):
        self.backend = backend
        self.cfg = cfg
        sysname = cfg.get('solver', 'system')
        self.system = SYSTEMS[sysname](backend, mesh, cfg, self.nregs,
                                       self.nwork)
        self.tcurr = cfg.getfloat('solver-time-integrator', 'tstart', 0.0)
        self._tcurr_comp = 0.0
        self.dt = cfg.getfloat('solver-time-integrator', 'dt')
        self.dtmin = cfg.getfloat('solver-time-integrator', 'dt-min', 1e-12)
        self._dt_lookhead = cfg.getint('solver-time-integrator',
                                       'dt-lookahead', 10)
        self.idxcurr = 0
        self.nsteps = 0
        self.nrhsevals = 0
        self._axnpby = {}
        self._wtime = 0.0

    def _rhs(self, t, uin, fout):
        self.system.rhs(t, uin, fout)
        self.nrhsevals += 1

    def _add(self, *args):
        """``pyfr/integrators/base.py:401-426``."""
        consts, regidxs = args[::2], args[1::2]
        if len(regidxs) != len(set(regidxs)):
            raise ValueError('Duplicate register indices')
        if regidxs not in self._axnpby:
            self._axnpby[regidxs] = [
                self.backend.kernel('axnpby', *[em[r] for r in regidxs])
                for em in self.system.ele_banks]
        ks = self._axnpby[regidxs]
        for k in ks:
            k.bind(*consts)
        self.backend.run_kernels(ks)

    def _advance_time(self, dt):
        y = dt - self._tcurr_comp
        t = self.tcurr + y
        self._tcurr_comp = (t - self.tcurr) - y
        self.tcurr = t

    def _clamp_dt(self, dt_want, t):
        remaining = t - self.tcurr
        nsteps = -(-remaining // dt_want)
        if nsteps > self._dt_lookhead:
            return dt_want
        return max(remaining/nsteps, self.dtmin)

    def advance_to(self, t):
        """``pyfr/integrators/explicit/controllers.py:48-60``."""
        if t < self.tcurr:
            raise ValueError('Advance time is in the past')
        while self.tcurr < t:
            dt = self._clamp_dt(self.dt, t)
            self.idxcurr = self.step(self.tcurr, dt)
            self._advance_time(dt)
            self.nsteps += 1

    def run_steps(self, nsteps, dt=None):
        """Take exactly nsteps fixed steps (no clamping); used for timing."""
        dt = self.dt if dt is None else dt
        for _ in range(nsteps):
            self.idxcurr = self.step(self.tcurr, dt)
            self._advance_time(dt)
            self.nsteps += 1

    @property
    def soln(self):
        return self.system.soln_upts(self.idxcurr)


class RK4Integrator(BaseIntegrator):
    """``pyfr/integrators/explicit/steppers.py:60-108``, with one register
    more than the reference: its stages 2-4 compute the RHS in place
    (``rhs(r2, r2)``); here they write r3, so the flux-fused kernels never
    read and write one bank.  Each combination keeps the reference's
    coefficients and operand order (``add(dt/2, r2, 1, r0)`` with r2 = k
    becomes ``add(0, r2, dt/2, r3, 1, r0)``: the same products and sum in
    the axnpby kernel), so the arithmetic is the reference's.  Where every
    element type runs a fused divergence kernel with RK4 epilogues (the
    flux-fused kernel of Euler tets and prisms, the generated kernel of
    Euler hexes), each stage's updates are done in that kernel's epilogue
    instead, the stage inputs alternating between r2 and r3."""
    nregs = 4
    nrhs_per_step = 4

    def step(self, t, dt):
        add, rhs = self._add, self._rhs
        r0, r1, r2, r3 = 0, 1, 2, 3
        if r0 != self.idxcurr:
            r0, r1 = r1, r0
        if self.system.fused_rk4:
            # the same updates in each stage's divergence epilogue; the
            # stage inputs alternate between r2 and r3, so no stage reads
            # and writes one bank
            st = self.system.rhs_rk4_stage
            st(t, 1, r0, r2, r1, r0, dt)
            st(t + dt/2.0, 2, r2, r3, r1, r0, dt)
            st(t + dt/2.0, 3, r3, r2, r1, r0, dt)
            st(t + dt, 4, r2, r2, r1, r0, dt)
            self.nrhsevals += 4
            return r1
        rhs(t, r0, r1)
        add(0.0, r2, 1.0, r0, dt/2.0, r1)
        rhs(t + dt/2.0, r2, r3)
        add(dt/6.0, r1, 1.0, r0, dt/3.0, r3)
        add(0.0, r2, dt/2.0, r3, 1.0, r0)
        rhs(t + dt/2.0, r2, r3)
        add(1.0, r1, dt/3.0, r3)
        add(0.0, r2, dt, r3, 1.0, r0)
        rhs(t + dt, r2, r3)
        add(1.0, r1, dt/6.0, r3)
        return r1


class MixedRK4Integrator(BaseIntegrator):
    """RK4 for the precision form: r0 (the step's initial state) and r1
    (the accumulator, then the new state) in FP64, the stage inputs and
    stage RHS in two FP32 arrays.  The updates are those of
    ``steppers.py:67-108`` formed in FP64 from the FP32 stage RHS; each
    stage's update writes the next stage's FP32 input, and the last writes
    the new state rounded to FP32, the next step's first input.  On hexes
    they run in the divergence kernel's epilogue, elsewhere in
    ``frhip_rk4_mixed`` after the RHS."""
    nregs = 2
    nwork = 2
    nrhs_per_step = 4

    def __init__(self, *args, **kwargs):
        super().__init__(*args, **kwargs)
        self._work_valid = False

    def _sync_work(self, r0):
        lib, be = self.backend.lib, self.backend
        for banks in self.system.ele_banks:
            acc, w = banks[r0], banks[2]
            lib.convert(1, w.data, 0, acc.data, acc.nrow*acc.leaddim,
                        be.stream)
        self._work_valid = True

    def step(self, t, dt):
        r0, r1 = self.idxcurr, 1 - self.idxcurr
        wa, wb = 2, 3
        if not self._work_valid:
            self._sync_work(r0)
        if self.system.fused_rk4_mixed:
            st = self.system.rhs_rk4_stage
            st(t, 1, wa, wb, r1, r0, dt, mixed=True)
            st(t + dt/2.0, 2, wb, wa, r1, r0, dt, mixed=True)
            st(t + dt/2.0, 3, wa, wb, r1, r0, dt, mixed=True)
            st(t + dt, 4, wb, wa, r1, r0, dt, mixed=True)
            self.nrhsevals += 4
            return r1
        lib, be = self.backend.lib, self.backend
        for stage, ts in ((1, t), (2, t + dt/2.0), (3, t + dt/2.0),
                          (4, t + dt)):
            self._rhs(ts, wa, wb)
            for banks in self.system.ele_banks:
                a = banks[r1]
                lib.rk4_mixed(a.nrow, a.ncol, a.leaddim, stage, a.data,
                              banks[r0].data, banks[wb].data, banks[wa].data,
                              dt, be.stream)
        return r1


class RK45Integrator(BaseIntegrator):
    """Low-storage RK45 (``steppers.py:111-190``, ``:218-243``)."""
    nregs = 2
    nrhs_per_step = 5
    a = [970286171893 / 4311952581923, 6584761158862 / 12103376702013,
         2251764453980 / 15575788980749, 26877169314380 / 34165994151039]
    b = [1153189308089 / 22510343858157, 1772645290293 / 4653164025191,
         -1672844663538 / 4480602732383, 2114624349019 / 3568978502595,
         5198255086312 / 14908931495163]

    def __init__(self, *args, **kwargs):
        super().__init__(*args, **kwargs)
        self.backend.pointwise.register(
            'pyfr.integrators.explicit.kernels.rkvdh2')
        self.c = [0.0] + [sum(self.b[:i]) + ai for i, ai in enumerate(self.a)]
        self._nstages = len(self.c)
        self._rk = {}

    def _rkvdh2_kerns(self, stage, r1, r2):
        key = (stage, r1, r2)
        if key not in self._rk:
            tplargs = {'a': self.a, 'b': self.b, 'stage': stage,
                       'nstages': self._nstages, 'nvars': 5, 'errest': False}
            self._rk[key] = [
                self.backend.kernel('rkvdh2', tplargs=tplargs,
                                    dims=[d[0], d[2]], r1=em[r1], r2=em[r2])
                for d, em in zip(self.system.ele_shapes.values(),
                                 self.system.ele_banks)]
        return self._rk[key]

    def step(self, t, dt):
        r1 = self.idxcurr
        r2 = 1 - r1
        last = self._nstages - 1
        fused = self.system.fused_stages
        for i, ci in enumerate(self.c):
            uin = r2 if i > 0 else r1
            if fused:
                # RHS with the stage update in its final kernel
                self.system.rhs_stage(t + ci*dt, uin, r2, r1, i,
                                      dt*self.a[i] if i < last else 0.0,
                                      dt*self.b[i], i == last)
                self.nrhsevals += 1
            else:
                self._rhs(t + ci*dt, uin, r2)
                kerns = self._rkvdh2_kerns(i, r1, r2)
                for k in kerns:
                    k.bind(dt=dt)
                self.backend.run_kernels(kerns)
            r1, r2 = r2, r1
        return r2


class MixedRK45Integrator(RK45Integrator):
    """RK45 for the precision form: the RHS runs in FP32 on one working
    array, the accumulated solution stays FP64.  Each stage update
    (``frhip_rkvdh2_mixed``) is done in FP64 and writes the next FP32 stage
    input in place of the stage RHS, so one FP64 and one FP32 array replace
    the two registers of ``steppers.py:111-190``."""
    nregs = 1
    nwork = 1

    def __init__(self, *args, **kwargs):
        super().__init__(*args, **kwargs)
        self._work_valid = False

    def _sync_work(self):
        lib, be = self.backend.lib, self.backend
        for banks in self.system.ele_banks:
            acc, w = banks[0], banks[1]
            lib.convert(1, w.data, 0, acc.data, acc.nrow*acc.leaddim,
                        be.stream)
        self._work_valid = True

    def step(self, t, dt):
        if not self._work_valid:
            self._sync_work()
        lib, be = self.backend.lib, self.backend
        last = self._nstages - 1
        fused = self.system.fused_stages
        for i, ci in enumerate(self.c):
            dta = dt*self.a[i] if i < last else 0.0
            dtb = dt*self.b[i]
            if fused:
                self.system.rhs_stage(t + ci*dt, 1, 1, 0, i, dta, dtb,
                                      i == last, mixed=True)
                self.nrhsevals += 1
                continue
            self._rhs(t + ci*dt, 1, 1)
            for banks in self.system.ele_banks:
                acc, w = banks[0], banks[1]
                lib.rkvdh2_mixed(acc.nrow, acc.ncol, acc.leaddim, acc.data,
                                 w.data, dta, dtb, int(i == last), be.stream)
        return 0


class RK34Integrator(RK45Integrator):
    """Low-storage RK34 (``steppers.py:193-215``): the rkvdh2 scheme with
    four stages, fused stage updates where RK45's are."""
    nrhs_per_step = 4
    a = [11847461282814 / 36547543011857, 3943225443063 / 7078155732230,
         -346793006927 / 4029903576067]
    b = [1017324711453 / 9774461848756, 8237718856693 / 13685301971492,
         57731312506979 / 19404895981398, -101169746363290 / 37734290219643]


class MixedRK34Integrator(MixedRK45Integrator):
    """RK34 for the precision form (FP64 accumulated solution)."""
    nrhs_per_step = 4
    a, b = RK34Integrator.a, RK34Integrator.b


class EulerIntegrator(BaseIntegrator):
    """Forward Euler (``steppers.py:11-23``): the state stays in one
    register, the RHS goes to the other, ``add(1, u, dt, f)``.  Where the
    RK45 stages are fused, the update is that epilogue's last-stage form
    (u += dt f in the divergence kernel)."""
    nregs = 2
    nrhs_per_step = 1

    def step(self, t, dt):
        ut = self.idxcurr
        f = 1 - ut
        if self.system.fused_stages:
            self.system.rhs_stage(t, ut, f, ut, 0, 0.0, dt, True)
            self.nrhsevals += 1
            return ut
        self._rhs(t, ut, f)
        self._add(1.0, ut, dt, f)
        return ut


class TVDRK3Integrator(BaseIntegrator):
    """TVD-RK3 (``steppers.py:26-57``), three registers, the reference's
    combinations and operand order."""
    nregs = 3
    nrhs_per_step = 3

    def step(self, t, dt):
        add, rhs = self._add, self._rhs
        r0, r1, r2 = 0, 1, 2
        if r0 != self.idxcurr:
            r0, r1 = r1, r0
        rhs(t, r0, r2)
        add(0.0, r1, 1.0, r0, dt, r2)
        rhs(t + dt, r1, r2)
        add(0.25, r1, 0.75, r0, 0.25*dt, r2)
        rhs(t + 0.5*dt, r1, r2)
        add(2.0/3.0, r1, 1.0/3.0, r0, 2.0/3.0*dt, r2)
        return r1


INTEGRATORS = {'euler': EulerIntegrator, 'tvd-rk3': TVDRK3Integrator,
               'rk4': RK4Integrator, 'rk34': RK34Integrator,
               'rk45': RK45Integrator}
MIXED_INTEGRATORS = {'rk4': MixedRK4Integrator, 'rk34': MixedRK34Integrator,
                     'rk45': MixedRK45Integrator}


def get_integrator(backend, mesh, cfg):
    scheme = cfg.get('solver-time-integrator', 'scheme')
    if cfg.get('solver-time-integrator', 'controller', 'none') != 'none':
        raise ValueError('Only the fixed-dt "none" controller is supported')
    if getattr(backend, 'precision', 'double') == 'mixed':
        if scheme not in MIXED_INTEGRATORS:
            raise NotImplementedError('The mixed precision form is '
                                      'implemented for rk4, rk34 and rk45')
        return MIXED_INTEGRATORS[scheme](backend, mesh, cfg)
    return INTEGRATORS[scheme](backend, mesh, cfg)
