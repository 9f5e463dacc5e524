"""Oracle: Euler and Navier-Stokes FR right-hand sides on meshes of
straight-sided tetrahedra, prisms and pyramids (test infrastructure).

Same kernel decompositions as the hex oracle: Euler ``pyfr/solvers/
baseadvec/system.py:57-136`` (disu -> intcflux -> tdisf -> tdivtpcorf ->
tdivtconf -> negdivconf) and NS ``pyfr/solvers/baseadvecdiff/system.py:
53-231`` (disu -> intconu -> tgradpcoru/tgradcoru -> tdisf_fused ->
gradcoru_fpts -> intcflux -> divergence), with the TetShape / PriShape
operators of ``oracle.basis`` and the boundary conditions of ``oracle.bcs``.

Geometry follows the reference's pipeline for linear elements
(``pyfr/solvers/base/elements.py:331-440``): physical points and the
Jacobian come from the order-1 nodal basis on the shape's ``std_ele(1)``
vertices (``pyfr/shapes.py:502-510``, ``:555-562``); S = |J| J^-1 by
cofactors (``smats.mako:13-25``); normals S^T n.  Interfaces are paired by
geometric brute force (physical flux-point coordinates modulo each block's
periods), independent of the product's node-based connectivity.
"""
import numpy as np

from oracle.basis import HexBasis, HexShapeOracle, PriBasis, PyrBasis, \
    SimplexShapeOracle, TetBasis
from oracle.bcs import bccflux_euler, bccflux_ns, bcconu
from oracle.geom import kopriva_metric
from oracle.physics import RSOLVERS, inviscid_flux, pri_to_con, \
    viscous_flux_add

STD_VERTS = {
    'tet': np.array([(-1, -1, -1), (1, -1, -1), (-1, 1, -1), (-1, -1, 1)],
                    dtype=float),
    'pri': np.array([(-1, -1, -1), (1, -1, -1), (-1, 1, -1), (-1, -1, 1),
                     (1, -1, 1), (-1, 1, 1)], dtype=float),
    'pyr': np.array([(-1, -1, -1), (1, -1, -1), (-1, 1, -1), (1, 1, -1),
                     (0, 0, 1)], dtype=float),
}
LINBASIS = {'tet': TetBasis, 'pri': PriBasis, 'pyr': PyrBasis,
            'hex': HexBasis}


def std_ele(name, k):
    """Equispaced order-k shape points: ``TetShape.std_ele``
    (``pyfr/shapes.py:502-510``), ``PriShape.std_ele`` (``:555-562``) and
    ``PyrShape.std_ele`` (``:619-627``: the 2k + 1 points of [-1, 1],
    every other one per layer)."""
    if name == 'pyr':
        x = np.linspace(-1, 1, 2*k + 1)
        return np.array([(p, q, r) for i, r in enumerate(x[::2])
                         for q in x[i:2*k + 1 - i:2]
                         for p in x[i:2*k + 1 - i:2]])
    x = np.linspace(-1, 1, k + 1)
    if name == 'hex':
        return np.array([(p, q, r) for r in x for q in x for p in x])
    if name == 'tet':
        return np.array([(p, q, r) for i, r in enumerate(x)
                         for j, q in enumerate(x[:k + 1 - i])
                         for p in x[:k + 1 - i - j]])
    return np.array([(p, q, r) for r in x for i, q in enumerate(x)
                     for p in x[:k + 1 - i]])


def shape_order(name, nspts):
    for k in range(1, 16):
        n = {'tet': (k + 1)*(k + 2)*(k + 3)//6, 'hex': (k + 1)**3,
             'pri': (k + 1)**2*(k + 2)//2,
             'pyr': (k + 1)*(k + 2)*(2*k + 3)//6}[name]
        if n == nspts:
            return k
    raise ValueError('Invalid shape point count')


class _Etype:
    def __init__(self, name, order, tables, verts):
        """tables as for SimplexShapeOracle, plus 'qrule' = (points,
        weights) of the element quadrature rule for flux anti-aliasing and
        'iqrules' = {face kind: (points, weights)} for surface-flux
        anti-aliasing."""
        self.name = name
        if name == 'hex':
            # GL solution and flux points (HexShapeOracle); its face points
            # are the tensor GL points the prisms' quad faces also carry
            if set(tables) - {'nq1', 'nf1', 'upts'}:
                raise ValueError('hex tables take only nq1 / nf1 / upts')
            sh = HexShapeOracle(order, tables.get('upts', 'gauss-legendre'))
            if 'nf1' in tables:
                sh.set_surface_quadrature(tables['nf1'])
        else:
            sh = SimplexShapeOracle(name, order, tables)
        self.sh = sh
        if 'iqrules' in tables:
            sh.set_surface_quadrature(tables['iqrules'])
        self.verts = verts                          # (nspts, 3, E)
        self.neles = verts.shape[-1]
        self.nupts, self.nfpts = sh.nupts, sh.nfpts
        # shape basis on the shape points (pyfr/shapes.py:49-50); curved
        # elements (order-k shape points) take the stored metric in the
        # invariant form on std_ele(max(p, 1)) (pyfr/shapes.py:52-61,
        # pyfr/solvers/base/elements.py:382-440)
        k = shape_order(name, verts.shape[0])
        self.curved = k > 1
        self._lb = LINBASIS[name](k, std_ele(name, k))
        if self.curved:
            mord = max(order, 1)
            mpts = std_ele(name, mord)
            self._mb = LINBASIS[name](mord, mpts)
            x = self.ploc(mpts)
            self._sm, self._dj = kopriva_metric(x, self._mb, mpts)
        self.M0 = sh.opmat('M0')
        self.M132 = sh.opmat('M1 - M3*M2')
        self.M3 = sh.opmat('M3')
        self.M460 = sh.opmat('M4 - M6*M0')
        self.M6 = sh.opmat('M6')
        self.smats, d = self._smats_djac(sh.upts)    # (Nu, 3, 3, E)
        if np.any(d <= 0):
            raise ValueError('negative Jacobian')
        self.rcpdjac = 1/d                           # (Nu, E)
        if 'qrule' in tables or 'nq1' in tables:
            if name == 'hex':
                sh.set_quadrature(tables['nq1'])
            else:
                sh.set_quadrature(*tables['qrule'])
            self.nqpts = len(sh.qpts)
            self.M7 = sh.opmat('M7')
            self.M1329 = sh.opmat('(M1 - M3*M2)*M9')
            self.smats_q, _ = self._smats_djac(sh.qpts)

        # Normals S^T n at the flux points (elements.py:364-380)
        sf, _ = self._smats_djac(sh.fpts)
        self.pnorm = np.einsum('pi,pike->pek', sh.norm_fpts, sf)  # (Nf,E,3)

    def _smats_djac(self, pts):
        """S = |J| J^-1 and |J| at reference points: by cofactors of the
        Jacobian j[a][b] = dX_b/dxi_a of the linear map, or interpolated
        from the pseudo-grid points for curved elements."""
        if self.curved:
            m0 = self._mb.nodal_basis_at(pts)
            s = np.einsum('pm,imje->pije', m0, self._sm)
            return s, m0 @ self._dj
        dl = self._lb.jac_nodal_basis_at(pts)        # (3, nv, npts)
        j = np.einsum('avp,vbe->pabe', dl, self.verts)
        s = np.empty_like(j)
        for i, (a, b) in enumerate([(1, 2), (2, 0), (0, 1)]):
            s[:, i, 0] = j[:, a, 1]*j[:, b, 2] - j[:, a, 2]*j[:, b, 1]
            s[:, i, 1] = j[:, a, 2]*j[:, b, 0] - j[:, a, 0]*j[:, b, 2]
            s[:, i, 2] = j[:, a, 0]*j[:, b, 1] - j[:, a, 1]*j[:, b, 0]
        d = (j[:, 0, 0]*s[:, 0, 0] + j[:, 0, 1]*s[:, 0, 1] +
             j[:, 0, 2]*s[:, 0, 2])
        return s, d

    def ploc(self, pts):
        return np.einsum('pv,vbe->pbe', self._lb.nodal_basis_at(pts),
                         self.verts)                 # (npts, 3, E)


def _wrap(x, period):
    """x modulo period along axes with a finite period (inf: no wrap)."""
    period = np.broadcast_to(period, x.shape)
    fin = np.isfinite(period)
    y = np.array(x, dtype=float)
    y[fin] = np.mod(y[fin], period[fin])
    y[fin & np.isclose(y, np.where(fin, period, 0), atol=1e-9, rtol=0)] = 0.0
    return y


def pair_faces_mixed(etypes, blocks, periodic=(True, True, True),
                     boundaries=None):
    """Brute-force pairing of faces (and their flux points) by physical
    coordinates.  blocks: list of (periods (3,), {etype: element indices});
    faces pair only within a block, wrapping only along periodic axes.
    Returns a list of (tL, eL, idxL, tR, eR, idxR) with flux-point index
    arrays in matched order; the left side is the smaller (etype order,
    face, element) key, the reference's lkey < rkey rule
    (pyfr/readers/native.py:404-406, :466-470).  An unpaired face must lie
    on a non-periodic side of its block (coordinate 0 or the block's
    extent, ``periods``); with ``boundaries`` (a dict) such faces are
    recorded as boundaries[side] = [(etype, element, fpt indices)], side
    'x0', 'x1', 'y0', ...; without it they are an error."""
    order = {n: i for i, n in enumerate(etypes)}
    per = np.asarray(periodic, dtype=bool)
    pairs = []
    for periods, members in blocks:
        periods = np.asarray(periods, dtype=float)
        wrap = np.where(per, periods, np.inf)
        keys = {}
        for tn, eidx in members.items():
            et = etypes[tn]
            off = np.cumsum([0] + et.sh.nfacefpts)
            pl = et.ploc(et.sh.fpts)                 # (Nf, 3, E)
            for e in eidx:
                for f in range(len(et.sh.nfacefpts)):
                    pts = pl[off[f]:off[f + 1], :, e]
                    c = _wrap(pts.mean(axis=0), wrap)
                    key = tuple(np.round(c, 6))
                    keys.setdefault(key, []).append((order[tn], tn, e, f))
        for key, faces in keys.items():
            if len(faces) == 1 and boundaries is not None:
                _, tn, e, f = faces[0]
                side = None
                for ax in np.flatnonzero(~per):
                    for hi, x in enumerate((0.0, periods[ax])):
                        if abs(key[ax] - x) < 1e-6:
                            side = f'{"xyz"[ax]}{hi}'
                if side is None:
                    raise ValueError(f'Unpaired interior face at {key}')
                off = np.cumsum([0] + etypes[tn].sh.nfacefpts)
                boundaries.setdefault(side, []).append(
                    (tn, e, np.arange(off[f], off[f + 1])))
                continue
            if len(faces) != 2:
                raise ValueError(f'Unpaired face at {key}: {faces}')
            (_, tL, eL, fL), (_, tR, eR, fR) = sorted(
                faces, key=lambda q: (q[0], q[3], q[2]))
            sl, sr = etypes[tL].sh, etypes[tR].sh
            offl = np.cumsum([0] + sl.nfacefpts)
            offr = np.cumsum([0] + sr.nfacefpts)
            pl = _wrap(etypes[tL].ploc(sl.fpts[offl[fL]:offl[fL + 1]])[..., eL],
                       wrap)
            pr = _wrap(etypes[tR].ploc(sr.fpts[offr[fR]:offr[fR + 1]])[..., eR],
                       wrap)
            dd = np.abs(pl[:, None, :] - pr[None, :, :])
            dd = np.where(per, np.minimum(dd, periods - dd), dd).sum(axis=2)
            jr = np.argmin(dd, axis=1)
            if (np.any(dd[np.arange(len(pl)), jr] > 1e-8) or
                    len(set(jr.tolist())) != len(pl)):
                raise ValueError('Flux points do not coincide')
            pairs.append((tL, eL, offl[fL] + np.arange(len(pl)),
                          tR, eR, offr[fR] + jr))
    return pairs


class OracleMixedSystem:
    """Euler or NS RHS on a mesh of tets and/or prisms.

    verts: {etype: (nverts, 3, E)}; tables: {etype: tables for
    SimplexShapeOracle}; blocks as for ``pair_faces_mixed``.  constants:
    gamma (Euler), plus mu, Pr, ldg-beta, ldg-tau (NS).  periodic: per-axis
    periodicity; each side of a non-periodic axis needs ``bcs[side] =
    (type, vals)`` as for the hex oracle (vals may hold callables of the
    flux-point coordinates)."""

    def __init__(self, verts, order, tables, blocks, rsolver='rusanov',
                 gamma=1.4, system='euler', constants=None,
                 periodic=(True, True, True), bcs=None, antialias=()):
        """antialias: ('flux',) forms the flux at each type's element
        quadrature rule (``tables[etype]['qrule']``) and projects it back
        (``pyfr/solvers/baseadvec/elements.py:78-94``,
        ``baseadvecdiff/elements.py:103-116``)."""
        self.names = list(verts)
        self.fluxaa = 'flux' in antialias
        self.et = {n: _Etype(n, order, tables[n], verts[n])
                   for n in self.names}
        self.system = system
        self.c = dict(constants or {'gamma': gamma})
        self.gamma = self.c['gamma']
        self.rsolver = RSOLVERS[rsolver]
        bfaces = {}
        pairs = pair_faces_mixed(self.et, blocks, periodic, bfaces)
        self.npairs_faces = len(pairs)

        # one flat flux-point index over all element types: (type, fpt,
        # element) -> off[type] + fpt*E + element
        self._off, o = {}, 0
        for n in self.names:
            self._off[n] = o
            o += self.et[n].nfpts*self.et[n].neles
        self._nflat = o
        g = self._gidx
        self.gl = np.concatenate([g(p[0], p[2], p[1]) for p in pairs])
        self.gr = np.concatenate([g(p[3], p[5], p[4]) for p in pairs])
        pn = self._flat({n: self.et[n].pnorm.transpose(0, 2, 1)
                         for n in self.names}, 3)    # (nflat, 3)
        self.nl = pn[self.gl].T                      # (3, npts)
        ploc = self._flat({n: self.et[n].ploc(self.et[n].sh.fpts)
                           for n in self.names}, 3)
        self.bcs = []
        for side, faces in sorted(bfaces.items()):
            bctype, vals = (bcs or {})[side]
            gb = np.concatenate([g(t, i, e) for t, e, i in faces])
            x, y, z = ploc[gb].T
            v = {k: (fn(x, y, z) if callable(fn) else fn)
                 for k, fn in vals.items()}
            self.bcs.append((side, bctype, gb, pn[gb].T, v))

    def _gidx(self, t, fpts, e):
        return self._off[t] + np.asarray(fpts)*self.et[t].neles + e

    def _flat(self, d, nv=5):
        """{etype: (..., Nf, nv, E)} -> (..., nflat, nv)"""
        return np.concatenate(
            [np.moveaxis(d[n], -1, -2).reshape(*d[n].shape[:-3], -1, nv)
             for n in self.names], axis=-2)

    def _unflat(self, a, nv=5):
        out = {}
        for n in self.names:
            et = self.et[n]
            k = et.nfpts*et.neles
            o = self._off[n]
            b = a[..., o:o + k, :].reshape(*a.shape[:-2], et.nfpts,
                                           et.neles, nv)
            out[n] = np.moveaxis(b, -2, -1)
        return out

    def ics(self, fns):
        """fns: dict of callables rho, u, v, w, p of (x, y, z)."""
        out = {}
        for n, et in self.et.items():
            x, y, z = et.ploc(et.sh.upts).transpose(1, 0, 2)
            pri = [fns[k](x, y, z) for k in ('rho', 'u', 'v', 'w', 'p')]
            out[n] = np.stack(pri_to_con(pri, self.gamma), axis=1)
        return out

    @staticmethod
    def _mul(A, B):
        return (A @ B.reshape(A.shape[1], -1)).reshape(A.shape[0],
                                                        *B.shape[1:])

    def _transform(self, et, f, s=None):
        """F[i] = sum_k S[i][k] f[k] (tflux.mako:24-28); f (3, 5, Np, E)
        at the solution points, or with s at other points."""
        s = et.smats if s is None else s
        F = np.zeros((3, s.shape[0], 5, et.neles))
        for i in range(3):
            for k in range(3):
                F[i] += s[:, i, k][:, None, :]*f[k].transpose(1, 0, 2)
        return F

    def _divergence(self, u, F, UF):
        uf = self._unflat(UF)
        out = {}
        for n in self.names:
            et = self.et[n]
            if self.fluxaa:
                div = self._mul(et.M1329, F[n].reshape(3*et.nqpts, 5, -1))
            else:
                div = self._mul(et.M132, F[n].reshape(3*et.nupts, 5, -1))
            div = self._mul(et.M3, uf[n]) + div
            out[n] = -et.rcpdjac[:, None, :]*div
        return out

    def rhs(self, u):
        """u: {etype: (Nu, 5, E)} -> {etype: (Nu, 5, E)}"""
        if self.system == 'navier-stokes':
            return self.rhs_ns(u)
        UF = self._flat({n: self._mul(self.et[n].M0, u[n])
                         for n in self.names})
        # interfaces (intcflux.mako:6-22) with the left side's normal
        ul, ur, nl = UF[self.gl].T, UF[self.gr].T, self.nl
        mag = np.sqrt(nl[0]*nl[0] + nl[1]*nl[1] + nl[2]*nl[2])
        fn = self.rsolver(ul, ur, nl/mag, self.gamma)
        bout = [bccflux_euler(bt, self.rsolver, UF[gb].T, bn, v, self.c)
                for _, bt, gb, bn, v in self.bcs]
        UF[self.gl] = (mag*fn).T
        UF[self.gr] = (-mag*fn).T
        for (_, _, gb, _, _), o in zip(self.bcs, bout):
            UF[gb] = o.T
        F = {}
        for n in self.names:
            et = self.et[n]
            if self.fluxaa:                                    # qptsu
                uq = self._mul(et.M7, u[n])
                f, _, _ = inviscid_flux(uq.transpose(1, 0, 2), self.gamma)
                F[n] = self._transform(et, f, et.smats_q)
            else:
                f, _, _ = inviscid_flux(u[n].transpose(1, 0, 2), self.gamma)
                F[n] = self._transform(et, f)
        return self._divergence(u, F, UF)

    def rhs_ns(self, u):
        """``pyfr/solvers/baseadvecdiff/system.py:53-231`` with the
        interface kernels ``intconu.mako:4-19`` and
        ``navstokes/kernels/intcflux.mako:10-54``."""
        c = self.c
        beta, tau = c['ldg-beta'], c['ldg-tau']
        UF = self._flat({n: self._mul(self.et[n].M0, u[n])
                         for n in self.names})                 # disu
        ul, ur = UF[self.gl].T, UF[self.gr].T
        COMM = np.zeros_like(UF)                               # intconu
        cu = ur*(0.5 + beta) + ul*(0.5 - beta)
        COMM[self.gl] = cu.T
        COMM[self.gr] = cu.T
        for _, bt, gb, bn, v in self.bcs:                      # bcconu
            COMM[gb] = bcconu(bt, UF[gb].T, bn, v, c).T
        comm = self._unflat(COMM)

        F, GF = {}, []
        for n in self.names:
            et = self.et[n]
            grad = self._mul(et.M460, u[n])                    # tgradpcoru
            grad = self._mul(et.M6, comm[n]) + grad            # tgradcoru
            grad = grad.reshape(3, et.nupts, 5, et.neles)
            # transform_grad.mako:3-10
            s = et.smats
            gp = np.zeros_like(grad)
            for i in range(3):
                acc = 0
                for k in range(3):
                    acc = acc + s[:, k, i][:, None, :]*grad[k]
                gp[i] = et.rcpdjac[:, None, :]*acc
            if self.fluxaa:
                # gradcoru_qpts and qptsu, the flux at the quadrature points
                uq = self._mul(et.M7, u[n])
                gq = np.stack([self._mul(et.M7, gp[d]) for d in range(3)])
                f, _, _ = inviscid_flux(uq.transpose(1, 0, 2), c['gamma'])
                viscous_flux_add(uq.transpose(1, 0, 2),
                                 gq.transpose(0, 2, 1, 3), f, c)
                F[n] = self._transform(et, f, et.smats_q)
            else:
                f, _, _ = inviscid_flux(u[n].transpose(1, 0, 2), c['gamma'])
                viscous_flux_add(u[n].transpose(1, 0, 2),
                                 gp.transpose(0, 2, 1, 3), f, c)
                F[n] = self._transform(et, f)
            GF.append(np.stack([self._mul(et.M0, gp[d])
                                for d in range(3)]))   # gradcoru_fpts
        GF = self._flat(dict(zip(self.names, GF)))             # (3, nflat, 5)

        gl = GF[:, self.gl].transpose(0, 2, 1)                 # (3, 5, np)
        gr = GF[:, self.gr].transpose(0, 2, 1)
        nl = self.nl
        mag = np.sqrt(nl[0]*nl[0] + nl[1]*nl[1] + nl[2]*nl[2])
        nn = (1/mag)*nl
        fi = self.rsolver(ul, ur, nn, c['gamma'])
        fvl = viscous_flux_add(ul, gl, np.zeros((3, 5, ul.shape[1])), c)
        fvr = viscous_flux_add(ur, gr, np.zeros((3, 5, ur.shape[1])), c)
        out = np.empty_like(fi)
        for i in range(5):
            nfl = nn[0]*fvl[0, i] + nn[1]*fvl[1, i] + nn[2]*fvl[2, i]
            nfr = nn[0]*fvr[0, i] + nn[1]*fvr[1, i] + nn[2]*fvr[2, i]
            if beta == -0.5:
                fv = nfr
            elif beta == 0.5:
                fv = nfl
            else:
                fv = (0.5 + beta)*nfl + (0.5 - beta)*nfr
            if tau != 0.0:
                fv = fv + tau*(ul[i] - ur[i])
            out[i] = fi[i] + fv
        bout = [bccflux_ns(bt, self.rsolver, UF[gb].T,
                           GF[:, gb].transpose(0, 2, 1), bn, v, c)
                for _, bt, gb, bn, v in self.bcs]
        UF[self.gl] = (mag*out).T
        UF[self.gr] = (-mag*out).T
        for (_, _, gb, _, _), o in zip(self.bcs, bout):        # bccflux
            UF[gb] = o.T
        return self._divergence(u, F, UF)
