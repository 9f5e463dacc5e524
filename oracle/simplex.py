"""Oracle: Euler FR right-hand side on meshes of straight-sided tetrahedra
and prisms (test infrastructure).

Same kernel decomposition as the hex oracle (``pyfr/solvers/baseadvec/
system.py:57-136``): disu -> intcflux -> tdisf -> tdivtpcorf -> tdivtconf ->
negdivconf, with the TetShape / PriShape operators of ``oracle.basis``.

Geometry follows the reference's pipeline for linear elements
(``pyfr/solvers/base/elements.py:331-440``): physical points and the
Jacobian come from the order-1 nodal basis on the shape's ``std_ele(1)``
vertices (``pyfr/shapes.py:502-510``, ``:555-562``); S = |J| J^-1 by
cofactors (``smats.mako:13-25``); normals S^T n.  Interfaces are paired by
geometric brute force (physical flux-point coordinates modulo each block's
periods), independent of the product's node-based connectivity.
"""
import numpy as np

from oracle.basis import PriBasis, SimplexShapeOracle, TetBasis
from oracle.physics import RSOLVERS, inviscid_flux, pri_to_con

STD_VERTS = {
    'tet': np.array([(-1, -1, -1), (1, -1, -1), (-1, 1, -1), (-1, -1, 1)],
                    dtype=float),
    'pri': np.array([(-1, -1, -1), (1, -1, -1), (-1, 1, -1), (-1, -1, 1),
                     (1, -1, 1), (-1, 1, 1)], dtype=float),
}
LINBASIS = {'tet': TetBasis, 'pri': PriBasis}


class _Etype:
    def __init__(self, name, order, tables, verts):
        self.name = name
        self.sh = sh = SimplexShapeOracle(name, order, tables)
        self.verts = verts                          # (nv, 3, E)
        self.neles = verts.shape[-1]
        self.nupts, self.nfpts = sh.nupts, sh.nfpts
        lb = LINBASIS[name](1, STD_VERTS[name])
        self._lb = lb
        self.M0 = sh.opmat('M0')
        self.M132 = sh.opmat('M1 - M3*M2')
        self.M3 = sh.opmat('M3')

        # Jacobian at the solution points: j[a][b] = dX_b/dxi_a
        dl = lb.jac_nodal_basis_at(sh.upts)          # (3, nv, Nu)
        j = np.einsum('avp,vbe->pabe', dl, verts)    # (Nu, 3, 3, E)
        s = np.empty_like(j)
        for i, (a, b) in enumerate([(1, 2), (2, 0), (0, 1)]):
            s[:, i, 0] = j[:, a, 1]*j[:, b, 2] - j[:, a, 2]*j[:, b, 1]
            s[:, i, 1] = j[:, a, 2]*j[:, b, 0] - j[:, a, 0]*j[:, b, 2]
            s[:, i, 2] = j[:, a, 0]*j[:, b, 1] - j[:, a, 1]*j[:, b, 0]
        d = (j[:, 0, 0]*s[:, 0, 0] + j[:, 0, 1]*s[:, 0, 1] +
             j[:, 0, 2]*s[:, 0, 2])
        if np.any(d <= 0):
            raise ValueError('negative Jacobian')
        self.smats = s                               # (Nu, 3, 3, E)
        self.rcpdjac = 1/d                           # (Nu, E)

        # Normals at the flux points from the Jacobian there
        dlf = lb.jac_nodal_basis_at(sh.fpts)
        jf = np.einsum('avp,vbe->pabe', dlf, verts)
        sf = np.empty_like(jf)
        for i, (a, b) in enumerate([(1, 2), (2, 0), (0, 1)]):
            sf[:, i] = np.cross(jf[:, a], jf[:, b], axis=1)
        self.pnorm = np.einsum('pi,pike->pek', sh.norm_fpts, sf)  # (Nf,E,3)

    def ploc(self, pts):
        return np.einsum('pv,vbe->pbe', self._lb.nodal_basis_at(pts),
                         self.verts)                 # (npts, 3, E)


def _wrap(x, period):
    y = np.mod(x, period)
    y[np.isclose(y, period, atol=1e-9, rtol=0)] = 0.0
    return y


def pair_faces_mixed(etypes, blocks):
    """Brute-force pairing of faces (and their flux points) by physical
    coordinates.  blocks: list of (periods (3,), {etype: element indices});
    faces pair only within a block.  Returns a list of
    (tL, eL, idxL, tR, eR, idxR) with flux-point index arrays in matched
    order; the left side is the smaller (etype order, face, element) key, the
    reference's lkey < rkey rule (pyfr/readers/native.py:404-406, :466-470)."""
    order = {n: i for i, n in enumerate(etypes)}
    pairs = []
    for periods, members in blocks:
        periods = np.asarray(periods, dtype=float)
        keys = {}
        for tn, eidx in members.items():
            et = etypes[tn]
            off = np.cumsum([0] + et.sh.nfacefpts)
            pl = et.ploc(et.sh.fpts)                 # (Nf, 3, E)
            for e in eidx:
                for f in range(len(et.sh.nfacefpts)):
                    pts = pl[off[f]:off[f + 1], :, e]
                    c = _wrap(pts.mean(axis=0), periods)
                    key = tuple(np.round(c, 6))
                    keys.setdefault(key, []).append((order[tn], tn, e, f))
        for key, faces in keys.items():
            if len(faces) != 2:
                raise ValueError(f'Unpaired face at {key}: {faces}')
            (_, tL, eL, fL), (_, tR, eR, fR) = sorted(
                faces, key=lambda q: (q[0], q[3], q[2]))
            sl, sr = etypes[tL].sh, etypes[tR].sh
            offl = np.cumsum([0] + sl.nfacefpts)
            offr = np.cumsum([0] + sr.nfacefpts)
            pl = _wrap(etypes[tL].ploc(sl.fpts[offl[fL]:offl[fL + 1]])[..., eL],
                       periods)
            pr = _wrap(etypes[tR].ploc(sr.fpts[offr[fR]:offr[fR + 1]])[..., eR],
                       periods)
            dd = np.abs(pl[:, None, :] - pr[None, :, :])
            dd = np.minimum(dd, periods - dd).sum(axis=2)
            jr = np.argmin(dd, axis=1)
            if (np.any(dd[np.arange(len(pl)), jr] > 1e-8) or
                    len(set(jr.tolist())) != len(pl)):
                raise ValueError('Flux points do not coincide')
            pairs.append((tL, eL, offl[fL] + np.arange(len(pl)),
                          tR, eR, offr[fR] + jr))
    return pairs


class OracleMixedSystem:
    """Euler RHS on a mesh of tets and/or prisms.

    verts: {etype: (nverts, 3, E)}; tables: {etype: tables for
    SimplexShapeOracle}; blocks as for ``pair_faces_mixed``."""

    def __init__(self, verts, order, tables, blocks, rsolver='rusanov',
                 gamma=1.4):
        self.names = list(verts)
        self.et = {n: _Etype(n, order, tables[n], verts[n])
                   for n in self.names}
        self.gamma = gamma
        self.rsolver = RSOLVERS[rsolver]
        pairs = pair_faces_mixed(self.et, blocks)
        self.npairs_faces = len(pairs)
        # flat flux-point pairs (etype, element, fpt) per side
        self.L = [(p[0], p[1], p[2]) for p in pairs]
        self.R = [(p[3], p[4], p[5]) for p in pairs]

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

    def rhs(self, u):
        """u: {etype: (Nu, 5, E)} -> {etype: (Nu, 5, E)}"""
        uf = {n: self._mul(self.et[n].M0, u[n]) for n in self.names}
        # interfaces (intcflux.mako:6-22) with the left side's normal
        lv, rv, nl = [], [], []
        for (tl, el, il), (tr, er, ir) in zip(self.L, self.R):
            lv.append(uf[tl][il, :, el])
            rv.append(uf[tr][ir, :, er])
            nl.append(self.et[tl].pnorm[il, el])
        ul, ur = np.concatenate(lv).T, np.concatenate(rv).T
        nl = np.concatenate(nl).T
        mag = np.sqrt(nl[0]*nl[0] + nl[1]*nl[1] + nl[2]*nl[2])
        fn = self.rsolver(ul, ur, nl/mag, self.gamma)
        q = 0
        for (tl, el, il), (tr, er, ir) in zip(self.L, self.R):
            k = len(il)
            uf[tl][il, :, el] = (mag[q:q + k]*fn[:, q:q + k]).T
            uf[tr][ir, :, er] = (-mag[q:q + k]*fn[:, q:q + k]).T
            q += k

        out = {}
        for n in self.names:
            et = self.et[n]
            f, _, _ = inviscid_flux(u[n].transpose(1, 0, 2), self.gamma)
            F = np.zeros((3, et.nupts, 5, et.neles))
            for i in range(3):
                for k in range(3):
                    F[i] += et.smats[:, i, k][:, None, :] * \
                        f[k].transpose(1, 0, 2)
            div = self._mul(et.M132, F.reshape(3*et.nupts, 5, -1))
            div = self._mul(et.M3, uf[n]) + div
            out[n] = -et.rcpdjac[:, None, :]*div
        return out
