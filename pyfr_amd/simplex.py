"""FR operators for tetrahedra, prisms and pyramids (dense operators).

Shapes follow the reference's ``TetShape``, ``PriShape`` and ``PyrShape``
(``pyfr/shapes.py:476-640``): reference elements, face tables (projections
and unnormalised reference normals), solution points (Shunn-Ham tets,
Williams-Shunn x Gauss-Legendre prisms, Gauss-Legendre pyramids), flux
points (Williams-Shunn triangles, Gauss-Legendre quads) and the operator
layouts m0..m6 (``pyfr/shapes.py:89-135``, ``:231-259``).

The orthonormal bases are built here independently of the reference's Jacobi
formulas: products of Legendre polynomials span the polynomial space
(P_p for tets and triangles, P_p(tri) x P_p(line) for prisms) and are
orthonormalised by Cholesky against an exact collapsed (Duffy) Gauss-Legendre
quadrature (for pyramids, homogenised Legendre products, ``_span_pyr``).  The
nodal operators and the correction functions (which use the space's
reproducing kernel) do not depend on which orthonormal basis spans
the space, so the result matches the reference's operators to rounding.
"""
from functools import cached_property
import itertools as it

import numpy as np

from pyfr_amd.quadtables import TABLES
from pyfr_amd.shapes import gauss_legendre


def _leg(n, x):
    """P_0..P_n and derivatives at x (arrays (n+1, len(x)))."""
    x = np.asarray(x, dtype=float)
    P = np.zeros((n + 1,) + x.shape)
    D = np.zeros_like(P)
    P[0] = 1.0
    if n >= 1:
        P[1] = x
        D[1] = 1.0
    for k in range(2, n + 1):
        P[k] = ((2*k - 1)*x*P[k - 1] - (k - 1)*P[k - 2])/k
        D[k] = D[k - 2] + (2*k - 1)*P[k - 1]
    return P, D


def _degrees(kind, p):
    if kind == 'tri':
        return [(i, j) for i in range(p + 1) for j in range(p + 1 - i)]
    if kind == 'tet':
        return [(i, j, k) for i in range(p + 1) for j in range(p + 1 - i)
                for k in range(p + 1 - i - j)]
    if kind == 'pri':
        return [(i, j, k) for i in range(p + 1) for j in range(p + 1 - i)
                for k in range(p + 1)]
    if kind == 'pyr':
        return [(i, j, k) for i in range(p + 1) for j in range(p + 1)
                for k in range(p + 1 - max(i, j))]
    if kind == 'quad':
        return [(i, j) for i in range(p + 1) for j in range(p + 1)]
    if kind == 'line':
        return [(i,) for i in range(p + 1)]
    raise ValueError(kind)


def _collapsed_rule(kind, n):
    """Exact (for degree <= 2n - 3 after the Duffy Jacobian) quadrature on
    the reference element from tensor Gauss-Legendre points."""
    g, w = gauss_legendre(n)
    if kind == 'line':
        return g[:, None], w
    if kind == 'quad':
        a, b = np.meshgrid(g, g, indexing='ij')
        wa, wb = np.meshgrid(w, w, indexing='ij')
        return np.stack([a.ravel(), b.ravel()], 1), (wa*wb).ravel()
    if kind == 'tri':
        a, b = np.meshgrid(g, g, indexing='ij')
        wa, wb = np.meshgrid(w, w, indexing='ij')
        a, b, wt = a.ravel(), b.ravel(), (wa*wb).ravel()
        x = (1 + a)*(1 - b)/2 - 1
        return np.stack([x, b], 1), wt*(1 - b)/2
    if kind == 'tet':
        a, b, c = np.meshgrid(g, g, g, indexing='ij')
        wa, wb, wc = np.meshgrid(w, w, w, indexing='ij')
        a, b, c = a.ravel(), b.ravel(), c.ravel()
        wt = (wa*wb*wc).ravel()
        y = (1 + b)*(1 - c)/2 - 1
        x = (1 + a)*(1 - b)*(1 - c)/4 - 1
        return np.stack([x, y, c], 1), wt*(1 - b)*(1 - c)**2/8
    if kind == 'pri':
        t, tw = _collapsed_rule('tri', n)
        pts = np.array([(*tp, z) for z in g for tp in t])
        wts = np.array([tw_*wz for wz in w for tw_ in tw])
        return pts, wts
    if kind == 'pyr':
        # (x, y, z) = (a s, b s, c) with s = (1 - c)/2 over the cube
        a, b, c = np.meshgrid(g, g, g, indexing='ij')
        wa, wb, wc = np.meshgrid(w, w, w, indexing='ij')
        a, b, c = a.ravel(), b.ravel(), c.ravel()
        s = (1 - c)/2
        return (np.stack([a*s, b*s, c], 1),
                (wa*wb*wc).ravel()*s**2)
    raise ValueError(kind)


def _homleg(n, x, s):
    """Q_k(x, s) = s^k P_k(x/s), k = 0..n, and its partial derivatives in
    x and s, by the Legendre recurrence multiplied through by s^k (a
    polynomial in x and s, regular where s = 0)."""
    Q = np.zeros((n + 1,) + x.shape)
    Qx, Qs = np.zeros_like(Q), np.zeros_like(Q)
    Q[0] = 1.0
    if n >= 1:
        Q[1], Qx[1] = x, 1.0
    for k in range(2, n + 1):
        a, b = (2*k - 1)/k, (k - 1)/k
        Q[k] = a*x*Q[k - 1] - b*s*s*Q[k - 2]
        Qx[k] = a*(Q[k - 1] + x*Qx[k - 1]) - b*s*s*Qx[k - 2]
        Qs[k] = a*x*Qs[k - 1] - b*(2*s*Q[k - 2] + s*s*Qs[k - 2])
    return Q, Qx, Qs


class OrthoBasis:
    """Orthonormal basis of the order-p space on a reference element."""

    def __init__(self, kind, order):
        self.kind, self.order = kind, order
        self.degs = _degrees(kind, order)
        # the Gram matrix's integrands have degree <= 2p in each collapsed
        # coordinate, plus 2(p + 1) in c for pyramids (s^(i + j) factors and
        # the s^2 Jacobian)
        qp, qw = _collapsed_rule(kind, 2*order + 3 if kind == 'pyr'
                                 else order + 3)
        V = self._span(qp)                               # (nb, nq)

        # phi = C @ span; the Legendre-product Gram matrix is ill
        # conditioned on simplices, so orthonormalise twice (the second
        # pass removes the first pass's rounding, as in re-orthogonalised
        # Gram-Schmidt)
        C = np.eye(len(V))
        for _ in range(2):
            W = C @ V
            C = np.linalg.inv(np.linalg.cholesky((W*qw) @ W.T)) @ C
        self._C = C

    def _span(self, pts, deriv=None):
        pts = np.atleast_2d(pts)
        p = self.order
        if self.kind == 'pyr':
            return self._span_pyr(pts, deriv)
        nd = pts.shape[1]
        PD = [_leg(p, pts[:, d]) for d in range(nd)]
        out = []
        for dg in self.degs:
            v = np.ones(len(pts))
            for d in range(nd):
                P, D = PD[d]
                v = v*(D[dg[d]] if d == deriv else P[dg[d]])
            out.append(v)
        return np.array(out)

    def _span_pyr(self, pts, deriv):
        """The pyramid space: Q_i(x, s) Q_j(y, s) P_k(z), s = (1 - z)/2,
        i, j <= p, k <= p - max(i, j).  For fixed (i, j) the reference's
        (1 - z)^(i + j) P_i(2x/(1 - z)) P_j(2y/(1 - z)) P_k^(2(i+j+1),0)(z)
        (pyfr/polys.py:332-392) is 2^(i + j) Q_i Q_j times a degree-k
        polynomial in z, so both span the same space."""
        x, y, z = pts.T
        s = (1 - z)/2
        p = self.order
        qx, qxx, qxs = _homleg(p, x, s)
        qy, qyy, qys = _homleg(p, y, s)
        L, D = _leg(p, z)
        out = []
        for i, j, k in self.degs:
            if deriv is None:
                v = qx[i]*qy[j]*L[k]
            elif deriv == 0:
                v = qxx[i]*qy[j]*L[k]
            elif deriv == 1:
                v = qx[i]*qyy[j]*L[k]
            else:
                v = (qx[i]*qy[j]*D[k] -
                     (qxs[i]*qy[j] + qx[i]*qys[j])*L[k]/2)
            out.append(v)
        return np.array(out)

    def at(self, pts):
        """(nbasis, npts)"""
        return self._C @ self._span(pts)

    def jac_at(self, pts):
        """(ndims, nbasis, npts)"""
        nd = np.atleast_2d(pts).shape[1]
        return np.array([self._C @ self._span(pts, d) for d in range(nd)])


class NodalBasis:
    """Lagrange basis through points ``pts`` of the order-p space."""

    def __init__(self, kind, order, pts):
        self.ortho = OrthoBasis(kind, order)
        self.pts = np.atleast_2d(np.asarray(pts, dtype=float))
        # l_n(x) = sum_b (V^-1)_nb phi_b(x) with V_bn = phi_b(x_n)
        self.invvdm = np.linalg.inv(self.ortho.at(self.pts))

    def nodal_basis_at(self, epts):
        """(npts, nnodes)"""
        return (self.invvdm @ self.ortho.at(epts)).T

    def jac_nodal_basis_at(self, epts):
        """(ndims, nnodes, npts): the layout of ``polys.py:63-66``."""
        return self.invvdm @ self.ortho.jac_at(epts)


def _proj(fn, pts2):
    return np.array([fn(s, t) for s, t in pts2], dtype=float)


def _flush(a, tol=1e-10):
    a = np.array(a, dtype=float, copy=True)
    a[np.abs(a) < tol] = 0
    return a


# Element quadrature rule of flux anti-aliasing per order: the rule the
# reference's default selection picks (pyfr/shapes.py:146-176: the fewest
# points of degree >= 2p + 1 for tets, >= 2(p + 2) for prisms)
AA_RULES = {
    'tet': {1: ('witherden-vincent', 8), 2: ('witherden-vincent', 14),
            3: ('witherden-vincent', 35), 4: ('witherden-vincent', 59)},
    'pri': {1: ('witherden-vincent', 28), 2: ('witherden', 46),
            3: ('witherden', 82), 4: ('witherden', 128)},
    # pyramids: degree >= 2(p + 1)
    'pyr': {1: ('witherden-vincent', 10), 2: ('witherden-vincent', 24),
            3: ('witherden', 44), 4: ('witherden', 76)},
}
# Triangle-face rule of surface-flux anti-aliasing per order (the fewest
# points among the symmetric rules of degree >= 2(p + 1),
# pyfr/shapes.py:159-185); quad faces take (p + 2)^2 Gauss-Legendre
SF_TRI_RULES = {1: ('williams-shunn', 6), 2: ('witherden-vincent', 12),
                3: ('witherden-vincent', 16), 4: ('witherden-vincent', 25)}


def qdeg_rule(shape, qdeg=None, name=None, npts=None):
    """(rule, npts) of a tabulated quadrature rule chosen as the
    reference chooses (``pyfr/quadrules/__init__.py:52-79``): among the
    rules of the shape of degree >= qdeg (and of the named family, and of
    npts points, where given), the fewest points when qdeg is given, the
    highest degree when only npts is; ties go to the first in name order.
    Tets, prisms and triangles need the chosen rule's points tabulated."""
    from pyfr_amd.quadtables import CATALOG, TABLES
    best = None
    for r, n, d, _ in CATALOG[shape]:
        if ((not name or name == r) and (not npts or npts == n) and
                (not qdeg or qdeg <= d)):
            if (best is None or (npts and d > best[2]) or
                    (qdeg and n < best[1])):
                best = (r, n, d)
    if best is None:
        raise ValueError(f'No suitable {shape} quadrature rule found')
    if shape in ('tet', 'pri', 'tri') and (shape, *best[:2]) not in TABLES:
        raise NotImplementedError(f'{shape} quadrature rule {best[:2]} is '
                                  'not tabulated')
    return best[:2]


class _SimplexShape:
    ndims = 3
    name = None
    faces = None

    # the default solution point family ([solver-elements-<name>]
    # soln-pts)
    upts_default = None

    def __init__(self, order, qrule=None, iqrules=None, upts_name=None,
                 tri_fpts=None):
        """qrule: (rule name, npts) of the element quadrature rule of flux
        anti-aliasing, or None.  iqrules: surface-flux anti-aliasing, the
        face rules {'tri': (rule name, npts), 'quad': Gauss-Legendre points
        per direction}; the flux points of each face are that rule's.
        upts_name: the solution point family (soln-pts), tri_fpts: the
        triangle-face flux point family (flux-pts); None for the defaults."""
        self.order = order
        self.upts_name = upts_name or self.upts_default
        self.tri_fpts = tri_fpts or 'williams-shunn'
        for key in ((self.name, self.upts_name, self._nupts_of(order)),
                    ('tri', self.tri_fpts, (order + 1)*(order + 2)//2)):
            if key not in TABLES:
                raise NotImplementedError(f'{key[0]} point set {key[1]} '
                                          f'with {key[2]} points is not '
                                          'tabulated')
        self.qrule = qrule
        self.iqrules = iqrules
        p = order
        self.nupts = len(self.upts)
        self.ubasis = NodalBasis(self.name, p, self.upts)
        self.nfacefpts = [len(self._face_rule(k)[0]) for k, _, _ in self.faces]
        self.nfpts = sum(self.nfacefpts)
        nf = np.cumsum([0] + self.nfacefpts)
        self.facefpts = [list(range(nf[i], nf[i + 1]))
                         for i in range(len(self.faces))]

    # standard face flux points: Williams-Shunn (tri) / Gauss-Legendre
    # (quad); the face bases live on them
    def _std_face_pts(self, kind):
        p = self.order
        if kind == 'tri':
            pts, _ = TABLES[('tri', self.tri_fpts, (p + 1)*(p + 2)//2)]
            return np.array(pts)
        g, _ = gauss_legendre(p + 1)
        return np.array([(g[i], g[j]) for j in range(p + 1)
                         for i in range(p + 1)])

    def _iq_rule(self, kind):
        """(points, weights) of the interface rule of surface-flux
        anti-aliasing."""
        if kind == 'tri':
            pts, wts = TABLES[('tri', *self.iqrules['tri'])]
            return np.array(pts), np.array(wts)
        n = self.iqrules['quad']
        g, w = gauss_legendre(n)
        return (np.array([(g[i], g[j]) for j in range(n) for i in range(n)]),
                np.array([w[i]*w[j] for j in range(n) for i in range(n)]))

    # the flux points of each face: the standard set, or the interface rule
    def _face_rule(self, kind):
        if self.iqrules:
            return self._iq_rule(kind)[0], None
        return self._std_face_pts(kind), None

    @cached_property
    def fpts(self):
        return np.vstack([_proj(proj, self._face_rule(kind)[0])
                          for kind, proj, _ in self.faces])

    @cached_property
    def norm_fpts(self):
        return np.vstack([[nrm]*n for (_, _, nrm), n in
                          zip(self.faces, self.nfacefpts)]).astype(float)

    @cached_property
    def facebases(self):
        out = {}
        for kind in {k for k, _, _ in self.faces}:
            out[kind] = NodalBasis(kind, self.order, self._std_face_pts(kind))
        return out

    @cached_property
    def gbasis_coeffs(self):
        """Correction functions from face quadrature
        (``pyfr/shapes.py:231-255``): tri faces use the 36-point
        Williams-Shunn rule, quad faces (p+1)^2 Gauss-Legendre."""
        p = self.order
        rows = []
        for kind, proj, _ in self.faces:
            if kind == 'tri':
                qp, qw = TABLES[('tri', 'williams-shunn', 36)]
                qp, qw = np.array(qp), np.array(qw)
            else:
                g, w = gauss_legendre(p + 1)
                qp = np.array([(g[i], g[j]) for j in range(p + 1)
                               for i in range(p + 1)])
                qw = np.array([w[i]*w[j] for j in range(p + 1)
                               for i in range(p + 1)])
            L = self.facebases[kind].nodal_basis_at(qp)      # (nq, nfp)
            M = self.ubasis.ortho.at(_proj(proj, qp))        # (nb, nq)
            rows.append(np.einsum('i,ik,ji->kj', qw, L, M))
        return np.vstack(rows)

    def nodal_basis_at(self, pts):
        """(npts, nupts): the solution's nodal basis at reference points."""
        return self.ubasis.nodal_basis_at(pts)

    @cached_property
    def m0(self):
        return self.ubasis.nodal_basis_at(self.fpts)

    @cached_property
    def m1(self):
        m = np.rollaxis(self.ubasis.jac_nodal_basis_at(self.upts), 2)
        return m.reshape(self.nupts, -1)

    @cached_property
    def m2(self):
        m = self.norm_fpts[..., None]*self.m0[:, None, :]
        return m.reshape(self.nfpts, -1)

    @cached_property
    def m3(self):
        m = (self.gbasis_coeffs @ self.ubasis.ortho.at(self.upts)).T
        if self.iqrules:
            # fed by the L2 projection of each face's interface-rule values
            # onto its face basis, proj_l2 (pyfr/shapes.py:103-113, :18-19)
            blocks = []
            for kind, _, _ in self.faces:
                fb = self.facebases[kind]
                qp, qw = self._iq_rule(kind)
                blocks.append(fb.ortho.at(fb.pts).T @ (qw*fb.ortho.at(qp)))
            bd = np.zeros((sum(b.shape[0] for b in blocks),
                           sum(b.shape[1] for b in blocks)))
            r = c = 0
            for b in blocks:
                bd[r:r + b.shape[0], c:c + b.shape[1]] = b
                r, c = r + b.shape[0], c + b.shape[1]
            m = m @ bd
        return m

    @cached_property
    def m4(self):
        m = self.m1.reshape(self.nupts, -1, self.nupts).swapaxes(0, 1)
        return m.reshape(-1, self.nupts)

    @cached_property
    def m6(self):
        m = self.norm_fpts.T[:, None, :]*self.m3
        return m.reshape(-1, self.nfpts)

    # -- flux anti-aliasing (pyfr/shapes.py:125-135, :18-19)
    @cached_property
    def _eqrule(self):
        pts, wts = TABLES[(self.name, *self.qrule)]
        return np.array(pts), np.array(wts)

    @property
    def qpts(self):
        return self._eqrule[0]

    @property
    def nqpts(self):
        return len(self._eqrule[0])

    @cached_property
    def m7(self):
        return self.ubasis.nodal_basis_at(self.qpts)

    @cached_property
    def m8(self):
        """L2 projection: values at the solution points of
        sum_b phi_b(x) sum_q w_q phi_b(x_q) f_q for the orthonormal phi_b."""
        qp, qw = self._eqrule
        ob = self.ubasis.ortho
        return ob.at(self.upts).T @ (qw*ob.at(qp))

    @property
    def m9(self):
        m8, (a, b) = self.m8, self.m8.shape
        out = np.zeros((3*a, 3*b))
        for d in range(3):
            out[d*a:(d + 1)*a, d*b:(d + 1)*b] = m8
        return out

    def opmat(self, expr):
        ids = (0, 1, 2, 3, 4, 6) + ((7, 8, 9) if self.qrule else ())
        env = {f'M{i}': getattr(self, f'm{i}') for i in ids}
        return _flush(eval(expr.replace('*', '@'), {'__builtins__': {}}, env))

    @cached_property
    def upts_wts(self):
        w = self._upts_rule[1]
        if w[0] is not None:
            return np.array(w)
        # point sets without weights: the interpolatory weights, the
        # integrals of the nodal basis functions over the element
        qp, qw = _collapsed_rule(self.name, self.order + 3)
        return self.ubasis.nodal_basis_at(qp).T @ qw


class TetShape(_SimplexShape):
    """``pyfr/shapes.py:476-518``."""
    name = 'tet'
    nverts = 4
    upts_default = 'shunn-ham'
    faces = [
        ('tri', lambda s, t: (s, t, -1.0), (0, 0, -1)),
        ('tri', lambda s, t: (s, -1.0, t), (0, -1, 0)),
        ('tri', lambda s, t: (-1.0, t, s), (-1, 0, 0)),
        ('tri', lambda s, t: (s, t, -s - t - 1), (1, 1, 1)),
    ]

    @staticmethod
    def _nupts_of(p):
        return (p + 1)*(p + 2)*(p + 3)//6

    @cached_property
    def _upts_rule(self):
        return TABLES[('tet', self.upts_name, self._nupts_of(self.order))]

    @cached_property
    def upts(self):
        return np.array(self._upts_rule[0])

    @staticmethod
    def std_ele(sptord):
        """Equispaced points (``TetShape.std_ele``, ``shapes.py:502-510``)."""
        pts1d = np.linspace(-1, 1, sptord + 1)
        return np.array([(p, q, r)
                         for i, r in enumerate(pts1d)
                         for j, q in enumerate(pts1d[:(sptord + 1 - i)])
                         for p in pts1d[:(sptord + 1 - i - j)]])


class PriShape(_SimplexShape):
    """``pyfr/shapes.py:521-570``."""
    name = 'pri'
    nverts = 6
    upts_default = 'williams-shunn~gauss-legendre'
    faces = [
        ('tri', lambda s, t: (s, t, -1.0), (0, 0, -1)),
        ('tri', lambda s, t: (s, t, 1.0), (0, 0, 1)),
        ('quad', lambda s, t: (s, -1.0, t), (0, -1, 0)),
        ('quad', lambda s, t: (-s, s, t), (1, 1, 0)),
        ('quad', lambda s, t: (-1.0, s, t), (-1, 0, 0)),
    ]

    @staticmethod
    def _nupts_of(p):
        return (p + 1)**2*(p + 2)//2

    @cached_property
    def _upts_rule(self):
        return TABLES[('pri', self.upts_name, self._nupts_of(self.order))]

    @cached_property
    def upts(self):
        return np.array(self._upts_rule[0])

    @staticmethod
    def std_ele(sptord):
        """``PriShape.std_ele`` (``shapes.py:555-562``)."""
        pts1d = np.linspace(-1, 1, sptord + 1)
        return np.array([(p, q, r)
                         for r in pts1d
                         for i, q in enumerate(pts1d)
                         for p in pts1d[:(sptord + 1 - i)]])


class PyrShape(_SimplexShape):
    """``pyfr/shapes.py:573-640``: a square base on z = -1 and the apex at
    (0, 0, 1); four triangular faces whose unnormalised reference normals
    have z component 1/2."""
    name = 'pyr'
    nverts = 5
    upts_default = 'gauss-legendre'
    faces = [
        ('quad', lambda s, t: (s, t, -1.0), (0, 0, -1)),
        ('tri', lambda s, t: (s + (t + 1)/2, (t - 1)/2, t), (0, -1, 0.5)),
        ('tri', lambda s, t: ((1 - t)/2, -s - (t + 1)/2, t), (1, 0, 0.5)),
        ('tri', lambda s, t: (-s - (t + 1)/2, (1 - t)/2, t), (0, 1, 0.5)),
        ('tri', lambda s, t: ((t - 1)/2, s + (t + 1)/2, t), (-1, 0, 0.5)),
    ]

    @staticmethod
    def _nupts_of(p):
        return (p + 1)*(p + 2)*(2*p + 3)//6

    @cached_property
    def _upts_rule(self):
        return TABLES[('pyr', self.upts_name, self._nupts_of(self.order))]

    @cached_property
    def upts(self):
        return np.array(self._upts_rule[0])

    @staticmethod
    def std_ele(sptord):
        """``PyrShape.std_ele`` (``shapes.py:619-627``): layer i at
        z = -1 + 2i/k holds (k + 1 - i)^2 points, x fastest, at every other
        one of the 2k + 1 equispaced points of [-1, 1] from the i-th."""
        k = sptord
        t = np.linspace(-1, 1, 2*k + 1)
        return np.array([(x, y, z) for i, z in enumerate(t[::2])
                         for y in t[i:2*k + 1 - i:2]
                         for x in t[i:2*k + 1 - i:2]])


SIMPLEX_SHAPES = {'tet': TetShape, 'pri': PriShape, 'pyr': PyrShape}
