"""Metric terms of hex elements (setup, host side).

Restates the reference's metric pipeline for hexes of any shape order (the
trilinear map of 8 vertices, or curved elements with (k + 1)^3 shape
points):

* physical locations by the shape-point basis (``elements.py:331-350``);
* S = |J| J^-1 at pseudo-grid points in Kopriva's invariant curl form and |J|
  (``pyfr/solvers/base/elements.py:382-440``), interpolated to any point set
  with the order-max(p, 1) basis on equispaced pseudo-grid points
  (``elements.py:276-307``);
* physical normals S^T n (``elements.py:364-380``).

The bases are tensor-product Lagrange polynomials on equispaced points,
evaluated directly rather than through an orthonormal Vandermonde inverse.
"""
import numpy as np


def _lagrange_1d(x, pts):
    """Values and derivatives of the Lagrange basis through x at pts."""
    n = len(x)
    L = np.ones((len(pts), n))
    dL = np.zeros((len(pts), n))
    for j in range(n):
        for m in range(n):
            if m == j:
                continue
            L[:, j] *= (pts - x[m])/(x[j] - x[m])
        for m in range(n):
            if m == j:
                continue
            t = np.full(len(pts), 1/(x[j] - x[m]))
            for q in range(n):
                if q != j and q != m:
                    t = t*(pts - x[q])/(x[j] - x[q])
            dL[:, j] += t
    return L, dL


class TensorLagrange:
    """Nodal basis on the (n1)^3 tensor grid of 1-D points x1, x fastest."""

    def __init__(self, x1):
        self.x1 = np.asarray(x1, dtype=float)
        self.n1 = len(x1)
        n = self.n1
        self.pts = np.array([(x1[i], x1[j], x1[k]) for k in range(n)
                             for j in range(n) for i in range(n)])

    def nodal_basis_at(self, pts):
        """(npts, nnodes)"""
        pts = np.atleast_2d(pts)
        Lx, _ = _lagrange_1d(self.x1, pts[:, 0])
        Ly, _ = _lagrange_1d(self.x1, pts[:, 1])
        Lz, _ = _lagrange_1d(self.x1, pts[:, 2])
        return np.einsum('pk,pj,pi->pkji', Lz, Ly, Lx).reshape(len(pts), -1)

    def jac_nodal_basis_at(self, pts):
        """(3, npts, nnodes): derivative along each reference axis."""
        pts = np.atleast_2d(pts)
        Lx, dLx = _lagrange_1d(self.x1, pts[:, 0])
        Ly, dLy = _lagrange_1d(self.x1, pts[:, 1])
        Lz, dLz = _lagrange_1d(self.x1, pts[:, 2])
        np_ = len(pts)
        d0 = np.einsum('pk,pj,pi->pkji', Lz, Ly, dLx).reshape(np_, -1)
        d1 = np.einsum('pk,pj,pi->pkji', Lz, dLy, Lx).reshape(np_, -1)
        d2 = np.einsum('pk,pj,pi->pkji', dLz, Ly, Lx).reshape(np_, -1)
        return np.stack([d0, d1, d2])


def hex_shape_order(nspts):
    """Shape order k of a hex with (k + 1)^3 shape points."""
    k = round(nspts ** (1/3)) - 1
    if k < 1 or (k + 1)**3 != nspts:
        raise ValueError(f'Invalid hex shape-point count {nspts}')
    return k


def hex_corner_index(k):
    """Indices of the 8 vertices (``HexShape`` order) among the order-k
    shape points (x fastest)."""
    n = k + 1
    return [((v & 1)*k) + n*((((v >> 1) & 1)*k) + n*(((v >> 2) & 1)*k))
            for v in range(8)]


def shape_basis():
    """Trilinear shape-point basis of a linear hex: nodes are the vertices
    in ``HexShape`` order (x fastest)."""
    return TensorLagrange(np.array([-1.0, 1.0]))


class HexGeometry:
    """Metric terms for hex elements with vertex array eles (8, neles, 3)."""

    def __init__(self, eles, order):
        self.eles = np.asarray(eles, dtype=float)
        self.nspts, self.neles, self.ndims = self.eles.shape
        self.sorder = hex_shape_order(self.nspts)
        self.order = order
        # shape basis on the order-k equispaced shape points (pyfr/shapes.py:
        # 49-50, :331-333); k = 1 is the trilinear map of the vertices
        self.sbasis = TensorLagrange(np.linspace(-1, 1, self.sorder + 1))
        mord = max(order, 1)
        self.mbasis = TensorLagrange(np.linspace(-1, 1, mord + 1))
        self._smats_djacs = None

    def ploc_at(self, pts):
        """Physical locations (npts, 3, neles)."""
        op = self.sbasis.nodal_basis_at(pts)
        ploc = op @ self.eles.reshape(self.nspts, -1)
        return ploc.reshape(len(pts), -1, self.ndims).swapaxes(1, 2)

    def smats_djacs_mpts(self, chunk=16384):
        if self._smats_djacs is not None:
            return self._smats_djacs
        mb = self.mbasis
        mpts, nm, ne = mb.pts, len(mb.pts), self.neles
        x = self.ploc_at(mpts)                               # (nm, 3, E)
        # jacop[(a, p), q] = d l_q / d xi_a at pseudo grid point p
        jacop = mb.jac_nodal_basis_at(mpts).reshape(3*nm, nm)
        smats = np.empty((3, nm, 3, ne))
        djacs = np.empty((nm, ne))
        for s0 in range(0, ne, chunk):
            s1 = min(ne, s0 + chunk)
            xc = np.ascontiguousarray(x[:, :, s0:s1])
            n = s1 - s0
            # jac[a, p, b, e] = dX_b/dxi_a
            jac = (jacop @ xc.reshape(nm, -1)).reshape(3, nm, 3, n)
            dtt = []
            for a in range(3):
                tt = np.cross(xc, jac[a], axisa=1, axisb=1, axisc=1)
                # dtt[a][c, p, k, e]: derivative along c of (X x X_a)_k
                dtt.append((jacop @ tt.reshape(nm, -1)).reshape(3, nm, 3, n))
            smats[0, :, :, s0:s1] = 0.5*(dtt[2][1] - dtt[1][2])
            smats[1, :, :, s0:s1] = 0.5*(dtt[0][2] - dtt[2][0])
            smats[2, :, :, s0:s1] = 0.5*(dtt[1][0] - dtt[0][1])
            x1cx2 = np.cross(jac[1], jac[2], axisa=1, axisb=1, axisc=1)
            djacs[:, s0:s1] = np.einsum('pbe,pbe->pe', jac[0], x1cx2)
        self._smats_djacs = (smats, djacs)
        return smats, djacs

    def smat_at(self, pts):
        """(3, npts, 3, neles): S[i][k] at each point."""
        smats, _ = self.smats_djacs_mpts()
        m0 = self.mbasis.nodal_basis_at(pts)
        nm, ne = smats.shape[1], smats.shape[3]
        return np.stack([(m0 @ smats[i].reshape(nm, -1)).reshape(-1, 3, ne)
                         for i in range(3)])

    def rcpdjac_at(self, pts):
        _, djacs = self.smats_djacs_mpts()
        djac = self.mbasis.nodal_basis_at(pts) @ djacs
        if np.any(djac < -1e-5):
            raise RuntimeError('Negative mesh Jacobians detected')
        return 1.0/djac

    def pnorm_at(self, pts, norms):
        """Physical normals (npts, neles, 3)."""
        s = self.smat_at(pts)                             # (3, np, 3, E)
        pn = np.zeros((len(pts), self.neles, 3))
        for i in range(3):
            pn += (norms[:, i, None, None]*s[i]).transpose(0, 2, 1)
        if np.any(np.sqrt(np.einsum('...i,...i', pn, pn)) < 1e-10):
            raise RuntimeError('Zero face normals detected')
        return pn


class AffineGeometry:
    """Metric terms of straight-sided tetrahedra, right prisms and
    pyramids with parallelogram bases, whose reference-to-physical maps are
    affine: the reference's metric pipeline
    (``elements.py:331-440``) then yields one S = |J| J^-1 and |J| per
    element at every point.

    eles: (nverts, neles, 3) in ``std_ele(1)`` vertex order; vertices 1, 2,
    3 are the neighbours of vertex 0 along the three reference axes for
    tets and prisms, so dX/dxi_a = (v_{a+1} - v_0)/2.
    """

    def __init__(self, eles, name):
        v = np.asarray(eles, dtype=float)
        self.name = name
        self.v0 = v[0]                                   # (E, 3)
        if name == 'pyr':
            # vertices 1, 2 neighbour vertex 0 on the base, 3 is opposite
            # it and 4 is the apex: dX/dxi = (v1 - v0)/2, (v2 - v0)/2 and
            # (v4 - base centre)/2 when the base is a parallelogram
            # (PyrShape.jac_exprs, pyfr/shapes.py:587-604)
            if np.abs(v[3] - v[1] - v[2] + v[0]).max() > 1e-10:
                raise ValueError('pyramids must have parallelogram bases')
            j = np.stack([(v[1] - v[0])/2, (v[2] - v[0])/2,
                          (v[4] - (v[1] + v[2])/2)/2], axis=1)
        else:
            j = np.stack([(v[a + 1] - v[0])/2 for a in range(3)], axis=1)
        if name == 'pri':
            # a right prism: the top triangle is the bottom one translated
            if (np.abs(v[4] - v[3] - (v[1] - v[0])).max() > 1e-12 or
                    np.abs(v[5] - v[3] - (v[2] - v[0])).max() > 1e-12):
                raise ValueError('prisms must be affine (right prisms)')
        elif name not in ('tet', 'pyr'):
            raise ValueError(f'no affine map for {name}')
        self.jac = j                                     # (E, a, b)
        s = np.empty_like(j)
        for i, (a, b) in enumerate([(1, 2), (2, 0), (0, 1)]):
            s[:, i] = np.cross(j[:, a], j[:, b])
        self.smats = s                                   # (E, i, k)
        self.djac = np.einsum('ek,ek->e', j[:, 0], s[:, 0])
        if np.any(self.djac <= 0):
            raise ValueError('Negative mesh Jacobians detected')

    def ploc_at(self, pts):
        """(npts, 3, E)"""
        pts = np.atleast_2d(pts)
        x = self.v0.T[None] + np.einsum('pa,eab->pbe', pts + 1, self.jac)
        return x

    def smat_at(self, pts):
        """(3, npts, 3, E): S[i][k] at every point."""
        n = len(np.atleast_2d(pts))
        return np.broadcast_to(self.smats.transpose(1, 2, 0)[:, None],
                               (3, n, 3, len(self.djac))).copy()

    def rcpdjac_at(self, pts):
        """(npts, E)"""
        n = len(np.atleast_2d(pts))
        return np.broadcast_to(1/self.djac, (n, len(self.djac))).copy()

    def pnorm_at(self, pts, norm):
        """Physical normals S^T n (``elements.py:364-380``): (npts, E, 3)."""
        return np.einsum('pi,eik->pek', np.asarray(norm, dtype=float),
                         self.smats)


class SimplexLagrange:
    """Nodal basis of the order-k tet / prism space through points pts,
    with the layouts of ``TensorLagrange``."""

    def __init__(self, kind, k, pts):
        from pyfr_amd.simplex import NodalBasis
        self.pts = np.asarray(pts, dtype=float)
        self._nb = NodalBasis(kind, k, self.pts)

    def nodal_basis_at(self, pts):
        """(npts, nnodes)"""
        return self._nb.nodal_basis_at(np.atleast_2d(pts))

    def jac_nodal_basis_at(self, pts):
        """(3, npts, nnodes)"""
        return self._nb.jac_nodal_basis_at(np.atleast_2d(pts)).swapaxes(1, 2)


def simplex_shape_order(name, nspts):
    """Shape order k of a tet ((k+1)(k+2)(k+3)/6 points), prism
    ((k+1)^2(k+2)/2 points) or pyramid ((k+1)(k+2)(2k+3)/6 points)."""
    for k in range(1, 16):
        n = {'tet': (k + 1)*(k + 2)*(k + 3)//6,
             'pri': (k + 1)**2*(k + 2)//2,
             'pyr': (k + 1)*(k + 2)*(2*k + 3)//6}[name]
        if n == nspts:
            return k
    raise ValueError(f'Invalid {name} shape point count {nspts}')


class CurvedSimplexGeometry(HexGeometry):
    """Metric terms of curved tets and prisms with order-k shape points
    (``std_ele(k)`` order): the same pipeline as for hexes, with the shape
    basis the order-k nodal basis on the shape points and the pseudo-grid
    basis the order-max(p, 1) one on ``std_ele(max(p, 1))``
    (``pyfr/shapes.py:49-61``, ``pyfr/solvers/base/elements.py:382-440``)."""

    def __init__(self, eles, order, name):
        from pyfr_amd.simplex import SIMPLEX_SHAPES
        self.eles = np.asarray(eles, dtype=float)
        self.nspts, self.neles, self.ndims = self.eles.shape
        self.name = name
        self.sorder = simplex_shape_order(name, self.nspts)
        self.order = order
        shape = SIMPLEX_SHAPES[name]
        self.sbasis = SimplexLagrange(name, self.sorder,
                                      shape.std_ele(self.sorder))
        mord = max(order, 1)
        self.mbasis = SimplexLagrange(name, mord, shape.std_ele(mord))
        self._smats_djacs = None
