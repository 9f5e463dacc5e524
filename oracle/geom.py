"""Oracle: metric terms for hex elements (test infrastructure).

* ``linear_smats_djac`` restates the on-the-fly metric of a linear element,
  ``pyfr/solvers/baseadvec/kernels/smats.mako:13-25`` with
  ``HexShape.jac_exprs`` (``pyfr/shapes.py:388-422``).
* ``kopriva_smats_djacs`` restates the stored (curved-element) metric,
  ``pyfr/solvers/base/elements.py:382-440``, used for p < 2 and for the
  physical normals and 1/|J| of every element (``:309-328``, ``:364-380``).
* ``hex_ploc`` maps reference points through shape points of any order
  (curved elements), the reference's ``sbasis``.
"""
import itertools as it

import numpy as np

from oracle.basis import HexBasis


def hex_std_ele(k):
    """Shape points of an order-k hex, ``TensorProdShape.std_ele``
    (``pyfr/shapes.py:331-333``): equispaced, x fastest."""
    pts1d = np.linspace(-1, 1, k + 1)
    return np.array([p[::-1] for p in it.product(pts1d, repeat=3)])


def hex_linspts():
    """Vertices of the reference hex, ``TensorProdShape.std_ele(1)``
    (``pyfr/shapes.py:278-282``)."""
    return hex_std_ele(1)


def hex_sorder(nspts):
    """Shape order k of a hex with nspts = (k + 1)^3 nodes."""
    k = round(nspts ** (1/3)) - 1
    if (k + 1)**3 != nspts:
        raise ValueError(f'Invalid hex node count {nspts}')
    return k


def hex_corners(spts):
    """The 8 vertices (in ``HexShape`` order) of order-k shape points
    (nspts, 3, E)."""
    k = hex_sorder(spts.shape[0])
    n = k + 1
    idx = [((v & 1)*k) + n*((((v >> 1) & 1)*k) + n*(((v >> 2) & 1)*k))
           for v in range(8)]
    return spts[idx]


def hex_ploc(spts, x):
    """Physical location of reference points x for shape points spts
    (nspts, 3, E) of any order: the shape basis ``sbasis`` of
    ``pyfr/shapes.py:49-50`` (here the oracle's orthonormal-basis nodal
    form), through ``ploc_at`` (``pyfr/solvers/base/elements.py:337-350``).
    -> (npts, 3, E)."""
    if spts.shape[0] == 8:
        return trilinear_ploc(spts, x)
    k = hex_sorder(spts.shape[0])
    sb = HexBasis(k, hex_std_ele(k))
    op = sb.nodal_basis_at(np.asarray(x))
    ne = spts.shape[-1]
    return (op @ spts.reshape(spts.shape[0], -1)).reshape(len(x), 3, ne)


# Sign of each reference coordinate at each of the 8 vertices
_VSGN = hex_linspts()


def hex_jac(verts, x):
    """j[a][b] = dX_b/dxi_a of the trilinear map at reference points x.

    verts: (8, 3, E); x: (npts, 3) -> (npts, 3, 3, E).  Same polynomial as
    ``HexShape.jac_exprs`` (``pyfr/shapes.py:388-422``) in factored form.
    """
    npts = len(x)
    out = np.zeros((npts, 3, 3, verts.shape[-1]))
    for v in range(8):
        s = _VSGN[v]
        f = 1 + s[None, :]*x                          # (npts, 3)
        dn = np.empty((npts, 3))
        dn[:, 0] = s[0]*f[:, 1]*f[:, 2]/8
        dn[:, 1] = s[1]*f[:, 0]*f[:, 2]/8
        dn[:, 2] = s[2]*f[:, 0]*f[:, 1]/8
        out += dn[:, :, None, None]*verts[v][None, None, :, :]
    return out


def linear_smats_djac(verts, x):
    """S (npts, 3, 3, E) with s[i] = j[(i+1)%3] x j[(i+2)%3] and |J|;
    ``smats.mako:13-25``."""
    j = hex_jac(verts, x)
    s = np.empty_like(j)
    for i, (a, b) in enumerate([(1, 2), (2, 0), (0, 1)]):
        s[:, i, 0] = j[:, a, 1]*j[:, b, 2] - j[:, a, 2]*j[:, b, 1]
        s[:, i, 1] = j[:, a, 2]*j[:, b, 0] - j[:, a, 0]*j[:, b, 2]
        s[:, i, 2] = j[:, a, 0]*j[:, b, 1] - j[:, a, 1]*j[:, b, 0]
    d = j[:, 0, 0]*s[:, 0, 0] + j[:, 0, 1]*s[:, 0, 1] + j[:, 0, 2]*s[:, 0, 2]
    return s, d


def trilinear_ploc(verts, x):
    """Physical location of reference points x; ``HexShape.interp_expr``.

    verts: (8, 3, E) -> (npts, 3, E)."""
    out = 0
    for v in range(8):
        s = _VSGN[v]
        w = np.prod(1 + s[None, :]*x, axis=1)/8
        out = out + w[:, None, None]*verts[v][None]
    return out


def kopriva_smats_djacs(verts, order):
    """Metric at the pseudo-grid points ``mpts`` with Kopriva's invariant
    curl form; ``pyfr/solvers/base/elements.py:382-440``.

    Returns (smats (3, nmpts, 3, E), djacs (nmpts, E), mbasis, mpts).
    """
    mord = max(order, 1)
    pts1d = np.linspace(-1, 1, mord + 1)
    mpts = np.array([p[::-1] for p in it.product(pts1d, repeat=3)])
    mbasis = HexBasis(mord, mpts)

    # Physical location of the pseudo grid points (nmpts, 3, E)
    x = hex_ploc(verts, mpts)
    smats, djacs = kopriva_metric(x, mbasis, mpts)
    return smats, djacs, mbasis, mpts


def kopriva_metric(x, mbasis, mpts):
    """The invariant curl form of S and |J| at the pseudo-grid points mpts
    from their physical locations x (nmpts, 3, E), with the nodal basis
    mbasis on mpts (any element type; ``elements.py:382-440``).
    Returns (smats (3, nmpts, 3, E), djacs (nmpts, E))."""
    nmpts, neles = len(mpts), x.shape[-1]
    jacop = np.rollaxis(mbasis.jac_nodal_basis_at(mpts), 2).reshape(-1, nmpts)
    jac = (jacop @ x.reshape(nmpts, -1)).reshape(nmpts, 3, 3, neles)
    jac = jac.transpose(1, 2, 0, 3)

    dtt = []
    for dx in jac:
        tt = np.cross(x, dx, axisa=1, axisb=0, axisc=1)
        dt = jacop @ tt.reshape(nmpts, -1)
        dtt.append(dt.reshape(nmpts, 3, 3, -1).swapaxes(0, 1))

    smats = np.empty((3, nmpts, 3, neles))
    smats[0] = 0.5*(dtt[2][1] - dtt[1][2])
    smats[1] = 0.5*(dtt[0][2] - dtt[2][0])
    smats[2] = 0.5*(dtt[1][0] - dtt[0][1])

    x0, x1, x2 = jac
    x1cx2 = np.cross(x1, x2, axisa=0, axisb=0, axisc=1)
    djacs = np.einsum('ij...,ji...->j...', x0, x1cx2)
    return smats, djacs


def smats_at(verts, order, pts):
    """Stored metric interpolated to pts: (3, npts, 3, E);
    ``elements.py:276-286``."""
    smats, _, mbasis, _ = kopriva_smats_djacs(verts, order)
    m0 = mbasis.nodal_basis_at(pts)
    return np.einsum('pm,imje->ipje', m0, smats)


def rcpdjac_at(verts, order, pts):
    """1/|J| interpolated from the pseudo-grid points; ``elements.py:297-307``."""
    _, djacs, mbasis, _ = kopriva_smats_djacs(verts, order)
    djac = mbasis.nodal_basis_at(pts) @ djacs
    if np.any(djac < -1e-5):
        raise RuntimeError('Negative mesh Jacobians detected')
    return 1.0/djac


def pnorm_at(verts, order, pts, norm):
    """Physical normals S^T n; ``elements.py:364-380``. -> (npts, E, 3)."""
    smats = smats_at(verts, order, pts).transpose(1, 3, 0, 2)
    pnorm = np.einsum('ijlk,il->ijk', smats, norm)
    if np.any(np.sqrt(np.einsum('...i,...i', pnorm, pnorm)) < 1e-10):
        raise RuntimeError('Zero face normals detected')
    return pnorm
