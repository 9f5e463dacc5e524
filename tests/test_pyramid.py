"""Pyramid operators, geometry and meshes (CPU).

The oracle restates the reference's pyramid basis (``PyrPolyBasis``,
``pyfr/polys.py:332-392``) and ``PyrShape`` (``pyfr/shapes.py:573-640``).
The reference holds no pyramid fixture, so the oracle is pinned by what the
mathematics fixes, each check built from numpy's Gauss-Legendre rules
collapsed onto the pyramid ((x, y, z) = (a s, b s, c), s = (1 - c)/2,
Jacobian s^2), independent of both the oracle and the product:

* the solution point sets are the layered Gauss-Legendre sets they are
  named for (layer i at the i-th Gauss node in z holding the tensor
  (n - i)-point rule scaled by (1 - z)/2);
* the oracle basis is orthonormal, its Jacobian matches finite differences,
  and its order-1 nodal basis on ``std_ele(1)`` is the reference's closed
  form ``PyrShape.interp_expr``;
* M0 / M4 reproduce degree-p polynomials and their gradients, the
  divergence operators are exact for a polynomial flux, and M3 / M6 satisfy
  the lifting identity;
* a planted mistake (a wrong Jacobi weight) fails the orthonormality check.

The product spans the space with homogenised Legendre products; it agrees
with the oracle to 1e-12 at p <= 3 and 1e-10 at p = 4, where the
reference's Jacobi form itself loses accuracy on the space's high-degree
members (shown against an exact member of the space).
"""
import numpy as np
import pytest

import oracle.basis as ob
from pyfr_amd.quadtables import TABLES
from pyfr_amd.simplex import PyrShape

OPS = ['M0', 'M1 - M3*M2', 'M3', 'M4 - M6*M0', 'M6']


def _rule(n):
    """Collapsed tensor Gauss-Legendre rule on the reference pyramid,
    exact for degree 2n - 1 in x / s, y / s and, after the s^2 Jacobian,
    in z."""
    g, w = np.polynomial.legendre.leggauss(n)
    a, b, c = (x.ravel() for x in np.meshgrid(g, g, g, indexing='ij'))
    s = (1 - c)/2
    return (np.stack([a*s, b*s, c], 1),
            np.einsum('i,j,k->ijk', w, w, w).ravel()*s**2)


def _tables(p):
    n = p + 1
    return {'upts': TABLES[('pyr', 'gauss-legendre', n*(n + 1)*(2*n + 1)//6)],
            'tri_fpts': TABLES[('tri', 'williams-shunn',
                                (p + 1)*(p + 2)//2)][0],
            'tri_q36': TABLES[('tri', 'williams-shunn', 36)]}


def _oracle(p):
    return ob.SimplexShapeOracle('pyr', p, _tables(p))


def _randpoly(rng, p):
    """A random polynomial of total degree p (inside the pyramid space) and
    its gradient."""
    exps = [(i, j, k) for i in range(p + 1) for j in range(p + 1 - i)
            for k in range(p + 1 - i - j)]
    c = rng.standard_normal(len(exps))

    def f(x):
        return sum(ci*x[:, 0]**i*x[:, 1]**j*x[:, 2]**k
                   for ci, (i, j, k) in zip(c, exps))

    def grad(x):
        g = np.zeros_like(x)
        for ci, (i, j, k) in zip(c, exps):
            if i:
                g[:, 0] += ci*i*x[:, 0]**(i - 1)*x[:, 1]**j*x[:, 2]**k
            if j:
                g[:, 1] += ci*j*x[:, 0]**i*x[:, 1]**(j - 1)*x[:, 2]**k
            if k:
                g[:, 2] += ci*k*x[:, 0]**i*x[:, 1]**j*x[:, 2]**(k - 1)
        return g
    return f, grad


@pytest.mark.parametrize('n', [2, 3, 4, 5])
def test_solution_points_are_layered_gauss_legendre(n):
    pts = np.array(TABLES[('pyr', 'gauss-legendre',
                           n*(n + 1)*(2*n + 1)//6)][0])
    gz = np.polynomial.legendre.leggauss(n)[0]
    want = []
    for i, z in enumerate(gz):
        g = np.polynomial.legendre.leggauss(n - i)[0]
        want += [(x*(1 - z)/2, y*(1 - z)/2, z) for y in g for x in g]
    assert np.abs(pts - np.array(want)).max() < 1e-15


@pytest.mark.parametrize('p', [1, 2, 3, 4, 5])
def test_oracle_basis_orthonormal(p):
    qp, qw = _rule(2*p + 3)
    phi = ob.PyrBasis(p, qp).ortho_basis_at(qp, do_clean=False)
    assert len(phi) == (p + 1)*(p + 2)*(2*p + 3)//6
    g = (phi*qw) @ phi.T
    assert np.abs(g - np.eye(len(g))).max() < 1e-13


def test_oracle_basis_jacobian_finite_difference():
    rng = np.random.default_rng(5)
    qp, _ = _rule(4)
    x = qp[rng.choice(len(qp), 8, replace=False)]
    b = ob.PyrBasis(4, x)
    j = b.jac_ortho_basis_at(x, do_clean=False)
    h = 1e-6
    for d in range(3):
        dx = np.zeros(3)
        dx[d] = h
        fd = (b.ortho_basis_at(x + dx, False) -
              b.ortho_basis_at(x - dx, False))/(2*h)
        assert np.abs(j[d] - fd).max() < 1e-6


def test_planted_jacobi_weight_mistake_is_caught():
    """A wrong weight in the z Jacobi polynomials breaks orthonormality."""
    class Planted(ob.PyrBasis):
        def _ortho(self, p, q, r):
            real = ob.jacobi
            ob.jacobi = lambda n, a, b, z: real(n, a + 1 if a else a, b, z)
            try:
                return super()._ortho(p, q, r)
            finally:
                ob.jacobi = real
    qp, qw = _rule(9)
    phi = Planted(3, qp).ortho_basis_at(qp, do_clean=False)
    g = (phi*qw) @ phi.T
    assert np.abs(g - np.eye(len(g))).max() > 1e-3


def test_linear_nodal_basis_is_reference_interp_expr():
    """The order-1 nodal basis on std_ele(1) is PyrShape.interp_expr
    (pyfr/shapes.py:606-612), the reference's closed form of the linear
    pyramid map; the product's and the oracle's std_ele(1) agree."""
    from oracle.simplex import std_ele
    v = std_ele('pyr', 1)
    assert np.array_equal(v, PyrShape.std_ele(1))
    rng = np.random.default_rng(1)
    x = _rule(3)[0][rng.choice(27, 10, replace=False)]
    a, b, c = x.T
    want = np.stack([2*a*b - 2*a - 2*b - c + 1, -2*a*b + 2*a - 2*b - c + 1,
                     -2*a*b - 2*a + 2*b - c + 1, 2*a*b + 2*a + 2*b - c + 1,
                     4*(1 + c)], 1)/8
    got = ob.PyrBasis(1, v).nodal_basis_at(x)
    assert np.abs(got - want).max() < 1e-14


@pytest.mark.parametrize('k', [1, 2, 3, 4])
def test_std_ele_counts_and_layers(k):
    pts = PyrShape.std_ele(k)
    assert len(pts) == (k + 1)*(k + 2)*(2*k + 3)//6
    for i in range(k + 1):
        lay = pts[np.isclose(pts[:, 2], -1 + 2*i/k)]
        assert len(lay) == (k + 1 - i)**2
        assert np.isclose(np.abs(lay[:, :2]).max(), 1 - i/k)


@pytest.mark.parametrize('p', [1, 2, 3, 4])
def test_oracle_operators_exact(p):
    o = _oracle(p)
    rng = np.random.default_rng(p)
    f, grad = _randpoly(rng, p)
    u = f(o.upts)
    assert np.abs(o.m0 @ u - f(o.fpts)).max() < 1e-11
    g = (o.m4 @ u).reshape(3, -1).T
    assert np.abs(g - grad(o.upts)).max() < 1e-10

    F = np.stack([f(o.upts)*(d + 1) for d in range(3)])
    fn = (o.norm_fpts*f(o.fpts)[:, None]*np.arange(1, 4)).sum(axis=1)
    div = o.opmat('M1 - M3*M2') @ F.reshape(-1) + o.opmat('M3') @ fn
    assert np.abs(div - grad(o.upts) @ np.arange(1, 4)).max() < 1e-10

    # lifting identity int_E g_j v = int_face L_j v, v of degree p (the
    # pyramid space's members reach degree 2p in x and y)
    qp, qw = _rule(2*p + 3)
    lu = o.ubasis.nodal_basis_at(qp)
    v, _ = _randpoly(rng, p)
    lhs = (lu @ o.m3).T @ (qw*v(qp))
    lhs6 = [(lu @ o.m6[d*o.nupts:(d + 1)*o.nupts]).T @ (qw*v(qp))
            for d in range(3)]
    rhs = []
    for kind, proj, _ in o.faces:
        g1, w1 = np.polynomial.legendre.leggauss(p + 2)
        a, b = (x.ravel() for x in np.meshgrid(g1, g1, indexing='ij'))
        if kind == 'quad':
            fq, fw = np.stack([a, b], 1), np.outer(w1, w1).ravel()
        else:
            fq = np.stack([(1 + a)*(1 - b)/2 - 1, b], 1)
            fw = np.outer(w1, w1).ravel()*(1 - b)/2
        L = o._facebasis(kind).nodal_basis_at(fq)
        rhs.append(L.T @ (fw*v(ob._proj(proj, fq))))
    rhs = np.concatenate(rhs)
    assert np.abs(lhs - rhs).max() < 1e-11
    for d in range(3):
        assert np.abs(lhs6[d] - o.norm_fpts[:, d]*rhs).max() < 1e-11


@pytest.mark.parametrize('p', [1, 2, 3, 4])
def test_product_operators_match_oracle(p):
    s, o = PyrShape(p), _oracle(p)
    assert np.array_equal(s.upts, o.upts)
    assert np.abs(s.fpts - o.fpts).max() < 1e-15
    assert np.array_equal(s.norm_fpts, o.norm_fpts)
    tol = 1e-12 if p <= 3 else 1e-10
    for op in OPS:
        a, b = s.opmat(op), o.opmat(op)
        assert a.shape == b.shape
        assert np.abs(a - b).max() < tol*max(1, np.abs(b).max()), op


def test_product_basis_exact_on_high_degree_member():
    """x^4-like members of the p = 4 space (s^4 P_4(x/s) s^4 P_4(y/s),
    s = (1 - z)/2, degree 8) are interpolated exactly by the product's M0
    (1e-14); the reference's Jacobi form is good to about 1e-11 there,
    which bounds the p = 4 agreement above."""
    from numpy.polynomial import legendre as L
    s = PyrShape(4)
    c = L.leg2poly([0, 0, 0, 0, 1])

    def f(x):
        sc = (1 - x[:, 2])/2
        q = [sum(c[m]*x[:, d]**m*sc**(4 - m) for m in range(5))
             for d in (0, 1)]
        return q[0]*q[1] + 0.3*q[0]*x[:, 1]
    assert np.abs(s.m0 @ f(s.upts) - f(s.fpts)).max() < 1e-14


@pytest.mark.parametrize('p', [1, 2, 3, 4])
def test_aa_rule_defaults_and_weights(p):
    """Flux anti-aliasing on pyramids defaults to the fewest-point rule of
    degree >= 2(p + 1) (pyfr/shapes.py:146-176); the interpolatory weights
    of the solution points integrate degree-p polynomials exactly."""
    from pyfr_amd.simplex import AA_RULES, qdeg_rule
    assert AA_RULES['pyr'][p] == qdeg_rule('pyr', 2*(p + 1))
    s = PyrShape(p)
    f, _ = _randpoly(np.random.default_rng(10 + p), p)
    qp, qw = _rule(p + 2)
    assert abs(s.upts_wts @ f(s.upts) - qw @ f(qp)) < 1e-12


@pytest.mark.parametrize('layers', [[('pyr', 2)], [('hex', 1), ('pyr', 1)],
                                    [('tet', 1), ('pyt', 1)],
                                    [('pyt', 1), ('pri', 1)]])
def test_pyramid_meshes(layers):
    """Pyramid box meshes fill the box (volumes from the affine metric)
    and pair every face once (the generator raises otherwise); pairs agree
    with brute-force matching of face centroids."""
    from pyfr_amd.geometry import AffineGeometry
    from pyfr_amd.mesh import FACE_VERTS, SimplexBoxMesh
    m = SimplexBoxMesh([((2, 2, 2), 0.5, layers)])
    vol = 0.0
    refvol = {'tet': 4/3, 'pri': 4.0, 'pyr': 8/3}
    for et, sp in m.spts.items():
        if et == 'hex':
            vol += sp.shape[1]*0.125
        else:
            vol += (AffineGeometry(sp, et).djac*refvol[et]).sum()
    assert abs(vol - 1.0) < 1e-13
    lhs, rhs = m.con

    def cen(t, e, f):
        v = m.spts[t][list(FACE_VERTS[t][f]), e]
        return v.mean(axis=0) % 1.0
    for a, b in zip(lhs, rhs):
        assert np.abs(cen(*a) - cen(*b)).max() < 1e-12
    nf = sum(len(FACE_VERTS[t])*v.shape[1] for t, v in m.spts.items())
    assert 2*len(lhs) == nf


def test_affine_geometry_rejects_skewed_base():
    from pyfr_amd.geometry import AffineGeometry
    v = PyrShape.std_ele(1)[:, None, :].copy()
    AffineGeometry(v, 'pyr')
    v[3, 0, 0] += 0.1
    with pytest.raises(ValueError):
        AffineGeometry(v, 'pyr')


@pytest.mark.parametrize('p', [1, 3])
def test_curved_pyramid_geometry_straight_equals_affine(p):
    """Straight pyramids given as order-2 shape points take the stored
    metric of the curved path; it equals the affine metric."""
    from pyfr_amd.geometry import AffineGeometry, CurvedSimplexGeometry
    from pyfr_amd.mesh import SimplexBoxMesh
    m1 = SimplexBoxMesh([((2, 2, 1), 1.0, [('pyr', 1)])])
    m2 = SimplexBoxMesh([((2, 2, 1), 1.0, [('pyr', 1)])], sorder=2)
    ga = AffineGeometry(m1.spts['pyr'], 'pyr')
    gc = CurvedSimplexGeometry(m2.spts['pyr'], p, 'pyr')
    s = PyrShape(p)
    assert np.abs(gc.ploc_at(s.upts) - ga.ploc_at(s.upts)).max() < 1e-13
    assert np.abs(gc.smat_at(s.upts) - ga.smat_at(s.upts)).max() < 1e-12
    assert np.abs(gc.rcpdjac_at(s.upts) - ga.rcpdjac_at(s.upts)).max() \
        < 1e-12


def _pyr_oracle(n, p, layers, **kw):
    from oracle.simplex import OracleMixedSystem
    from pyfr_amd.mesh import SimplexBoxMesh
    from tests.test_simplex import _tables as st
    L = 2*np.pi
    m = SimplexBoxMesh([((n, n, n), L/n, layers)])
    verts = {et: v.transpose(0, 2, 1) for et, v in m.spts.items()}
    idx = {et: np.arange(v.shape[1]) for et, v in m.spts.items()}
    tabs = {et: _tables(p) if et == 'pyr' else st(et, p) for et in verts}
    return m, OracleMixedSystem(verts, p, tabs, [((L, L, L), idx)], **kw)


@pytest.mark.parametrize('layers', [[('pyr', 2)], [('tet', 1), ('pyt', 1)]])
def test_oracle_pyramid_free_stream(layers):
    """A uniform state stays steady on pyramid and pyramid-tet meshes.
    (Totals are not conserved to rounding on pyramids: the space's members
    restricted to a triangular face reach degree 2p, beyond the face
    basis, so the lifted interior flux differs from its face integral.)"""
    _, o = _pyr_oracle(2, 2, layers)
    fns = {k: (lambda v: lambda x, y, z: v + 0*x)(v) for k, v in
           {'rho': 1.1, 'u': 0.3, 'v': -0.2, 'w': 0.1, 'p': 0.9}.items()}
    r = o.rhs(o.ics(fns))
    assert max(np.abs(a).max() for a in r.values()) < 1e-12


def _density_wave_error(n, p, plant=False):
    U = np.array([0.5, 0.3, -0.2])
    _, o = _pyr_oracle(n, p, [('pyr', n)])
    if plant:
        o.et['pyr'].M3 = -o.et['pyr'].M3
    fns = {'rho': lambda x, y, z: 1 + 0.2*np.sin(x + y + z),
           'u': lambda x, y, z: U[0] + 0*x, 'v': lambda x, y, z: U[1] + 0*x,
           'w': lambda x, y, z: U[2] + 0*x, 'p': lambda x, y, z: 1 + 0*x}
    r = o.rhs(o.ics(fns))['pyr']
    oe = o.et['pyr']
    x, y, z = oe.ploc(oe.sh.upts).transpose(1, 0, 2)
    ex = -U.sum()*0.2*np.cos(x + y + z)
    return np.sqrt(((r[:, 0] - ex)**2).mean())


def test_oracle_pyramid_rhs_converges():
    """The oracle's pyramid Euler RHS for an advected density wave converges
    at about order p (p = 2, n = 2 -> 4); a sign error planted in its
    correction operator does not."""
    e2, e4 = _density_wave_error(2, 2), _density_wave_error(4, 2)
    assert np.log2(e2/e4) > 2 - 0.7, (e2, e4)
    p2, p4 = _density_wave_error(2, 2, True), _density_wave_error(4, 2, True)
    assert np.log2(p2/p4) < 2 - 0.7


@pytest.mark.parametrize('p', [1, 2, 3])
def test_curved_pyramid_geometry_matches_oracle(p):
    """Warped order-2 pyramids: the product's stored metric (its pyramid
    nodal bases) against the oracle's (the reference's Jacobi bases).
    Warped pyramids do not in general meet their neighbours exactly: the
    pyramid space restricted to a triangular face reaches degree 2k, so a
    face's shape is not fixed by the shape points on it alone."""
    from oracle.simplex import _Etype
    from pyfr_amd.geometry import CurvedSimplexGeometry
    from pyfr_amd.mesh import SimplexBoxMesh
    from tests.test_simplex import _warp
    L = 2*np.pi
    m = SimplexBoxMesh([((2, 2, 2), L/2, [('pyr', 2)])], sorder=2,
                       warp=_warp)
    v = m.spts['pyr']
    oe = _Etype('pyr', p, _tables(p), v.transpose(0, 2, 1))
    g = CurvedSimplexGeometry(v, p, 'pyr')
    assert oe.curved
    sh = oe.sh
    assert np.abs(g.ploc_at(sh.upts) - oe.ploc(sh.upts)).max() < 1e-13
    assert np.abs(g.smat_at(sh.upts).transpose(1, 0, 2, 3) -
                  oe.smats).max() < 1e-12
    assert np.abs(g.rcpdjac_at(sh.upts) - oe.rcpdjac).max() < 1e-12
    assert np.abs(g.pnorm_at(sh.fpts, sh.norm_fpts) - oe.pnorm).max() \
        < 1e-12


def test_pyramid_conservation_defect():
    """Totals over a periodic pyramid box: conserved to rounding at p = 1;
    from p = 2 the space's members beyond degree p leave a defect that is
    truncation error, falling with the cell size (measured 1.8e-3 -> 1.2e-5
    from n = 2 to 4 at p = 2)."""
    from tests.fr_cases import eval_ics, mode_ics
    out = {}
    for p, n in ((1, 2), (2, 2), (2, 4)):
        _, o = _pyr_oracle(n, p, [('pyr', n)])
        r = o.rhs(o.ics(eval_ics(mode_ics(seed=4))))['pyr']
        w = PyrShape(p).upts_wts[:, None]/o.et['pyr'].rcpdjac
        out[p, n] = max(abs((w*r[:, v]).sum()) for v in range(5))
    assert out[1, 2] < 1e-12
    assert out[2, 2] > 1e-5 and out[2, 4] < out[2, 2]/64
