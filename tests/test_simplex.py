"""Tetrahedron / prism operators (CPU).

The oracle restates the reference's Jacobi-polynomial bases
(``pyfr/polys.py:117-338``) and ``TetShape`` / ``PriShape``
(``pyfr/shapes.py:476-570``).  The reference holds no tet/prism fixture, so
the oracle is pinned here by what the mathematics fixes, each check built
from numpy's Gauss-Legendre rules collapsed onto the simplex (independent of
both the oracle and the product):

* the tabulated point sets integrate monomials to their stated degree;
* the bases are orthonormal and their Jacobians match finite differences;
* M0 / M1 / M4 reproduce interpolants and gradients of degree-p polynomials;
* M3 and M6 satisfy the lifting identity  int_E g_j v = int_face L_j v  for
  every v of degree p (this is what the correction functions are);
* a planted mistake (a wrong Jacobi weight) fails the orthonormality check.

The product builds its bases independently (Legendre products orthonormalised
on the same collapsed rules) and must agree with the oracle to rounding.
"""
import itertools as it

import numpy as np
import pytest

import oracle.basis as ob
from pyfr_amd.quadtables import TABLES
from pyfr_amd.simplex import PriShape, TetShape

OPS = ['M0', 'M1 - M3*M2', 'M3', 'M4 - M6*M0', 'M6']


def _gl(n):
    return np.polynomial.legendre.leggauss(n)


def _rule(kind, n):
    """Collapsed tensor Gauss-Legendre rule on the reference element."""
    g, w = _gl(n)
    if kind == 'tri':
        a, b = (x.ravel() for x in np.meshgrid(g, g, indexing='ij'))
        wt = np.outer(w, w).ravel()*(1 - b)/2
        return np.stack([(1 + a)*(1 - b)/2 - 1, b], 1), wt
    if kind == 'quad':
        a, b = (x.ravel() for x in np.meshgrid(g, g, indexing='ij'))
        return np.stack([a, b], 1), np.outer(w, w).ravel()
    if kind == 'tet':
        a, b, c = (x.ravel() for x in np.meshgrid(g, g, g, indexing='ij'))
        wt = np.einsum('i,j,k->ijk', w, w, w).ravel()
        wt = wt*(1 - b)*(1 - c)**2/8
        y = (1 + b)*(1 - c)/2 - 1
        x = (1 + a)*(1 - b)*(1 - c)/4 - 1
        return np.stack([x, y, c], 1), wt
    if kind == 'pri':
        tp, tw = _rule('tri', n)
        pts = np.array([(*p, z) for z in g for p in tp])
        return pts, np.outer(w, tw).ravel()
    raise ValueError(kind)


def _tables(name, p):
    nu = (p + 1)*(p + 2)*(p + 3)//6 if name == 'tet' else (p + 1)**2*(p + 2)//2
    key = (('tet', 'shunn-ham', nu) if name == 'tet' else
           ('pri', 'williams-shunn~gauss-legendre', nu))
    return {'upts': TABLES[key],
            'tri_fpts': TABLES[('tri', 'williams-shunn',
                                (p + 1)*(p + 2)//2)][0],
            'tri_q36': TABLES[('tri', 'williams-shunn', 36)]}


def _oracle(name, p):
    return ob.SimplexShapeOracle(name, p, _tables(name, p))


def _randpoly(rng, ndim, deg, kind):
    """A random polynomial of the order-`deg` space of `kind` and its
    gradient, as callables of an (n, ndim) array."""
    if kind == 'pri':
        exps = [(i, j, k) for i in range(deg + 1) for j in range(deg + 1 - i)
                for k in range(deg + 1)]
    else:
        exps = [e for e in it.product(range(deg + 1), repeat=ndim)
                if sum(e) <= deg]
    c = rng.standard_normal(len(exps))

    def f(x):
        return sum(ci*np.prod([x[:, d]**e[d] for d in range(ndim)], axis=0)
                   for ci, e in zip(c, exps))

    def grad(x):
        out = np.zeros_like(x)
        for ci, e in zip(c, exps):
            for d in range(ndim):
                if e[d]:
                    t = ci*e[d]*x[:, d]**(e[d] - 1)
                    for dd in range(ndim):
                        if dd != d:
                            t = t*x[:, dd]**e[dd]
                    out[:, d] += t
        return out
    return f, grad


# Degrees stated in the reference's rule-file names (d2, d3, ...)
@pytest.mark.parametrize('key,deg', [
    (('tet', 'shunn-ham', 4), 2), (('tet', 'shunn-ham', 10), 3),
    (('tet', 'shunn-ham', 20), 5), (('tet', 'shunn-ham', 35), 6),
    (('tri', 'williams-shunn', 3), 2), (('tri', 'williams-shunn', 6), 4),
    (('tri', 'williams-shunn', 10), 5), (('tri', 'williams-shunn', 15), 7),
    (('tri', 'williams-shunn', 36), 12),
    (('pri', 'williams-shunn~gauss-legendre', 6), 2),
    (('pri', 'williams-shunn~gauss-legendre', 18), 4),
    (('pri', 'williams-shunn~gauss-legendre', 40), 5),
    (('pri', 'williams-shunn~gauss-legendre', 75), 7)])
def test_quadrature_tables_exact_to_stated_degree(key, deg):
    pts, wts = (np.array(x, dtype=float) for x in TABLES[key])
    qp, qw = _rule(key[0], deg + 2)
    nd = pts.shape[1]
    for e in it.product(range(deg + 1), repeat=nd):
        if key[0] == 'pri':
            if e[0] + e[1] > deg or e[2] > deg:
                continue
        elif sum(e) > deg:
            continue
        f = lambda x: np.prod([x[:, d]**e[d] for d in range(nd)], axis=0)
        assert abs(wts @ f(pts) - qw @ f(qp)) < 1e-13, e


@pytest.mark.parametrize('cls,kind,nd', [
    (ob.TriBasis, 'tri', 2), (ob.TetBasis, 'tet', 3), (ob.PriBasis, 'pri', 3)])
@pytest.mark.parametrize('p', [1, 2, 3, 4])
def test_oracle_basis_orthonormal(cls, kind, nd, p):
    qp, qw = _rule(kind, p + 3)
    phi = cls(p, qp).ortho_basis_at(qp, do_clean=False)
    g = (phi*qw) @ phi.T
    assert np.abs(g - np.eye(len(g))).max() < 1e-13


@pytest.mark.parametrize('cls,kind', [(ob.TetBasis, 'tet'),
                                      (ob.PriBasis, 'pri')])
def test_oracle_basis_jacobian_finite_difference(cls, kind):
    rng = np.random.default_rng(3)
    qp, _ = _rule(kind, 3)
    x = qp[rng.choice(len(qp), 6, replace=False)]
    b = cls(4, x)
    j = b.jac_ortho_basis_at(x, do_clean=False)     # (ndim, nb, npts)
    h = 1e-6
    for d in range(3):
        dx = np.zeros(3)
        dx[d] = h
        fd = (b.ortho_basis_at(x + dx, False) -
              b.ortho_basis_at(x - dx, False))/(2*h)
        assert np.abs(j[d] - fd).max() < 1e-6


def test_planted_jacobi_weight_mistake_is_caught():
    """A wrong Jacobi weight in the tet basis breaks orthonormality."""
    class Planted(ob.TetBasis):
        def _ortho(self, p, q, r):
            real = ob.jacobi
            ob.jacobi = lambda n, a, b, z: real(n, a if a < 2 else a + 1, b, z)
            try:
                return super()._ortho(p, q, r)
            finally:
                ob.jacobi = real
    qp, qw = _rule('tet', 6)
    phi = Planted(3, qp).ortho_basis_at(qp, do_clean=False)
    g = (phi*qw) @ phi.T
    assert np.abs(g - np.eye(len(g))).max() > 1e-3


def _face_rule(kind, n):
    return _rule(kind, n)


@pytest.mark.parametrize('name', ['tet', 'pri'])
@pytest.mark.parametrize('p', [1, 2, 3, 4])
def test_oracle_operators_exact(name, p):
    o = _oracle(name, p)
    rng = np.random.default_rng(p)
    f, grad = _randpoly(rng, 3, p, name)
    u = f(o.upts)

    # M0: interpolation to flux points; M4: gradient at solution points
    assert np.abs(o.m0 @ u - f(o.fpts)).max() < 1e-11
    g = (o.m4 @ u).reshape(3, -1).T
    assert np.abs(g - grad(o.upts)).max() < 1e-10

    # Divergence of a polynomial flux with a consistent normal flux
    F = np.stack([f(o.upts)*(d + 1) for d in range(3)])
    fn = (o.norm_fpts*f(o.fpts)[:, None]*np.arange(1, 4)).sum(axis=1)
    div = o.opmat('M1 - M3*M2') @ F.reshape(-1) + o.opmat('M3') @ fn
    assert np.abs(div - grad(o.upts) @ np.arange(1, 4)).max() < 1e-10

    # Lifting identity for M3 (and M6 = n (x) M3) with v of degree p
    qp, qw = _rule(name, p + 2)
    lu = o.ubasis.nodal_basis_at(qp)                    # (nq, nupts)
    v, _ = _randpoly(rng, 3, p, name)
    lhs = (lu @ o.m3).T @ (qw*v(qp))                    # (nfpts,)
    lhs6 = [(lu @ o.m6[d*o.nupts:(d + 1)*o.nupts]).T @ (qw*v(qp))
            for d in range(3)]
    rhs = []
    for (kind, proj, _), nf in zip(o.faces, o.nfacefpts):
        fq, fw = _face_rule(kind, p + 2)
        L = o._facebasis(kind).nodal_basis_at(fq)       # (nq, nf)
        rhs.append(L.T @ (fw*v(ob._proj(proj, fq))))
    rhs = np.concatenate(rhs)
    assert np.abs(lhs - rhs).max() < 1e-11
    for d in range(3):
        assert np.abs(lhs6[d] - o.norm_fpts[:, d]*rhs).max() < 1e-11


@pytest.mark.parametrize('name,cls', [('tet', TetShape), ('pri', PriShape)])
@pytest.mark.parametrize('p', [1, 2, 3, 4])
def test_product_operators_match_oracle(name, cls, p):
    s, o = cls(p), _oracle(name, p)
    assert np.array_equal(s.upts, o.upts)
    # quad-face Gauss-Legendre points: the oracle takes numpy's rule
    assert np.abs(s.fpts - o.fpts).max() < 1e-15
    assert np.array_equal(s.norm_fpts, o.norm_fpts)
    for op in OPS:
        a, b = s.opmat(op), o.opmat(op)
        assert a.shape == b.shape
        assert np.array_equal(a != 0, b != 0), op
        assert np.abs(a - b).max() < 1e-12*max(1, np.abs(b).max()), op


def test_operator_shapes_tet_pri_p4():
    t, pr = TetShape(4), PriShape(4)
    assert t.opmat('M0').shape == (60, 35)
    assert t.opmat('M1 - M3*M2').shape == (35, 105)
    assert t.opmat('M3').shape == (35, 60)
    assert pr.opmat('M0').shape == (105, 75)
    assert pr.opmat('M1 - M3*M2').shape == (75, 225)
    assert pr.opmat('M3').shape == (75, 105)


@pytest.mark.parametrize('layout', [(2, 1, 1), (3, 2, 1), (2, 0, 2),
                                    (2, 2, 0)])
def test_mixed_connectivity_against_bruteforce(layout):
    """Node-free product pairing (structured face keys + fuzzy-sorted flux
    points) against the oracle's geometric brute force, on stacked tet/prism
    meshes with tet-prism interfaces at both ends of the prism slab."""
    from oracle.simplex import pair_faces_mixed, _Etype
    from pyfr_amd.inifile import Inifile
    from pyfr_amd.mesh import SimplexBoxMesh
    from pyfr_amd.solvers import SHAPES, FluidElements
    n, nt, npr = layout
    p = 2
    m = SimplexBoxMesh.stacked(n, nt, npr)
    cfg = Inifile.from_dict({'solver': {'order': p},
                             'constants': {'gamma': 1.4}})
    eles = {et: FluidElements(SHAPES[et], v, cfg)
            for et, v in m.spts.items()}
    lhs, rhs = m.con
    nfaces = sum(len(ets.basis.nfacefpts)*ets.neles for ets in eles.values())
    assert 2*len(lhs) == nfaces
    ours = set()
    for (tl, el, fl), (tr, er, fr) in zip(lhs, rhs):
        pl = eles[tl].srtd_face_fpts[fl][el]
        pr = eles[tr].srtd_face_fpts[fr][er]
        for a, b in zip(pl, pr):
            ours.add(tuple(sorted(((tl, el, int(a)), (tr, er, int(b))))))
    L = 2*np.pi
    ets = {et: _Etype(et, p, _tables(et, p), v.transpose(0, 2, 1))
           for et, v in m.spts.items()}
    idx = {et: np.arange(v.shape[1]) for et, v in m.spts.items()}
    ref = set()
    for tl, el, il, tr, er, ir in pair_faces_mixed(ets, [((L, L, L), idx)]):
        for a, b in zip(il, ir):
            ref.add(tuple(sorted(((tl, el, int(a)), (tr, er, int(b))))))
    assert ours == ref


def test_mixed_mesh_geometry_matches_oracle():
    from oracle.simplex import _Etype
    from pyfr_amd.geometry import AffineGeometry
    from pyfr_amd.mesh import SimplexBoxMesh
    m = SimplexBoxMesh.stacked(3, 2, 1)
    for et, v in m.spts.items():
        g = AffineGeometry(v, et)
        o = _Etype(et, 3, _tables(et, 3), v.transpose(0, 2, 1))
        sh = o.sh
        assert np.abs(g.ploc_at(sh.upts) - o.ploc(sh.upts)).max() < 1e-13
        assert np.abs(g.rcpdjac_at(sh.upts) - o.rcpdjac).max() < 1e-13
        assert np.abs(g.smat_at(sh.upts).transpose(1, 0, 2, 3) -
                      o.smats).max() < 1e-13
        assert np.abs(g.pnorm_at(sh.fpts, sh.norm_fpts) - o.pnorm).max() \
            < 1e-13
    # cell volumes: each cube splits into pieces of total volume h^3
    h = 2*np.pi/3
    vol = sum(AffineGeometry(v, et).djac.sum()*(4/3 if et == 'tet' else 4)
              for et, v in m.spts.items())
    assert abs(vol - 27*h**3) < 1e-9


def test_disjoint_bench_mesh_counts():
    from pyfr_amd.mesh import SimplexBoxMesh
    m = SimplexBoxMesh.disjoint(4, 2)
    assert m.spts['tet'].shape == (4, 6*4**3, 3)
    assert m.spts['pri'].shape == (6, 2*2**3, 3)
    # every face paired once, never across the two boxes
    lhs, rhs = m.con
    assert 2*len(lhs) == 4*6*4**3 + 5*2*2**3
    assert all(a[0] == b[0] for a, b in zip(lhs, rhs))


def _density_wave_error(n, p, kind, plant=False):
    from oracle.simplex import OracleMixedSystem
    from pyfr_amd.mesh import SimplexBoxMesh
    U = np.array([0.5, 0.3, -0.2])
    m = SimplexBoxMesh.stacked(n, n if kind == 'tet' else 0,
                               0 if kind == 'tet' else n)
    L = 2*np.pi
    verts = {et: v.transpose(0, 2, 1) for et, v in m.spts.items()}
    idx = {et: np.arange(v.shape[1]) for et, v in m.spts.items()}
    o = OracleMixedSystem(verts, p, {et: _tables(et, p) for et in verts},
                          [((L, L, L), idx)])
    if plant:
        for et in o.et.values():
            et.M3 = -et.M3
    fns = {'rho': lambda x, y, z: 1 + 0.2*np.sin(x + y + z),
           'u': lambda x, y, z: U[0] + 0*x, 'v': lambda x, y, z: U[1] + 0*x,
           'w': lambda x, y, z: U[2] + 0*x, 'p': lambda x, y, z: 1 + 0*x}
    r = o.rhs(o.ics(fns))
    e2 = cnt = 0
    for et in verts:
        x, y, z = o.et[et].ploc(o.et[et].sh.upts).transpose(1, 0, 2)
        ex = -U.sum()*0.2*np.cos(x + y + z)
        e2 += ((r[et][:, 0] - ex)**2).sum()
        cnt += ex.size
    return np.sqrt(e2/cnt)


@pytest.mark.parametrize('kind', ['tet', 'pri'])
def test_oracle_simplex_rhs_converges(kind):
    """The oracle's tet/prism Euler RHS for an advected density wave
    converges at order p (p = 2, n = 4 -> 8: measured 1.95 tets, 2.05
    prisms); a sign error planted in its correction operator does not."""
    e4, e8 = _density_wave_error(4, 2, kind), _density_wave_error(8, 2, kind)
    assert np.log2(e4/e8) > 2 - 0.7
    p4 = _density_wave_error(4, 2, kind, True)
    p8 = _density_wave_error(8, 2, kind, True)
    assert np.log2(p4/p8) < 2 - 0.7


# -- Navier-Stokes and boundaries on tets / prisms (oracle) -----------------

def _mixed_oracle(n, nt, npr, p, periodic=(True, True, True), **kw):
    from oracle.simplex import OracleMixedSystem
    from pyfr_amd.mesh import SimplexBoxMesh
    m = SimplexBoxMesh.stacked(n, nt, npr, periodic=periodic)
    L = 2*np.pi
    verts = {et: v.transpose(0, 2, 1) for et, v in m.spts.items()}
    idx = {et: np.arange(v.shape[1]) for et, v in m.spts.items()}
    o = OracleMixedSystem(verts, p, {et: _tables(et, p) for et in verts},
                          [((L, L, (nt + npr)*L/n), idx)], periodic=periodic,
                          **kw)
    return m, o


def _shear_error(n, p, kind, mu, plant=False):
    """Shear flow rho = 1, p = 1, w = A sin(x): the exact viscous part of
    the NS RHS is (0, 0, 0, -mu A sin x, mu A^2 cos 2x) and the inviscid
    part vanishes.  The discrete viscous part is RHS(mu) - RHS(0), which
    removes the inviscid discretisation error."""
    A = 0.3
    fns = {'rho': lambda x, y, z: 1 + 0*x, 'u': lambda x, y, z: 0*x,
           'v': lambda x, y, z: 0*x, 'w': lambda x, y, z: A*np.sin(x),
           'p': lambda x, y, z: 1 + 0*x}
    r = []
    for m in (mu, 0.0):
        c = {'gamma': 1.4, 'mu': m, 'Pr': 0.71, 'ldg-beta': 0.0,
             'ldg-tau': 0.0}
        _, o = _mixed_oracle(n, n if kind == 'tet' else 0,
                             0 if kind == 'tet' else n, p,
                             system='navier-stokes', constants=c)
        if plant:
            for et in o.et.values():
                et.smats = et.smats.swapaxes(1, 2)
        r.append(o.rhs(o.ics(fns)))
    e2 = cnt = 0
    for et, oe in o.et.items():
        x, y, z = oe.ploc(oe.sh.upts).transpose(1, 0, 2)
        ex = np.zeros_like(r[0][et])
        ex[:, 3] = -mu*A*np.sin(x)
        ex[:, 4] = mu*A*A*np.cos(2*x)
        e2 += ((r[0][et] - r[1][et] - ex)**2).sum()
        cnt += ex.shape[0]*ex.shape[2]
    return np.sqrt(e2/cnt)


@pytest.mark.parametrize('kind', ['tet', 'pri'])
def test_oracle_simplex_ns_converges(kind):
    """The viscous part of the oracle's tet/prism NS RHS (LDG gradients
    with beta = tau = 0, viscous fluxes) for a manufactured shear flow
    converges to the exact one (p = 2, n = 4 -> 8: measured order 1.7 tets,
    2.2 prisms; the BR1-like central LDG flux is suboptimal); a transposed
    metric planted in the gradient transform diverges."""
    mu = 0.05
    e4, e8 = _shear_error(4, 2, kind, mu), _shear_error(8, 2, kind, mu)
    assert e8 < 0.15*mu*0.3
    assert np.log2(e4/e8) > 1.3
    b4 = _shear_error(4, 2, kind, mu, plant=True)
    b8 = _shear_error(8, 2, kind, mu, plant=True)
    assert b8 > b4 and b8 > 10*e8


def test_oracle_simplex_ns_inviscid_limit():
    """mu = 0: the NS RHS is the Euler RHS."""
    from tests.fr_cases import eval_ics, mode_ics
    c = {'gamma': 1.4, 'mu': 0.0, 'Pr': 0.71, 'ldg-beta': 0.0,
         'ldg-tau': 0.0}
    _, oe = _mixed_oracle(2, 1, 1, 2)
    _, on = _mixed_oracle(2, 1, 1, 2, system='navier-stokes', constants=c)
    u = oe.ics(eval_ics(mode_ics(seed=2)))
    re, rn = oe.rhs(u), on.rhs(u)
    for et in re:
        assert np.abs(re[et] - rn[et]).max() < 1e-12*np.abs(re[et]).max()


def test_simplex_boundaries_match_mesh_and_stay_steady():
    """Boundary faces: the oracle's geometric detection and the product
    mesh's bcon agree side by side, with outward normals; uniform states
    stay steady with characteristic and slip-wall sides (Euler) and with
    no-slip walls at rest (NS)."""
    per = (False, True, False)
    uinf = {'rho': 1.1, 'u': 0.0, 'v': 0.2, 'w': 0.0, 'p': 1.3}
    bcs = {'x0': ('char-riem-inv', uinf), 'x1': ('slp-adia-wall', {}),
           'z0': ('slp-adia-wall', {}), 'z1': ('char-riem-inv', uinf)}
    m, o = _mixed_oracle(3, 2, 1, 2, per, bcs=bcs)
    assert sorted(m.bcon) == ['x0', 'x1', 'z0', 'z1']
    got = {}
    for side, bt, gb, nl, v in o.bcs:
        ax, hi = 'xyz'.index(side[0]), side[1] == '1'
        nn = nl/np.linalg.norm(nl, axis=0)
        assert np.abs(nn[ax] - (1 if hi else -1)).max() < 1e-12
        got[side] = len(gb)
    for side, faces in m.bcon.items():
        npts = sum(o.et[t].sh.nfacefpts[f] for t, e, f in faces)
        assert got[side] == npts
    fns = {k: (lambda v: lambda x, y, z: v + 0*x)(v) for k, v in uinf.items()}
    u = o.ics(fns)
    r = o.rhs(u)
    assert max(np.abs(a).max() for a in r.values()) < 1e-12
    c = {'gamma': 1.4, 'mu': 0.01, 'Pr': 0.71, 'ldg-beta': 0.0,
         'ldg-tau': 0.0}
    walls = {s: ('no-slp-adia-wall', {}) for s in ('x0', 'x1', 'z0', 'z1')}
    _, on = _mixed_oracle(3, 2, 1, 2, per, system='navier-stokes',
                          constants=c, bcs=walls)
    rest = dict(uinf, v=0.0)
    fns = {k: (lambda v: lambda x, y, z: v + 0*x)(v) for k, v in rest.items()}
    r = on.rhs(on.ics(fns))
    assert max(np.abs(a).max() for a in r.values()) < 1e-12


# -- flux anti-aliasing on tets / prisms ------------------------------------

# Element rules of flux anti-aliasing: total degree stated in the rule-file
# names (witherden-vincent-n8-d3, ...)
AA_DEGS = {('tet', 'witherden-vincent', 8): 3,
           ('tet', 'witherden-vincent', 14): 5,
           ('tet', 'witherden-vincent', 35): 7,
           ('tet', 'witherden-vincent', 59): 9,
           ('pri', 'witherden-vincent', 28): 6, ('pri', 'witherden', 46): 8,
           ('pri', 'witherden', 82): 10, ('pri', 'witherden', 128): 12}


@pytest.mark.parametrize('key', list(AA_DEGS))
def test_aa_rules_exact_to_stated_degree(key):
    deg = AA_DEGS[key]
    pts, wts = (np.array(x, dtype=float) for x in TABLES[key])
    qp, qw = _rule(key[0], deg + 2)
    for e in it.product(range(deg + 1), repeat=3):
        if sum(e) > deg:
            continue
        f = lambda x: np.prod([x[:, d]**e[d] for d in range(3)], axis=0)
        assert abs(wts @ f(pts) - qw @ f(qp)) < 1e-12, e


def _aa_tables(name, p):
    from pyfr_amd.simplex import AA_RULES
    t = _tables(name, p)
    key = (name, *AA_RULES[name][p])
    assert AA_DEGS[key] >= (2*p + 1 if name == 'tet' else 2*(p + 2))
    t['qrule'] = TABLES[key]
    return t


@pytest.mark.parametrize('name', ['tet', 'pri'])
@pytest.mark.parametrize('p', [1, 2, 3, 4])
def test_aa_operators_match_oracle(name, p):
    """Product M7 / M8 against the oracle's reference-form construction;
    M8 M7 = I wherever the rule integrates products of two basis functions
    exactly (tets: degree 2p; prisms: 4p, so p <= 2 with the default
    rules, which the reference also uses at p = 3, 4)."""
    from pyfr_amd.simplex import AA_RULES, SIMPLEX_SHAPES
    t = _aa_tables(name, p)
    o = ob.SimplexShapeOracle(name, p, t)
    o.set_quadrature(*t['qrule'])
    sh = SIMPLEX_SHAPES[name](p, AA_RULES[name][p])
    for e in ('M7', 'M8', '(M1 - M3*M2)*M9'):
        a, b = o.opmat(e), sh.opmat(e)
        assert np.abs(a - b).max() < 1e-11*max(1, np.abs(a).max()), e
    if name == 'tet' or p <= 2:
        assert np.abs(o.m8 @ o.m7 - np.eye(o.nupts)).max() < 1e-12


@pytest.mark.parametrize('layout,p', [((2, 2, 0), 1), ((2, 2, 0), 2),
                                      ((2, 2, 0), 3), ((2, 2, 0), 4),
                                      ((2, 0, 2), 1), ((2, 0, 2), 2),
                                      ((2, 1, 1), 2)])
def test_oracle_simplex_flux_aa(layout, p):
    """With uniform velocity and pressure the Euler flux is linear in the
    state, so where the projection is exact the anti-aliased RHS equals
    the plain one; a nonlinear field changes it; a uniform state stays
    steady."""
    from oracle.simplex import OracleMixedSystem
    from pyfr_amd.mesh import SimplexBoxMesh
    from tests.fr_cases import eval_ics, mode_ics
    n, nt, npr = layout
    m = SimplexBoxMesh.stacked(n, nt, npr)
    L = 2*np.pi
    verts = {et: v.transpose(0, 2, 1) for et, v in m.spts.items()}
    idx = {et: np.arange(v.shape[1]) for et, v in m.spts.items()}
    blocks = [((L, L, (nt + npr)*L/n), idx)]
    o = OracleMixedSystem(verts, p, {et: _tables(et, p) for et in verts},
                          blocks)
    oa = OracleMixedSystem(verts, p, {et: _aa_tables(et, p) for et in verts},
                           blocks, antialias=('flux',))
    U = np.array([0.5, 0.3, -0.2])
    fns = {'rho': lambda x, y, z: 1 + 0.2*np.sin(x + y + z),
           'u': lambda x, y, z: U[0] + 0*x, 'v': lambda x, y, z: U[1] + 0*x,
           'w': lambda x, y, z: U[2] + 0*x, 'p': lambda x, y, z: 1 + 0*x}
    u = o.ics(fns)
    r, ra = o.rhs(u), oa.rhs(u)
    for et in r:
        assert np.abs(ra[et] - r[et]).max() < 1e-11*np.abs(r[et]).max()
    u = o.ics(eval_ics(mode_ics(seed=4)))
    r, ra = o.rhs(u), oa.rhs(u)
    num = sum(np.linalg.norm(ra[et] - r[et])**2 for et in r)
    den = sum(np.linalg.norm(r[et])**2 for et in r)
    assert np.sqrt(num/den) > 1e-6
    fns = {k: (lambda v: lambda x, y, z: v + 0*x)(v) for k, v in
           {'rho': 1.1, 'u': 0.3, 'v': -0.2, 'w': 0.1, 'p': 0.9}.items()}
    ra = oa.rhs(oa.ics(fns))
    assert max(np.abs(a).max() for a in ra.values()) < 1e-12


# -- surface-flux anti-aliasing on tets / prisms ----------------------------

def _iq_tables(p):
    from pyfr_amd.shapes import gauss_legendre
    from pyfr_amd.simplex import SF_TRI_RULES
    g, w = gauss_legendre(p + 2)
    quad = (np.array([(g[i], g[j]) for j in range(p + 2)
                      for i in range(p + 2)]),
            np.array([w[i]*w[j] for j in range(p + 2) for i in range(p + 2)]))
    return {'tri': TABLES[('tri', *SF_TRI_RULES[p])], 'quad': quad}


@pytest.mark.parametrize('p', [1, 2, 3, 4])
def test_sf_tri_rules_exact(p):
    """The default triangle-face rules integrate to degree 2(p + 1)."""
    pts, wts = (np.array(x, dtype=float) for x in _iq_tables(p)['tri'])
    deg = 2*(p + 1)
    qp, qw = _rule('tri', deg + 2)
    for e in it.product(range(deg + 1), repeat=2):
        if sum(e) <= deg:
            f = lambda x: x[:, 0]**e[0]*x[:, 1]**e[1]
            assert abs(wts @ f(pts) - qw @ f(qp)) < 1e-13, e


@pytest.mark.parametrize('name', ['tet', 'pri'])
@pytest.mark.parametrize('p', [1, 2, 3, 4])
def test_sf_operators(name, p):
    """Product operators against the oracle's reference-form construction;
    for the oracle: face data in the order-p face space give the standard
    correction (the projection reproduces them), and each interface-rule
    value is integrated with its weight."""
    from pyfr_amd.simplex import SF_TRI_RULES, SIMPLEX_SHAPES
    iq = _iq_tables(p)
    o = ob.SimplexShapeOracle(name, p, _tables(name, p))
    o.set_surface_quadrature(iq)
    sh = SIMPLEX_SHAPES[name](p, None, {'tri': SF_TRI_RULES[p],
                                        'quad': p + 2})
    for e in ('M0', 'M3', 'M1 - M3*M2', 'M4 - M6*M0', 'M6'):
        a, b = o.opmat(e), sh.opmat(e)
        assert np.abs(a - b).max() < 1e-11*max(1, np.abs(a).max()), e

    std = _oracle(name, p)
    rng = np.random.default_rng(p)
    gs, gq = np.empty(std.nfpts), np.empty(o.nfpts)
    offs, offq = np.cumsum([0] + std.nfacefpts), np.cumsum([0] + o.nfacefpts)
    for f, (kind, _, _) in enumerate(o.faces):
        g, _ = _randpoly(rng, 2, p, 'pri' if kind == 'quad' else 'tri')
        if kind == 'quad':
            # the quad face space is the tensor one
            c = rng.standard_normal((p + 1, p + 1))
            g = lambda x, c=c: sum(c[i, j]*x[:, 0]**i*x[:, 1]**j
                                   for i in range(p + 1)
                                   for j in range(p + 1))
        gs[offs[f]:offs[f + 1]] = g(std._frule[kind])
        gq[offq[f]:offq[f + 1]] = g(np.asarray(iq[kind][0], dtype=float))
    assert np.abs(o.m3 @ gq - std.m3 @ gs).max() < 1e-11*np.abs(gs).max()

    w = np.array(_tables(name, p)['upts'][1])
    col = w @ o.m3
    ref = np.concatenate([np.asarray(iq[k][1]) for k, _, _ in o.faces])
    assert np.abs(np.abs(col) - ref).max() < 1e-12


@pytest.mark.parametrize('layout,p', [((2, 2, 0), 2), ((2, 0, 2), 2),
                                      ((2, 1, 1), 3), ((2, 1, 1), 1)])
def test_oracle_simplex_sf_rhs_identity(layout, p):
    """Uniform density and pressure, a different constant velocity in each
    element: each face sees two constant states, so the correction input is
    a constant on each face, which the projection reproduces: the
    surface-flux RHS equals the standard one."""
    from oracle.simplex import OracleMixedSystem
    from pyfr_amd.mesh import SimplexBoxMesh
    n, nt, npr = layout
    m = SimplexBoxMesh.stacked(n, nt, npr)
    L = 2*np.pi
    verts = {et: v.transpose(0, 2, 1) for et, v in m.spts.items()}
    idx = {et: np.arange(v.shape[1]) for et, v in m.spts.items()}
    blocks = [((L, L, (nt + npr)*L/n), idx)]
    o = OracleMixedSystem(verts, p, {et: _tables(et, p) for et in verts},
                          blocks)
    tabs = {et: dict(_tables(et, p), iqrules=_iq_tables(p)) for et in verts}
    osf = OracleMixedSystem(verts, p, tabs, blocks)
    rng = np.random.default_rng(p)
    u = {}
    for et, oe in o.et.items():
        v = 0.3*rng.standard_normal((3, oe.neles))
        uu = np.empty((oe.nupts, 5, oe.neles))
        uu[:, 0] = 1.2
        uu[:, 1:4] = 1.2*v[None]
        uu[:, 4] = (1/1.4)/0.4 + 0.6*(v**2).sum(axis=0)[None]
        u[et] = uu
    r, rs = o.rhs(u), osf.rhs(u)
    for et in r:
        assert np.abs(r[et]).max() > 1e-2
        assert np.abs(rs[et] - r[et]).max() < 1e-11*np.abs(r[et]).max()


# -- curved tets and prisms -------------------------------------------------

def _warp(x, y, z):
    return (x + 0.12*np.sin(y)*np.sin(z), y + 0.08*np.sin(x)*np.cos(z),
            z + 0.05*np.sin(x + y))


def _curved_oracle(n, nt, npr, p, sorder=2, warp=_warp, **kw):
    from oracle.simplex import OracleMixedSystem
    from pyfr_amd.mesh import SimplexBoxMesh
    m = SimplexBoxMesh.stacked(n, nt, npr, sorder=sorder, warp=warp)
    L = 2*np.pi
    verts = {et: v.transpose(0, 2, 1) for et, v in m.spts.items()}
    idx = {et: np.arange(v.shape[1]) for et, v in m.spts.items()}
    o = OracleMixedSystem(verts, p, {et: _tables(et, p) for et in verts},
                          [((L, L, (nt + npr)*L/n), idx)], **kw)
    return m, o


@pytest.mark.parametrize('p', [1, 2, 3])
def test_curved_simplex_geometry_matches_oracle(p):
    from pyfr_amd.geometry import CurvedSimplexGeometry
    m, o = _curved_oracle(2, 1, 1, p)
    for et, v in m.spts.items():
        g = CurvedSimplexGeometry(v, p, et)
        oe = o.et[et]
        assert oe.curved
        sh = oe.sh
        assert np.abs(g.ploc_at(sh.upts) - oe.ploc(sh.upts)).max() < 1e-13
        assert np.abs(g.smat_at(sh.upts).transpose(1, 0, 2, 3) -
                      oe.smats).max() < 1e-12
        assert np.abs(g.rcpdjac_at(sh.upts) - oe.rcpdjac).max() < 1e-12
        assert np.abs(g.pnorm_at(sh.fpts, sh.norm_fpts) - oe.pnorm).max() \
            < 1e-12


@pytest.mark.parametrize('p', [1, 3])
def test_curved_simplex_straight_equals_affine(p):
    """Straight elements given as order-2 shape points take the stored
    metric; it equals the affine one, and so does the RHS."""
    from tests.fr_cases import eval_ics, mode_ics
    _, oc = _curved_oracle(2, 1, 1, p, warp=None)
    _, oa = _curved_oracle(2, 1, 1, p, sorder=1, warp=None)
    assert all(e.curved for e in oc.et.values())
    u = oa.ics(eval_ics(mode_ics(seed=6)))
    rc, ra = oc.rhs(u), oa.rhs(u)
    for et in ra:
        assert np.abs(rc[et] - ra[et]).max() < 1e-11*np.abs(ra[et]).max()


@pytest.mark.parametrize('layout', [(2, 2, 0), (2, 0, 2), (2, 1, 1)])
def test_curved_simplex_free_stream_and_conservation(layout):
    """On a warped mesh the invariant-form metric keeps a uniform state
    steady (a metric taken point by point from the exact Jacobian, planted
    in its place, does not) and the totals are conserved."""
    from tests.fr_cases import eval_ics, mode_ics
    n, nt, npr = layout
    _, o = _curved_oracle(n, nt, npr, 2)
    fns = {k: (lambda v: lambda x, y, z: v + 0*x)(v) for k, v in
           {'rho': 1.1, 'u': 0.3, 'v': -0.2, 'w': 0.1, 'p': 0.9}.items()}
    r = o.rhs(o.ics(fns))
    assert max(np.abs(a).max() for a in r.values()) < 1e-11
    u = o.ics(eval_ics(mode_ics(seed=2)))
    r = o.rhs(u)
    for v in range(5):
        tot = sum((np.array(_tables(et, 2)['upts'][1])[:, None] *
                   r[et][:, v]/oe.rcpdjac).sum() for et, oe in o.et.items())
        assert abs(tot) < 1e-10
    # plant: S and |J| by cofactors of the exact Jacobian at every point
    for oe in o.et.values():
        sh = oe.sh
        lb = oe._lb
        for pts, attr in ((sh.upts, 'smats'), (sh.fpts, None)):
            dl = lb.jac_nodal_basis_at(pts)
            j = np.einsum('avp,vbe->pabe', dl, oe.verts)
            s = np.empty_like(j)
            for i, (a, b) in enumerate([(1, 2), (2, 0), (0, 1)]):
                s[:, i] = np.cross(j[:, a], j[:, b], axis=1)
            if attr:
                oe.smats = s
            else:
                oe.pnorm = np.einsum('pi,pike->pek', sh.norm_fpts, s)
    r = o.rhs(o.ics(fns))
    assert max(np.abs(a).max() for a in r.values()) > 1e-6


def _mono_exact(shape, a, b, c=0):
    """Exact integral of x^a y^b (z^c) over the reference element: the
    triangle / tet with corners at -1 and the prism (triangle x [-1, 1]),
    from the unit-simplex moments i! j! (k!) / (i + j (+ k) + 2 (3))!; the
    pyramid layer by layer."""
    from math import comb, factorial

    def simplex(powers):
        # integral of prod (2 u_i - 1)^p_i over the unit simplex, scaled
        # to the reference element (Jacobian 2^dim)
        d = len(powers)
        tot = 0.0
        terms = [[(comb(p, k), k, p - k) for k in range(p + 1)]
                 for p in powers]
        import itertools as it
        for choice in it.product(*terms):
            coef, sign = 1.0, 1
            ks = []
            for cb, k, rest in choice:
                coef *= cb*2.0**k
                sign *= (-1)**rest
                ks.append(k)
            mom = np.prod([factorial(k) for k in ks])/factorial(sum(ks) + d)
            tot += sign*coef*mom
        return 2.0**d*tot
    if shape == 'tri':
        return simplex([a, b])
    if shape == 'tet':
        return simplex([a, b, c])
    if shape == 'pyr':
        # the layer at height z is the square |x|, |y| <= s = (1 - z)/2:
        # int x^a y^b = 4 s^(a+b+2)/((a+1)(b+1)) (a, b even), then z = 1 - 2s
        if a % 2 or b % 2:
            return 0.0
        m = a + b + 2
        zint = 2*sum(comb(c, k)*(-2.0)**k/(m + k + 1) for k in range(c + 1))
        return 4*zint/((a + 1)*(b + 1))
    line = (1 - (-1)**(c + 1))/(c + 1)
    return simplex([a, b])*line


@pytest.mark.parametrize('shape', ['tri', 'tet', 'pri', 'pyr'])
def test_quadrature_catalog_degrees(shape):
    """Every tabulated rule with weights integrates every monomial up to the
    degree its catalog entry states exactly (1e-12), and misses some
    monomial of the next degree (the stated degree is the rule's own);
    this pins the point / weight tables and the catalog together."""
    from pyfr_amd.quadtables import CATALOG, TABLES
    deg = {(r, n): d for r, n, d, _ in CATALOG[shape]}
    nd = 2 if shape == 'tri' else 3
    checked = 0
    for (s, r, n), (pts, wts) in TABLES.items():
        if s != shape or wts[0] is None:
            continue
        d = deg[r, n]
        x = np.array(pts)
        w = np.array(wts)
        assert len(w) == n
        worst, nxt = 0.0, 0.0
        for tot in range(d + 2):
            for a in range(tot + 1):
                for b in range(tot + 1 - a):
                    c = tot - a - b
                    if nd == 2 and c:
                        continue
                    f = x[:, 0]**a*x[:, 1]**b*(x[:, 2]**c if nd == 3 else 1)
                    e = abs(w @ f - _mono_exact(shape, a, b, c))
                    if tot <= d:
                        worst = max(worst, e)
                    else:
                        nxt = max(nxt, e)
        assert worst < 1e-12, (r, n, worst)
        # prism products of a triangle and a Gauss line rule can exceed
        # their stated degree in some monomials but not in all
        assert nxt > 1e-12, (r, n)
        checked += 1
    assert checked >= 5


def test_quad_deg_selection():
    """quad-deg picks the rule the reference's search picks: the defaults
    of flux and surface-flux anti-aliasing come back from their degrees
    (2p + 1 tets, 2(p + 2) prisms, 2(p + 1) triangles); a named family,
    a fixed point count, and an untabulated choice."""
    from pyfr_amd.cases import make_simplex_cfg
    from pyfr_amd.simplex import AA_RULES, SF_TRI_RULES, qdeg_rule
    from pyfr_amd.solvers import FluidElements
    for p in range(1, 5):
        assert qdeg_rule('tet', 2*p + 1) == AA_RULES['tet'][p]
        assert qdeg_rule('pri', 2*(p + 2)) == AA_RULES['pri'][p]
        assert qdeg_rule('tri', 2*(p + 1)) == SF_TRI_RULES[p]
    assert qdeg_rule('hex', 5, 'gauss-legendre') == ('gauss-legendre', 27)
    assert qdeg_rule('hex', 5) == ('witherden-vincent', 14)
    assert qdeg_rule('tet', npts=35) == ('witherden-vincent', 35)
    assert qdeg_rule('tet', 8) == ('witherden-vincent', 46)
    with pytest.raises(NotImplementedError):
        qdeg_rule('tet', 15)
    cfg = make_simplex_cfg(3)
    cfg.set('solver-elements-tet', 'quad-deg', '8')
    assert FluidElements._simplex_qrule(cfg, 'tet', 3) == \
        ('witherden-vincent', 46)
    cfg.set('solver-elements-hex', 'quad-deg', '7')
    with pytest.raises(NotImplementedError):
        FluidElements._gl_n1(cfg, 'solver-elements-hex', 3, 3)
    cfg.set('solver-elements-hex', 'quad-pts', 'gauss-legendre')
    assert FluidElements._gl_n1(cfg, 'solver-elements-hex', 3, 3) == 4
