"""The other solution / flux point families (CPU): [solver-elements-tet]
soln-pts = alpha-opt, [solver-elements-pri] soln-pts =
alpha-opt~gauss-legendre-lobatto | williams-shunn~gauss-legendre-lobatto,
[solver-elements-pyr] soln-pts = gauss-legendre-lobatto and
[solver-interfaces-tri] flux-pts = alpha-opt (pyfr/shapes.py:188-218).

The point sets are the reference's tables; they are pinned by their
structure where it is fixed (the prism sets are a triangle set times the
Gauss-Legendre-Lobatto line, the pyramid set is layered on the
Gauss-Lobatto nodes, every set lies in its element and is symmetric), and
the operators they give are checked against the oracle and for exactness on
degree-p polynomials.  Hex solution points and quad flux points other than
Gauss-Legendre are declined loudly.  Hex solution points may be
Gauss-Legendre-Lobatto: the operators (the tensor forms, with the face
basis on the Gauss-Legendre face points off the solution-point lines) match
the oracle's general construction (``pyfr/shapes.py:89-135``, ``:231-259``).
"""
import numpy as np
import pytest

import oracle.basis as ob
from pyfr_amd.quadtables import TABLES
from pyfr_amd.simplex import PriShape, PyrShape, TetShape

CASES = [(TetShape, 'alpha-opt', 'williams-shunn'),
         (TetShape, 'shunn-ham', 'alpha-opt'),
         (PriShape, 'alpha-opt~gauss-legendre-lobatto', 'alpha-opt'),
         (PriShape, 'williams-shunn~gauss-legendre-lobatto', 'williams-shunn'),
         (PyrShape, 'gauss-legendre-lobatto', 'alpha-opt')]
OPS = ['M0', 'M1 - M3*M2', 'M3', 'M4 - M6*M0', 'M6']


def _gll(n):
    """Gauss-Legendre-Lobatto nodes: the ends and the roots of P'_(n-1)."""
    from numpy.polynomial import legendre as L
    c = np.zeros(n)
    c[-1] = 1
    return np.concatenate([[-1.0], np.sort(L.legroots(L.legder(c))), [1.0]])


def _oracle(cls, p, up, fp):
    return ob.SimplexShapeOracle(cls.name, p, {
        'upts': TABLES[(cls.name, up, cls._nupts_of(p))],
        'tri_fpts': TABLES[('tri', fp, (p + 1)*(p + 2)//2)][0],
        'tri_q36': TABLES[('tri', 'williams-shunn', 36)]})


@pytest.mark.parametrize('cls,up,fp', CASES)
@pytest.mark.parametrize('p', [1, 2, 3, 4])
def test_operators_match_oracle(cls, up, fp, p):
    s, o = cls(p, upts_name=up, tri_fpts=fp), _oracle(cls, p, up, fp)
    assert np.array_equal(s.upts, o.upts)
    assert np.abs(s.fpts - o.fpts).max() < 1e-15
    tol = 1e-12 if (cls is not PyrShape or p <= 3) else 1e-10
    for op in OPS:
        a, b = s.opmat(op), o.opmat(op)
        assert np.abs(a - b).max() < tol*max(1, np.abs(b).max()), op


@pytest.mark.parametrize('cls,up,fp', CASES)
@pytest.mark.parametrize('p', [2, 4])
def test_operators_exact_on_polynomials(cls, up, fp, p):
    """M0 interpolates and M4 differentiates degree-p polynomials exactly,
    and the divergence of a polynomial flux with its own normal flux is
    exact, on every family."""
    s = cls(p, upts_name=up, tri_fpts=fp)
    rng = np.random.default_rng(p)
    ex = [(i, j, k) for i in range(p + 1) for j in range(p + 1 - i)
          for k in range(p + 1 - i - j)]
    c = rng.standard_normal(len(ex))

    def f(x):
        return sum(ci*x[:, 0]**i*x[:, 1]**j*x[:, 2]**k
                   for ci, (i, j, k) in zip(c, ex))

    def grad(x):
        h = 1e-6
        return np.stack([(f(x + h*e) - f(x - h*e))/(2*h)
                         for e in np.eye(3)], 1)
    u = f(s.upts)
    assert np.abs(s.m0 @ u - f(s.fpts)).max() < 1e-11
    g = (s.m4 @ u).reshape(3, -1).T
    assert np.abs(g - grad(s.upts)).max() < 1e-7
    F = np.stack([f(s.upts)*(d + 1) for d in range(3)])
    fn = (s.norm_fpts*f(s.fpts)[:, None]*np.arange(1, 4)).sum(axis=1)
    div = s.opmat('M1 - M3*M2') @ F.reshape(-1) + s.opmat('M3') @ fn
    assert np.abs(div - grad(s.upts) @ np.arange(1, 4)).max() < 1e-7


@pytest.mark.parametrize('p', [1, 2, 3, 4])
def test_point_set_structure(p):
    """Prism GLL sets: the p + 1 Gauss-Lobatto layers in z, each the
    named triangle set; pyramid GLL sets: the layers sit at Gauss-Lobatto
    nodes in z, each a square grid; every set inside its element and
    invariant under the element's symmetries (x <-> y)."""
    z = _gll(p + 1)
    for tri in ('alpha-opt', 'williams-shunn'):
        pts = np.array(TABLES[('pri', f'{tri}~gauss-legendre-lobatto',
                                PriShape._nupts_of(p))][0])
        t = np.array(TABLES[('tri', tri, (p + 1)*(p + 2)//2)][0])
        layers = sorted({round(v, 12) for v in pts[:, 2]})
        assert np.allclose(layers, z, atol=1e-14)
        for zz in layers:
            lay = pts[np.isclose(pts[:, 2], zz)][:, :2]
            d = np.abs(lay[:, None] - t[None]).max(axis=2).min(axis=1)
            assert d.max() < 1e-14
    pts = np.array(TABLES[('pyr', 'gauss-legendre-lobatto',
                           PyrShape._nupts_of(p))][0])
    zs = sorted({round(v, 12) for v in pts[:, 2]})
    assert len(zs) == p + 1 and np.isclose(zs[0], -1) and np.isclose(zs[-1], 1)
    for i, zz in enumerate(zs):
        lay = pts[np.isclose(pts[:, 2], zz)]
        assert len(lay) == (p + 1 - i)**2
    sw = pts[:, [1, 0, 2]]
    d = np.abs(pts[:, None] - sw[None]).max(axis=2).min(axis=1)
    assert d.max() < 1e-14
    for key in (('tet', 'alpha-opt', TetShape._nupts_of(p)),
                ('tri', 'alpha-opt', (p + 1)*(p + 2)//2)):
        x = np.array(TABLES[key][0])
        assert np.all(x >= -1 - 1e-14) and np.all(x.sum(axis=1) <= 2 - x.shape[1] + 1e-12)


def test_declined_families():
    from pyfr_amd.cases import make_cfg, make_simplex_cfg
    from pyfr_amd.mesh import BoxMesh, SimplexBoxMesh
    from pyfr_amd.solvers import SHAPES, SYSTEMS
    cls = SYSTEMS['euler'].elementscls
    cfg = make_cfg('euler', 2)
    cfg.set('solver-elements-hex', 'soln-pts', 'alpha-opt')
    with pytest.raises(NotImplementedError, match='hex soln-pts'):
        cls(SHAPES['hex'], BoxMesh(2).spts['hex'], cfg)
    cfg = make_simplex_cfg(2)
    cfg.set('solver-interfaces-quad', 'flux-pts', 'gauss-legendre-lobatto')
    m = SimplexBoxMesh.stacked(2, 1, 1)
    with pytest.raises(NotImplementedError, match='quad flux-pts'):
        cls(SHAPES['pri'], m.spts['pri'], cfg)
    cfg = make_simplex_cfg(2)
    cfg.set('solver-elements-tet', 'soln-pts', 'no-such-set')
    with pytest.raises(NotImplementedError, match='not tabulated'):
        cls(SHAPES['tet'], m.spts['tet'], cfg)


@pytest.mark.parametrize('p', [1, 2, 3, 4])
@pytest.mark.parametrize('nf1', [None, 'p+2'])
def test_hex_lobatto_operators(p, nf1):
    """Gauss-Legendre-Lobatto hexes, plain and with surface-flux
    anti-aliasing: the product's tensor operators against the oracle's
    general construction; the Lobatto rule is the ends and the roots of
    P'_p with weights exact to degree 2p - 1."""
    from oracle.basis import HexShapeOracle
    from pyfr_amd.shapes import HexShape, gauss_legendre_lobatto
    x, w = gauss_legendre_lobatto(p + 1)
    g = _gll(p + 1)
    assert np.abs(x - g).max() < 1e-15
    for k in range(2*p):
        assert abs(w @ x**k - (1 - (-1)**(k + 1))/(k + 1)) < 1e-14
    nf = p + 2 if nf1 else None
    s = HexShape(p, None, nf, 'gauss-legendre-lobatto')
    o = HexShapeOracle(p, 'gauss-legendre-lobatto')
    if nf:
        o.set_surface_quadrature(nf)
    assert np.abs(s.upts - o.upts).max() < 1e-15
    for op in OPS:
        a, b = s.opmat(op), o.opmat(op)
        assert np.abs(a - b).max() < 1e-13*max(1, np.abs(b).max()), op
