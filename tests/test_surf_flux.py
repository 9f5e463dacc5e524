"""Surface-flux anti-aliasing on hexes (``anti-alias = surf-flux``,
``pyfr/shapes.py:103-113``, ``:196-224``, ``:297-303``): the flux points of
each face move to an nf1 x nf1 Gauss-Legendre rule, and the correction is
fed by the L2 projection of the face values onto the order-p face basis.

The oracle's construction (the reference's generic ``gbasis_at @
block_diag(proj_l2(...))``) is pinned by what the mathematics fixes:
projection exactness on the face space, the reduction to the standard
scheme at nf1 = p + 1, the face weights carried by the correction, and an
RHS that must equal the standard one when the face flux is a face-space
polynomial.  The product's closed-form operators are checked against it."""
import numpy as np
import pytest

from oracle.basis import HexShapeOracle, gauss_legendre_quad, QuadBasis
from oracle.rhs import OracleHexSystem


OPS = ('M0', 'M3', 'M1 - M3*M2', 'M4 - M6*M0', 'M6')


def _face_poly(order, rng):
    """A random polynomial of degree <= order in each face coordinate."""
    c = rng.standard_normal((order + 1, order + 1))
    return lambda s, t: sum(c[i, j]*s**i*t**j for i in range(order + 1)
                            for j in range(order + 1))


@pytest.mark.parametrize('order', [1, 2, 3, 4])
def test_product_operators_match_oracle(order):
    from pyfr_amd.shapes import HexShape
    for nf1 in (order + 1, order + 2, order + 3):
        o = HexShapeOracle(order)
        o.set_surface_quadrature(nf1)
        h = HexShape(order, nf1=nf1)
        assert h.nfpts == o.nfpts == 6*nf1*nf1
        assert np.abs(h.fpts - o.fpts).max() < 1e-15
        assert np.abs(h.norm_fpts - o.norm_fpts).max() == 0
        for e in OPS:
            assert np.abs(h.opmat(e) - o.opmat(e)).max() < 1e-12, (nf1, e)


@pytest.mark.parametrize('order', [1, 2, 3, 4])
def test_projection_exact_on_face_space(order):
    """Face data in the order-p tensor space: the projection reproduces it
    at the standard flux points, so M3 gives the standard correction; M0
    interpolates order-p solutions exactly at the quadrature points."""
    rng = np.random.default_rng(order)
    std = HexShapeOracle(order)
    sf = HexShapeOracle(order)
    sf.set_surface_quadrature(order + 2)
    gs = np.empty(std.nfpts)
    gq = np.empty(sf.nfpts)
    for f in range(6):
        g = _face_poly(order, rng)
        s, t = std.face_pts2.T
        gs[std.facefpts[f]] = g(s, t)
        s, t = sf.iq_pts.T
        gq[sf.facefpts[f]] = g(s, t)
    assert np.abs(sf.m3 @ gq - std.m3 @ gs).max() < 1e-12*np.abs(gs).max()

    c = rng.standard_normal((order + 1,)*3)

    def u(p):
        x, y, z = p.T
        return sum(c[i, j, k]*x**i*y**j*z**k
                   for i in range(order + 1) for j in range(order + 1)
                   for k in range(order + 1))
    assert np.abs(sf.m0 @ u(sf.upts) - u(sf.fpts)).max() < 1e-12


@pytest.mark.parametrize('order', [1, 2, 3])
def test_reduces_to_standard_and_carries_face_weights(order):
    std = HexShapeOracle(order)
    same = HexShapeOracle(order)
    same.set_surface_quadrature(order + 1)
    for e in OPS:
        assert np.abs(same.opmat(e) - std.opmat(e)).max() < 1e-13, e
    # the correction integrates each face value with its quadrature
    # weight: sum_u w_u M3[u, q] = -w_q (the reference's sign)
    for nf1 in (order + 2, order + 3):
        sf = HexShapeOracle(order)
        sf.set_surface_quadrature(nf1)
        col = sf.upts_wts @ sf.m3
        _, iw = gauss_legendre_quad(nf1)
        ref = np.abs(np.concatenate([iw]*6))
        assert np.abs(np.abs(col) - ref).max() < 1e-13
        # and the face basis it projects onto is the order-p one
        fb = QuadBasis(order, std.face_pts2)
        assert fb.vdm.shape == (std.nfacefpts[0],)*2


@pytest.mark.parametrize('order,nf1', [(1, 3), (2, 4), (3, 5), (3, 6)])
def test_rhs_equals_standard_for_face_space_flux(order, nf1):
    """Uniform density and pressure, a different constant velocity in each
    cell, an affine mesh: each face sees two constant states, so the
    correction input (common flux minus the interpolated normal flux) is a
    nonzero constant on each face, which the projection must reproduce."""
    n = 4
    o = OracleHexSystem(n, order)
    osf = OracleHexSystem(n, order, antialias=('surf-flux',), nf1=nf1)
    rng = np.random.default_rng(order)
    v = 0.3*rng.standard_normal((3, o.neles))
    rho, p = 1.2, 1/1.4
    u = np.empty((o.nupts, 5, o.neles))
    u[:, 0] = rho
    u[:, 1:4] = rho*v[None]
    u[:, 4] = p/0.4 + 0.5*rho*(v**2).sum(axis=0)[None]
    r, rs = o.rhs(u), osf.rhs(u)
    assert np.abs(r).max() > 1e-2
    assert np.abs(rs - r).max() < 1e-12*np.abs(r).max()


@pytest.mark.parametrize('system', ['euler', 'navier-stokes'])
def test_free_stream_and_conservation(system):
    from oracle.mesh import box_hex_verts
    from tests.fr_cases import oracle_consts
    c = oracle_consts(system)
    # a nonlinear field changes the RHS; totals are conserved
    o = OracleHexSystem(3, 2, system, constants=c)
    osf = OracleHexSystem(3, 2, system, constants=c,
                          antialias=('surf-flux',))
    r, rs = o.rhs(o.ics()), osf.rhs(osf.ics())
    assert np.linalg.norm(rs - r) > 1e-6*np.linalg.norm(r)
    w = osf.shape.upts_wts
    assert max(abs((w[:, None]*rs[:, v]/osf.rcpdjac).sum())
               for v in range(5)) < 1e-10
    # a uniform state stays steady on a warped mesh, also with both
    # anti-aliasing kinds
    wv = box_hex_verts(3, warp=lambda x, y, z: (x + 0.1*np.sin(y)*np.sin(z),
                                                y, z))
    for aa in (('surf-flux',), ('flux', 'surf-flux')):
        ow = OracleHexSystem(3, 2, system, constants=c, verts=wv,
                             antialias=aa)
        uu = np.empty((ow.nupts, 5, ow.neles))
        uu[:] = np.array([1.0, 0.3, -0.2, 0.1, 6.32])[None, :, None]
        assert np.abs(ow.rhs(uu)).max() < 1e-10, aa
