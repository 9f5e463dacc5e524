"""Curved hexes (CPU): shape points of order k = 2, 3 moved by a smooth
warp.  The oracle's curved geometry is pinned by what the mathematics
fixes -- shape points reproduce themselves and a linear element's are its
trilinear map; Kopriva's invariant metric preserves a free stream exactly
on curved meshes, for Euler and NS, whole-curved or mixed with linear
elements -- and the product's host geometry (S, 1/|J|, normals, point
locations) and mesh (curved-first order, curved flags, interface pairing)
are checked against it."""
import numpy as np
import pytest

from oracle.geom import (hex_corners, hex_ploc, hex_std_ele, pnorm_at,
                         rcpdjac_at, smats_at, trilinear_ploc)
from oracle.mesh import pair_faces_bruteforce
from oracle.rhs import OracleHexSystem
from pyfr_amd.geometry import HexGeometry
from pyfr_amd.mesh import BoxMesh
from pyfr_amd.shapes import HexShape


def warp(x, y, z):
    """Smooth, periodic on [0, 2 pi]^3, zero for x in [pi, 2 pi]."""
    b = np.maximum(np.sin(x), 0)**2
    return (x + 0.12*np.sin(y)*np.sin(z)*b, y + 0.10*np.sin(2*z)*b,
            z + 0.08*np.sin(y)*b)


def _mesh(n, k, periodic=(True, True, True)):
    return BoxMesh(n, sorder=k, warp=warp, periodic=periodic)


@pytest.mark.parametrize('k', [2, 3])
def test_oracle_shape_map(k):
    rng = np.random.default_rng(k)
    spts = rng.standard_normal(((k + 1)**3, 3, 5))
    # the shape points reproduce themselves
    assert np.allclose(hex_ploc(spts, hex_std_ele(k)), spts, atol=1e-12)
    # a linear element given by order-k points maps as its trilinear map
    v = rng.standard_normal((8, 3, 5))
    lin = trilinear_ploc(v, hex_std_ele(k))
    x = rng.uniform(-1, 1, (20, 3))
    assert np.allclose(hex_ploc(lin, x), trilinear_ploc(v, x), atol=1e-12)
    assert np.array_equal(hex_corners(lin), v)


@pytest.mark.parametrize('system,order,k', [('euler', 2, 2), ('euler', 3, 3),
                                            ('navier-stokes', 3, 2),
                                            ('euler', 1, 3)])
def test_free_stream_on_curved_mesh(system, order, k):
    """The metric's invariant form makes a uniform state an exact steady
    solution on curved elements, whole-curved (curved=None at p < 2 is
    all stored metric) and mixed with linear ones."""
    m = _mesh(3, k)
    c = {'gamma': 1.4} if system == 'euler' else \
        {'gamma': 1.4, 'mu': 0.01, 'Pr': 0.71, 'ldg-beta': 0.0,
         'ldg-tau': 0.0}
    v = m.spts['hex'].transpose(0, 2, 1)
    o = OracleHexSystem(3, order, system, constants=c, verts=v,
                        curved=m.spts_curved['hex'])
    u = np.empty((o.nupts, 5, o.neles))
    u[:] = np.array([1.0, 0.3, -0.2, 0.1,
                     2.5/0.4 + 0.5*(0.09 + 0.04 + 0.01)])[None, :, None]
    assert np.abs(o.rhs(u)).max() < 1e-10


def test_curved_mesh_order_and_pairing():
    n, k = 4, 3
    m = _mesh(n, k)
    cf = m.spts_curved['hex']
    # curved first, some of each
    assert 0 < cf.sum() < n**3
    assert np.all(cf[:cf.sum()]) and not np.any(cf[cf.sum():])
    # linear elements were snapped onto their trilinear map
    sp = m.spts['hex'].transpose(0, 2, 1)
    lin = trilinear_ploc(hex_corners(sp), hex_std_ele(k))
    assert np.abs(sp[..., ~cf] - lin[..., ~cf]).max() < 1e-14
    assert np.abs(sp[..., cf] - lin[..., cf]).max() > 1e-3
    # interface pairs from the structured keys match geometric brute force
    # on the curved flux points
    from pyfr_amd.solvers import FluidElements
    from pyfr_amd.inifile import Inifile
    p = 2
    cfg = Inifile.from_dict({'solver': {'order': p},
                             'constants': {'gamma': 1.4}})
    e = FluidElements(HexShape, m.spts['hex'], cfg)
    sf = e.srtd_face_fpts
    ours = set()
    for (_, el, fl), (_, er, fr) in zip(*m.con):
        for a, b in zip(sf[fl][el], sf[fr][er]):
            ours.add(tuple(sorted(((el, a), (er, b)))))
    from oracle.basis import HexShapeOracle
    o = HexShapeOracle(p)
    pl = hex_ploc(sp, o.fpts)
    ref = set()
    for eL, fL, iL, eR, fR, iR in pair_faces_bruteforce(pl, 6, o.nfacefpts,
                                                        2*np.pi):
        for a, b in zip(iL, iR):
            ref.add(tuple(sorted(((eL, a), (eR, b)))))
    assert ours == ref


@pytest.mark.parametrize('k,p', [(2, 2), (3, 3), (3, 1)])
def test_curved_geometry_against_oracle(k, p):
    m = _mesh(3, k)
    sp = m.spts['hex']
    sh = HexShape(p)
    g = HexGeometry(sp, p)
    ov = sp.transpose(0, 2, 1)
    assert np.allclose(g.ploc_at(sh.upts), hex_ploc(ov, sh.upts),
                       rtol=0, atol=1e-12)
    s_o = smats_at(ov, p, sh.upts)                        # (3, Nu, 3, E)
    s_p = g.smat_at(sh.upts)
    assert np.allclose(s_p.reshape(s_o.shape), s_o, rtol=0, atol=1e-11)
    r_o = rcpdjac_at(ov, p, sh.upts)
    assert np.allclose(g.rcpdjac_at(sh.upts), r_o, rtol=1e-11)
    pn_o = pnorm_at(ov, p, sh.fpts, sh.norm_fpts)        # (Nf, E, 3)
    pn_p = g.pnorm_at(sh.fpts, sh.norm_fpts)
    assert np.allclose(pn_p.reshape(pn_o.shape), pn_o, rtol=0, atol=1e-11)


@pytest.mark.parametrize('order', [1, 2, 3])
def test_flux_antialias_oracle(order):
    """Flux anti-aliasing (quadrature-point flux, M8 projection): with
    uniform velocity and pressure the Euler flux is linear in the state, so
    on an affine mesh the projected flux is the flux itself and the
    anti-aliased RHS equals the plain one; on a warped mesh a uniform state
    stays steady and the totals are conserved."""
    from oracle.mesh import box_hex_verts
    U = np.array([0.5, 0.3, -0.2])
    o = OracleHexSystem(3, order)
    oa = OracleHexSystem(3, order, antialias=('flux',))
    x, y, z = trilinear_ploc(o.verts, o.shape.upts).transpose(1, 0, 2)
    rho = 1 + 0.2*np.sin(x + y + z)
    u = np.stack([rho, rho*U[0], rho*U[1], rho*U[2],
                  1.0/0.4 + 0.5*rho*(U**2).sum()], axis=1)
    r, ra = o.rhs(u), oa.rhs(u)
    assert np.abs(ra - r).max() < 1e-12*np.abs(r).max()
    # a nonlinear field: the projection changes the RHS
    r, ra = o.rhs(o.ics()), oa.rhs(oa.ics())
    assert np.linalg.norm(ra - r) > 1e-6*np.linalg.norm(r)
    w = oa.shape.upts_wts
    assert max(abs((w[:, None]*ra[:, v]/oa.rcpdjac).sum())
               for v in range(5)) < 1e-10
    wv = box_hex_verts(3, warp=lambda x, y, z: (x + 0.1*np.sin(y)*np.sin(z),
                                                y, z))
    ow = OracleHexSystem(3, order, verts=wv, antialias=('flux',))
    uu = np.empty((ow.nupts, 5, ow.neles))
    uu[:] = np.array([1.0, 0.3, -0.2, 0.1, 6.32])[None, :, None]
    assert np.abs(ow.rhs(uu)).max() < 1e-11
