"""Hexes and prisms in one mesh (CPU): a box of hexes below a slab of
prisms extruded along x, meeting on whole quad faces, as in meshes of a
hex core with prism layers.  The reference runs such meshes with one
element set per type and interfaces between them on their quad faces
(``pyfr/solvers/base/system.py:60-110``, ``pyfr/readers/native.py:404-495``).

Pins: the mesh's connectivity (every face paired once, quad faces across
the hex-prism planes, the Gmsh round trip); the oracle's hex element type
inside its mixed-mesh system reproduces the hex oracle exactly on an
all-hex box, straight or curved; on the curved hex-prism mesh a uniform
state is steady and the weighted RHS totals vanish (a wrongly oriented element breaks the steady
state; an unpaired face, or flux points paired across unequal quadrature
weights, breaks conservation)."""
import collections
import io

import numpy as np
import pytest

from pyfr_amd.mesh import SimplexBoxMesh
from tests.fr_cases import eval_ics, mode_ics, oracle_consts


def _tables(name, p, aa=()):
    """Oracle tables of a hex or prism at order p, with the default
    anti-aliasing rules for aa: Gauss-Legendre (p + 2) per direction for
    the hex element and all quad faces, the reference's default prism and
    triangle rules."""
    from pyfr_amd.quadtables import TABLES
    from pyfr_amd.shapes import gauss_legendre
    from pyfr_amd.simplex import AA_RULES, SF_TRI_RULES
    if name == 'hex':
        return ({'nq1': p + 2} if 'flux' in aa else {}) | (
            {'nf1': p + 2} if 'surf-flux' in aa else {})
    nu = (p + 1)**2*(p + 2)//2
    t = {'upts': TABLES[('pri', 'williams-shunn~gauss-legendre', nu)],
         'tri_fpts': TABLES[('tri', 'williams-shunn',
                             (p + 1)*(p + 2)//2)][0],
         'tri_q36': TABLES[('tri', 'williams-shunn', 36)]}
    if 'flux' in aa:
        t['qrule'] = TABLES[('pri', *AA_RULES['pri'][p])]
    if 'surf-flux' in aa:
        g, w = gauss_legendre(p + 2)
        r = range(p + 2)
        t['iqrules'] = {
            'tri': TABLES[('tri', *SF_TRI_RULES[p])],
            'quad': (np.array([(g[i], g[j]) for j in r for i in r]),
                     np.array([w[i]*w[j] for j in r for i in r]))}
    return t


def warp(x, y, z):
    """A smooth periodic warp of the box (every period 2 pi)."""
    return (x + 0.12*np.sin(y)*np.sin(z), y + 0.08*np.sin(x)*np.cos(z),
            z + 0.05*np.sin(x + y))


def _oracle(mesh, order, lz, system='euler', periodic=(True, True, True),
            bcs=None, aa=()):
    from oracle.simplex import OracleMixedSystem
    L = 2*np.pi
    verts = {et: v.transpose(0, 2, 1) for et, v in mesh.spts.items()}
    idx = {et: np.arange(v.shape[1]) for et, v in mesh.spts.items()}
    return OracleMixedSystem(verts, order, {et: _tables(et, order, aa)
                                            for et in verts},
                             [((L, L, lz), idx)], system=system,
                             constants=oracle_consts(system),
                             periodic=periodic, bcs=bcs,
                             antialias=tuple(a for a in aa if a == 'flux'))


def test_connectivity():
    m = SimplexBoxMesh.hex_prism(3, 2, 1)
    assert {k: v.shape for k, v in m.spts.items()} == {'hex': (8, 18, 3),
                                                       'pri': (6, 18, 3)}
    lhs, rhs = m.con
    # 18*6 + 18*5 face sides, each paired once
    assert len(lhs) == (18*6 + 18*5)//2
    pairs = collections.Counter((l[0], r[0]) for l, r in zip(lhs, rhs))
    # 9 quads on the hex-prism plane plus 9 across the periodic z wrap
    assert pairs[('hex', 'pri')] == 18 and ('pri', 'hex') not in pairs
    # those are prism quad faces (PriShape faces 2..4)
    assert all(r[2] >= 2 for l, r in zip(lhs, rhs) if r[0] != l[0])
    # positive Jacobians of the x-extruded prisms
    v = m.spts['pri']
    d = np.linalg.det(np.stack([v[1] - v[0], v[2] - v[0], v[3] - v[0]], 1))
    assert np.all(d > 0)
    b = SimplexBoxMesh.hex_prism(3, 2, 1, periodic=(False, True, False))
    assert {k: len(f) for k, f in b.bcon.items()} == {
        'x0': 6 + 6, 'x1': 6 + 6, 'z0': 9, 'z1': 9}
    assert {t for t, _, _ in b.bcon['z0']} == {'hex'}
    assert {t for t, _, _ in b.bcon['z1']} == {'pri'}
    with pytest.raises(ValueError):
        SimplexBoxMesh([((2, 2, 2), 1.0, [('tet', 1), ('hex', 1)])])


@pytest.mark.parametrize('version', ['2.2', '4.1'])
def test_gmsh_round_trip(version):
    from pyfr_amd.gmsh import GmshMesh
    from tests.msh_files import box_sides, mesh_to_msh
    n = 2
    ref = SimplexBoxMesh.hex_prism(n, 1, 1, periodic=(False, True, False))
    raw = SimplexBoxMesh.hex_prism(n, 1, 1, periodic=(False, False, False))
    spec = {'x0': 'inlet', 'x1': 'outlet', 'y0': 'periodic_y_l',
            'y1': 'periodic_y_r', 'z0': 'bottom', 'z1': 'top'}
    m = GmshMesh(io.StringIO(mesh_to_msh(raw, box_sides(raw, spec),
                                         version)))
    assert m.etypes == ['hex', 'pri']
    for t in m.etypes:
        assert np.allclose(m.spts[t], ref.spts[t], rtol=0, atol=1e-12)
    assert sorted(zip(*m.con)) == sorted(zip(*ref.con))
    for side, name in (('x0', 'inlet'), ('x1', 'outlet'), ('z0', 'bottom'),
                       ('z1', 'top')):
        assert sorted(m.bcon[name]) == sorted(ref.bcon[side])


@pytest.mark.parametrize('system,order,aa', [
    ('euler', 1, ()), ('euler', 3, ()), ('navier-stokes', 2, ()),
    ('euler', 2, ('flux',)), ('euler', 2, ('surf-flux',)),
    ('navier-stokes', 2, ('flux', 'surf-flux'))])
def test_oracle_hex_type_matches_hex_oracle(system, order, aa):
    from oracle.rhs import OracleHexSystem
    n = 3
    mesh = SimplexBoxMesh.hex_prism(n, n, 0)
    orc = _oracle(mesh, order, 2*np.pi, system, aa=aa)
    ref = OracleHexSystem(n, order, system, constants=oracle_consts(system),
                          antialias=aa, nq1=order + 2, nf1=order + 2)
    assert np.array_equal(mesh.spts['hex'].transpose(0, 2, 1), ref.verts)
    u = orc.ics(eval_ics(mode_ics(seed=2)))
    r = orc.rhs(u)['hex']
    rr = ref.rhs(u['hex'])
    assert np.abs(rr).max() > 1e-2
    assert np.abs(r - rr).max() < 1e-12*np.abs(rr).max()


@pytest.mark.parametrize('system,order,sorder', [
    ('euler', 1, 2), ('euler', 2, 2), ('euler', 3, 3),
    ('navier-stokes', 2, 2)])
def test_oracle_curved_hex_type_matches_hex_oracle(system, order, sorder):
    """Curved hexes (order-k shape points of a warped box, all flagged
    curved): the mixed system's invariant metric on std_ele(max(p, 1))
    gives the hex oracle's curved RHS."""
    from oracle.rhs import OracleHexSystem
    n = 3
    mesh = SimplexBoxMesh.hex_prism(n, n, 0, sorder=sorder, warp=warp)
    assert mesh.spts_curved['hex'].all()
    orc = _oracle(mesh, order, 2*np.pi, system)
    assert orc.et['hex'].curved
    ref = OracleHexSystem(n, order, system, constants=oracle_consts(system),
                          verts=mesh.spts['hex'].transpose(0, 2, 1),
                          curved=np.ones(n**3, dtype=bool))
    u = orc.ics(eval_ics(mode_ics(seed=2)))
    r = orc.rhs(u)['hex']
    rr = ref.rhs(u['hex'])
    assert np.abs(rr).max() > 1e-2
    assert np.abs(r - rr).max() < 1e-12*np.abs(rr).max()


@pytest.mark.parametrize('system,order,aa', [
    ('euler', 2, ()), ('euler', 4, ()), ('navier-stokes', 3, ()),
    ('euler', 2, ('flux', 'surf-flux')), ('navier-stokes', 2, ('surf-flux',))])
def test_oracle_free_stream_and_conservation(system, order, aa):
    n, nh, npr = 3, 2, 1
    mesh = SimplexBoxMesh.hex_prism(n, nh, npr, sorder=2, warp=warp)
    orc = _oracle(mesh, order, (nh + npr)*2*np.pi/n, system, aa=aa)
    uu = {}
    for et, e in orc.et.items():
        uu[et] = np.empty((e.nupts, 5, e.neles))
        uu[et][:] = np.array([1.0, 0.3, -0.2, 0.1, 6.32])[None, :, None]
    assert max(np.abs(r).max() for r in orc.rhs(uu).values()) < 1e-11
    u = orc.ics(eval_ics(mode_ics(seed=4)))
    r = orc.rhs(u)
    tot = np.zeros(5)
    for et, e in orc.et.items():
        w = (e.sh.upts_wts if et == 'hex' else
             np.array(_tables(et, order)['upts'][1], dtype=float))
        tot += np.einsum('u,uve->v', w, r[et]/e.rcpdjac[:, None, :])
    scale = max(np.abs(r[et]).max() for et in r)
    assert scale > 1e-2
    assert np.abs(tot).max() < 1e-11*scale*orc.et['hex'].neles
