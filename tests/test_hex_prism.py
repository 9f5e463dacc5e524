"""Hexes and prisms in one mesh (CPU): a box of hexes below a slab of
prisms extruded along x, meeting on whole quad faces, as in meshes of a
hex core with prism layers.  The reference runs such meshes with one
element set per type and interfaces between them on their quad faces
(``pyfr/solvers/base/system.py:60-110``, ``pyfr/readers/native.py:404-495``).

Pins: the mesh's connectivity (every face paired once, quad faces across
the hex-prism planes, the Gmsh round trip); the oracle's hex element type
inside its mixed-mesh system reproduces the hex oracle exactly on an
all-hex box; on the hex-prism mesh a uniform state is steady and the
weighted RHS totals vanish (a wrongly oriented element breaks the steady
state; an unpaired face, or flux points paired across unequal quadrature
weights, breaks conservation)."""
import collections
import io

import numpy as np
import pytest

from pyfr_amd.mesh import SimplexBoxMesh
from tests.fr_cases import eval_ics, mode_ics, oracle_consts


def _tables(name, p):
    from pyfr_amd.quadtables import TABLES
    if name == 'hex':
        return {}
    nu = (p + 1)**2*(p + 2)//2
    return {'upts': TABLES[('pri', 'williams-shunn~gauss-legendre', nu)],
            'tri_fpts': TABLES[('tri', 'williams-shunn',
                                (p + 1)*(p + 2)//2)][0],
            'tri_q36': TABLES[('tri', 'williams-shunn', 36)]}


def _oracle(mesh, order, lz, system='euler', periodic=(True, True, True),
            bcs=None):
    from oracle.simplex import OracleMixedSystem
    L = 2*np.pi
    verts = {et: v.transpose(0, 2, 1) for et, v in mesh.spts.items()}
    idx = {et: np.arange(v.shape[1]) for et, v in mesh.spts.items()}
    return OracleMixedSystem(verts, order, {et: _tables(et, order)
                                            for et in verts},
                             [((L, L, lz), idx)], system=system,
                             constants=oracle_consts(system),
                             periodic=periodic, bcs=bcs)


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


@pytest.mark.parametrize('system,order', [('euler', 1), ('euler', 3),
                                          ('navier-stokes', 2)])
def test_oracle_hex_type_matches_hex_oracle(system, order):
    from oracle.rhs import OracleHexSystem
    n = 3
    mesh = SimplexBoxMesh.hex_prism(n, n, 0)
    orc = _oracle(mesh, order, 2*np.pi, system)
    ref = OracleHexSystem(n, order, system, constants=oracle_consts(system))
    assert np.array_equal(mesh.spts['hex'].transpose(0, 2, 1), ref.verts)
    u = orc.ics(eval_ics(mode_ics(seed=2)))
    r = orc.rhs(u)['hex']
    rr = ref.rhs(u['hex'])
    assert np.abs(rr).max() > 1e-2
    assert np.abs(r - rr).max() < 1e-12*np.abs(rr).max()


@pytest.mark.parametrize('system,order', [('euler', 2), ('euler', 4),
                                          ('navier-stokes', 3)])
def test_oracle_free_stream_and_conservation(system, order):
    n, nh, npr = 3, 2, 1
    mesh = SimplexBoxMesh.hex_prism(n, nh, npr)
    orc = _oracle(mesh, order, (nh + npr)*2*np.pi/n, system)
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
