"""Gmsh import (CPU): meshes written out as .msh 2.2 / 4.1 text and read
back give the generated meshes exactly: vertices, interface pairs with the
reference's left/right orientation, periodic pairs and named boundaries;
unsupported and malformed inputs are declined."""
import io

import numpy as np
import pytest

from pyfr_amd.gmsh import FACES, GmshMesh
from pyfr_amd.mesh import BoxMesh, SimplexBoxMesh
from tests.msh_files import box_sides, mesh_to_msh


def _pairs(m):
    return sorted(zip(m.con[0], m.con[1]))


@pytest.mark.parametrize('version', ['2.2', '4.1'])
@pytest.mark.parametrize('periodic', [(True, False, True),
                                      (True, True, True),
                                      (False, False, False)])
def test_hex_box_round_trip(version, periodic):
    n = 3
    ref = BoxMesh(n, periodic=periodic)
    raw = BoxMesh(n, periodic=(False, False, False))
    spec = {}
    for a, ax in enumerate('xyz'):
        if periodic[a]:
            spec[f'{ax}0'], spec[f'{ax}1'] = f'periodic_{ax}_l', \
                f'periodic_{ax}_r'
        else:
            spec[f'{ax}0'], spec[f'{ax}1'] = f'{ax}lo', f'{ax}hi'
    m = GmshMesh(io.StringIO(mesh_to_msh(raw, box_sides(raw, spec),
                                         version)))
    assert m.etypes == ['hex'] and m.neles == n**3
    assert np.array_equal(m.spts['hex'], ref.spts['hex'])
    assert _pairs(m) == _pairs(ref)
    for a, ax in enumerate('xyz'):
        if not periodic[a]:
            assert sorted(m.bcon[f'{ax}lo']) == sorted(ref.bcon[f'{ax}0'])
            assert sorted(m.bcon[f'{ax}hi']) == sorted(ref.bcon[f'{ax}1'])


def _simplex_sides(mesh, L):
    """Faces of a SimplexBoxMesh block on the six box sides."""
    out = {}
    for t in mesh.spts:
        pts = mesh.spts[t]                         # (nv, neles, 3)
        for f, fv in enumerate(FACES[t]):
            c = pts[list(fv)].mean(axis=0)         # (neles, 3)
            for a, ax in enumerate('xyz'):
                for s, x in ((0, 0.0), (1, L[a])):
                    on = np.flatnonzero(np.isclose(c[:, a], x))
                    name = f'periodic_{ax}_{"lr"[s]}'
                    out.setdefault(name, []).extend(
                        (t, int(e), f) for e in on)
    return out


@pytest.mark.parametrize('version', ['2.2', '4.1'])
def test_tet_prism_round_trip(version):
    n, nt, npr = 2, 1, 1
    ref = SimplexBoxMesh.stacked(n, nt, npr)
    L = 2*np.pi
    sides = _simplex_sides(ref, (L, L, (nt + npr)*L/n))
    m = GmshMesh(io.StringIO(mesh_to_msh(ref, sides, version)))
    assert m.etypes == ['tet', 'pri']
    for t in m.etypes:
        assert np.allclose(m.spts[t], ref.spts[t], rtol=0, atol=1e-12)
    assert _pairs(m) == _pairs(ref)
    assert m.bcon == {}


def _tiny(version='4.1'):
    raw = BoxMesh(1, periodic=(False, False, False))
    spec = {s: 'wall' for s in ('x0', 'x1', 'y0', 'y1', 'z0', 'z1')}
    return mesh_to_msh(raw, box_sides(raw, spec), version)


def test_declines_and_errors():
    txt = _tiny('2.2')
    m = GmshMesh(io.StringIO(txt))
    assert len(m.bcon['wall']) == 6 and m.con == ([], [])
    # a second-order hex (Gmsh type 12) is declined
    bad = txt.replace(' 5 2 1 1 ', ' 12 2 1 1 ')
    with pytest.raises(NotImplementedError):
        GmshMesh(io.StringIO(bad))
    # a boundary face without a physical name leaves a face unpaired
    lines = txt.split('\n')
    i = lines.index('$Elements')
    cut = lines[:i + 1] + [str(int(lines[i + 1]) - 1)] + lines[i + 3:]
    with pytest.raises(ValueError, match='Unpaired faces'):
        GmshMesh(io.StringIO('\n'.join(cut)))
    # a one-sided periodic pair
    one = txt.replace('"wall"', '"periodic_a_l"')
    with pytest.raises(ValueError, match='Unpaired periodic'):
        GmshMesh(io.StringIO(one))
    # binary files and other versions
    with pytest.raises(ValueError):
        GmshMesh(io.StringIO(txt.replace('2.2 0 8', '2.2 1 8')))
    with pytest.raises(ValueError):
        GmshMesh(io.StringIO(txt.replace('2.2 0 8', '3.0 0 8')))
