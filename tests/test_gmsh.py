"""Gmsh import (CPU): meshes written out as .msh 2.2 / 4.1 text and read
back give the generated meshes exactly: vertices (or the shape points of
high-order elements), interface pairs with the reference's left/right
orientation, periodic pairs, named boundaries, the curved / linear split
and the curved-first order; unsupported and malformed inputs are declined.
The Gmsh node maps derived from Gmsh's ordering rules are pinned by the
digests of the reference's tables (tests/golden/gmsh_nodemaps.json)."""
import hashlib
import io
import json
import os
import re

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
    assert m.etypes == ['pri', 'tet']          # name order, as loaded
    for t in m.etypes:
        assert np.allclose(m.spts[t], ref.spts[t], rtol=0, atol=1e-12)
    assert _pairs(m) == _pairs(ref)
    assert m.bcon == {}


@pytest.mark.parametrize('version', ['2.2', '4.1'])
@pytest.mark.parametrize('layers,k', [([('pyr', 2)], 1),
                                      ([('hex', 1), ('pyr', 1)], 2),
                                      ([('tet', 1), ('pyt', 1)], 3),
                                      ([('pyt', 1), ('pri', 1)], 4)])
def test_pyramid_round_trip(version, layers, k):
    """Pyramids alone and with hexes, tets and prisms, first order and
    warped of shape order k (quad faces with hofaces at k = 3): shape
    points, element types and pairs as generated."""
    L = 2*np.pi
    blk = [((2, 2, 2), L/2, layers)]
    flat = SimplexBoxMesh(blk, sorder=k)
    sides = _sim_sides(flat, (L, L, L))
    ref = SimplexBoxMesh(blk, sorder=k, warp=_swarp if k > 1 else None)
    m = GmshMesh(io.StringIO(mesh_to_msh(ref, sides, version,
                                         hofaces=k == 3)))
    assert m.etypes == sorted(ref.spts)
    for t in m.etypes:
        assert np.allclose(m.spts[t], ref.spts[t], rtol=0, atol=1e-12)
    assert _pairs(m) == _pairs(ref)


def _tiny(version='4.1'):
    raw = BoxMesh(1, periodic=(False, False, False))
    spec = {s: 'wall' for s in ('x0', 'x1', 'y0', 'y1', 'z0', 'z1')}
    return mesh_to_msh(raw, box_sides(raw, spec), version)


def test_declines_and_errors():
    txt = _tiny('2.2')
    m = GmshMesh(io.StringIO(txt))
    assert len(m.bcon['wall']) == 6 and m.con == ([], [])
    # a pyramid (Gmsh type 7) whose base is not a parallelogram (the
    # hex's Gmsh corners 0, 1, 2, 4) is declined, as by the reference; a
    # second-order hex (type 12) with 8 nodes is malformed
    hx = re.search(r' 5 2 1 1 ((\d+ ){7}\d+)', txt)[1].split()
    bad = ' '.join(hx[i] for i in (0, 1, 2, 4, 6))
    with pytest.raises(ValueError, match='parallelogram'):
        GmshMesh(io.StringIO(txt.replace(' 5 2 1 1 ' + ' '.join(hx),
                                         ' 7 2 1 1 ' + bad, 1)))
    with pytest.raises(ValueError, match='node count'):
        GmshMesh(io.StringIO(txt.replace(' 5 2 1 1 ', ' 12 2 1 1 ')))
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


GOLDEN = os.path.join(os.path.dirname(__file__), 'golden',
                      'gmsh_nodemaps.json')


def test_nodemaps_pinned():
    """Every node map the reference tabulates for these shapes (orders 1-5
    hexes and quads, 1-6 tets, prisms and pyramids, 1-4 triangles) equals
    the one derived from Gmsh's ordering rules."""
    from pyfr_amd.gmsh import gmsh_nodemap, npts_order
    with open(GOLDEN) as f:
        want = json.load(f)
    assert len(want) == 32
    for key, dig in want.items():
        kind, n = key.split('-')
        m = gmsh_nodemap(kind, npts_order(kind, int(n)))
        assert sorted(m) == list(range(int(n)))
        got = hashlib.sha256(np.asarray(m, dtype=np.int64).tobytes())
        assert got.hexdigest() == dig, key


def test_nodemap_rules():
    """What Gmsh's ordering fixes independently of any table: corners
    first, then each edge's k - 1 nodes from its first corner to its
    second (tet edge 0 joins corners 0 and 1; hex edge 1 joins 0 and 3),
    the face and volume interiors last."""
    from pyfr_amd.gmsh import _vertex_index, gmsh_nodemap, std_ele
    for kind, nv in (('tet', 4), ('hex', 8), ('pri', 6)):
        for k in (1, 2, 3, 4):
            m = np.asarray(gmsh_nodemap(kind, k))
            vi = _vertex_index(kind, k)
            assert sorted(m[vi]) == list(range(nv))
    def at(kind, k, x):
        p = std_ele(kind, k)
        return gmsh_nodemap(kind, k)[int(np.abs(p - x).max(axis=1).argmin())]
    assert at('tet', 3, (-1/3, -1, -1)) == 4
    assert at('tet', 3, (1/3, -1, -1)) == 5
    assert at('hex', 3, (-1, -1/3, -1)) == 10
    assert at('hex', 3, (-1, 1/3, -1)) == 11
    # interior nodes come last: the hex's centre-most nodes at k = 4
    p, m = std_ele('hex', 4), np.asarray(gmsh_nodemap('hex', 4))
    inner = np.all(np.abs(p) < 0.99, axis=1)
    assert sorted(m[inner]) == list(range(125 - 27, 125))


def _hex_spec(periodic):
    spec = {}
    for a, ax in enumerate('xyz'):
        if periodic[a]:
            spec[f'{ax}0'], spec[f'{ax}1'] = (f'periodic_{ax}_l',
                                              f'periodic_{ax}_r')
        else:
            spec[f'{ax}0'], spec[f'{ax}1'] = f'{ax}lo', f'{ax}hi'
    return spec


@pytest.mark.parametrize('version', ['2.2', '4.1'])
@pytest.mark.parametrize('k,hofaces', [(2, False), (3, True), (4, True)])
def test_high_order_hex_round_trip(version, k, hofaces):
    """Warped hexes of shape order k (curved for x < pi, linear beyond):
    shape points, curved flags and order, pairs and boundaries."""
    from tests.test_curved import warp
    n, periodic = 3, (True, False, True)
    ref = BoxMesh(n, periodic=periodic, sorder=k, warp=warp)
    raw = BoxMesh(n, periodic=(False, False, False), sorder=k, warp=warp)
    assert 0 < ref.spts_curved['hex'].sum() < n**3
    txt = mesh_to_msh(raw, box_sides(raw, _hex_spec(periodic)), version,
                      hofaces=hofaces)
    m = GmshMesh(io.StringIO(txt))
    assert m.etypes == ['hex'] and m.sorders == {'hex': k}
    assert np.array_equal(m.spts_curved['hex'], ref.spts_curved['hex'])
    assert np.allclose(m.spts['hex'], ref.spts['hex'], rtol=0, atol=1e-12)
    assert _pairs(m) == _pairs(ref)
    assert sorted(m.bcon['ylo']) == sorted(ref.bcon['y0'])
    assert sorted(m.bcon['yhi']) == sorted(ref.bcon['y1'])


def test_high_order_hex_curved_first():
    """Elements written linear-first come back curved-first (stable), the
    interfaces renumbered with them."""
    from tests.test_curved import warp
    n = 3
    ref = BoxMesh(n, periodic=(False, False, False), sorder=2, warp=warp)
    c = ref.spts_curved['hex']
    # the same mesh with its element order reversed
    rev = BoxMesh(n, periodic=(False, False, False), sorder=2, warp=warp)
    E = n**3
    rev.spts = {'hex': ref.spts['hex'][:, ::-1].copy()}
    rev.bcon = {s: [(t, E - 1 - e, f) for t, e, f in fs]
                for s, fs in ref.bcon.items()}
    m = GmshMesh(io.StringIO(mesh_to_msh(
        rev, box_sides(rev, {s: s for s in rev.bcon}), '4.1')))
    assert m.spts_curved['hex'].sum() == c.sum()
    assert m.spts_curved['hex'][:c.sum()].all()
    # file element e (reversed order) sits at new[e]
    fc = c[::-1]
    new = np.empty(E, dtype=int)
    new[np.argsort(~fc, kind='stable')] = np.arange(E)
    assert np.allclose(m.spts['hex'][:, new], rev.spts['hex'], atol=1e-12)
    for s in ref.bcon:
        assert sorted(m.bcon[s]) == sorted((t, int(new[E - 1 - e]), f)
                                           for t, e, f in ref.bcon[s])
    pairs = {(a[1], a[2], b[1], b[2]) for a, b in zip(*m.con)}
    assert len(pairs) == len(ref.con[0])
    for a, b in zip(*ref.con):
        ea, eb = int(new[E - 1 - a[1]]), int(new[E - 1 - b[1]])
        assert ((ea, a[2], eb, b[2]) in pairs or
                (eb, b[2], ea, a[2]) in pairs)


def _sim_sides(mesh, L):
    """Faces of an (unwarped) SimplexBoxMesh block on the six box sides
    as periodic pairs."""
    from pyfr_amd.gmsh import _vertex_index, npts_order
    out = {}
    for t in mesh.spts:
        pts = mesh.spts[t]                         # (nsp, neles, 3)
        vi = _vertex_index(t, npts_order(t, pts.shape[0]))
        for f, fv in enumerate(FACES[t]):
            c = pts[[vi[v] for v in fv]].mean(axis=0)
            for a, ax in enumerate('xyz'):
                for s, x in ((0, 0.0), (1, L[a])):
                    on = np.flatnonzero(np.isclose(c[:, a], x))
                    name = f'periodic_{ax}_{"lr"[s]}'
                    out.setdefault(name, []).extend(
                        (t, int(e), f) for e in on)
    return out


def _swarp(x, y, z):
    return (x + 0.12*np.sin(y)*np.sin(z), y + 0.08*np.sin(x)*np.cos(z),
            z + 0.05*np.sin(x + y))


@pytest.mark.parametrize('version', ['2.2', '4.1'])
@pytest.mark.parametrize('k', [2, 3, 4])
def test_high_order_simplex_round_trip(version, k):
    """Warped tets and prisms of shape order k (every element curved):
    shape points in order, pairs with the reference's orientation."""
    n, nt, npr = 2, 1, 1
    L = 2*np.pi
    flat = SimplexBoxMesh.stacked(n, nt, npr, sorder=k)
    sides = _sim_sides(flat, (L, L, (nt + npr)*L/n))
    ref = SimplexBoxMesh.stacked(n, nt, npr, sorder=k, warp=_swarp)
    m = GmshMesh(io.StringIO(mesh_to_msh(ref, sides, version,
                                         hofaces=k == 3)))
    assert m.etypes == ['pri', 'tet'] and m.sorders == {'tet': k, 'pri': k}
    for t in m.etypes:
        assert np.allclose(m.spts[t], ref.spts[t], rtol=0, atol=1e-12)
    assert _pairs(m) == _pairs(ref)
    assert not hasattr(m, 'spts_curved')


def test_high_order_simplex_linear_snaps_to_corners():
    """Order-2 tets and prisms on their affine maps are linear: the mesh
    keeps their corners only, as a first-order file would give."""
    n, nt, npr = 2, 1, 1
    L = 2*np.pi
    ref = SimplexBoxMesh.stacked(n, nt, npr)
    raw = SimplexBoxMesh.stacked(n, nt, npr, sorder=2)
    sides = _sim_sides(raw, (L, L, (nt + npr)*L/n))
    m = GmshMesh(io.StringIO(mesh_to_msh(raw, sides, '4.1')))
    for t in m.etypes:
        assert np.allclose(m.spts[t], ref.spts[t], rtol=0, atol=1e-12)
    assert _pairs(m) == _pairs(ref)


def test_high_order_simplex_lintol():
    """A warp of 1e-7 of the element size is below the default lintol of
    1e-5 (snapped: corners only), one of 1e-3 above it (curved); lintol
    moves the split."""
    n, nt, npr = 2, 1, 1
    L = 2*np.pi
    flat = SimplexBoxMesh.stacked(n, nt, npr, sorder=2)
    sides = _sim_sides(flat, (L, L, (nt + npr)*L/n))
    for eps, curved in ((1e-7, False), (1e-3, True)):
        def w(x, y, z, eps=eps):
            return (x + eps*np.sin(y)*np.sin(z), y, z)
        raw = SimplexBoxMesh.stacked(n, nt, npr, sorder=2, warp=w)
        txt = mesh_to_msh(raw, sides, '4.1')
        m = GmshMesh(io.StringIO(txt))
        assert (m.spts['tet'].shape[0] == 10) == curved
        m = GmshMesh(io.StringIO(txt), lintol=1e-9)
        assert m.spts['tet'].shape[0] == 10


def test_high_order_errors():
    n = 2
    raw = BoxMesh(n, periodic=(False, False, False), sorder=2)
    spec = {s: 'wall' for s in ('x0', 'x1', 'y0', 'y1', 'z0', 'z1')}
    txt = mesh_to_msh(raw, box_sides(raw, spec), '2.2')
    m = GmshMesh(io.StringIO(txt))
    assert m.spts['hex'].shape == (8, n**3, 3)
    # one hex of another order (first order, its corners) is refused
    lines = txt.split('\n')
    i = next(j for j, ln in enumerate(lines) if ln.split()[1:2] == ['12'])
    f = lines[i].split()
    g = np.asarray(f[5:], dtype=int)
    from pyfr_amd.gmsh import gmsh_nodemap
    lines[i] = ' '.join(f[:1] + ['5'] + f[2:5] +
                        [str(x) for x in g[:8]])
    with pytest.raises(NotImplementedError, match='shape order'):
        GmshMesh(io.StringIO('\n'.join(lines)))
    assert gmsh_nodemap('hex', 2)[:2] == [0, 8]
