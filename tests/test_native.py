"""The native mesh file (CPU): meshes imported from Gmsh written as .pyfrm
(format version 2, through the HDF5 C library) and read back give the same
solver mesh -- shape points, curved flags and order, interface pairs with
their orientation and order, boundaries; the file holds the reference's
datasets (``pyfr/readers/base.py:67-105, 384-396``) with hand-checkable
contents on a two-hex mesh; the element colouring gives neighbours
different colours."""
import io

import numpy as np
import pytest

from pyfr_amd.gmsh import GmshMesh, std_ele
from pyfr_amd.h5 import File
from pyfr_amd.mesh import BoxMesh, SimplexBoxMesh
from pyfr_amd.native import NativeMesh, RawMesh
from tests.msh_files import box_sides, mesh_to_msh
from tests.test_curved import warp
from tests.test_gmsh import _hex_spec, _sim_sides, _swarp


def _same(a, b):
    assert a.etypes == b.etypes
    for t in a.etypes:
        assert np.array_equal(a.spts[t], b.spts[t])
    assert (getattr(a, 'spts_curved', {}).keys() ==
            getattr(b, 'spts_curved', {}).keys())
    for t, c in getattr(a, 'spts_curved', {}).items():
        assert np.array_equal(c, b.spts_curved[t])
    assert a.con == b.con
    assert a.bcon == b.bcon


@pytest.mark.parametrize('k', [1, 3])
def test_hex_round_trip(tmp_path, k):
    raw = BoxMesh(3, periodic=(False, False, False), sorder=k,
                  warp=warp if k > 1 else None)
    spec = _hex_spec((True, False, False))
    g = GmshMesh(io.StringIO(mesh_to_msh(raw, box_sides(raw, spec))))
    fn = str(tmp_path/'box.pyfrm')
    g.write_pyfrm(fn)
    m = NativeMesh(fn)
    _same(g, m)
    assert m.uuid == g.raw.uuid and m.codec == g.codec


def test_simplex_round_trip(tmp_path):
    n, nt, npr = 2, 1, 1
    L = 2*np.pi
    flat = SimplexBoxMesh.stacked(n, nt, npr, sorder=2)
    sides = _sim_sides(flat, (L, L, (nt + npr)*L/n))
    gen = SimplexBoxMesh.stacked(n, nt, npr, sorder=2, warp=_swarp)
    g = GmshMesh(io.StringIO(mesh_to_msh(gen, sides, '2.2')))
    fn = str(tmp_path/'tp.pyfrm')
    g.write_pyfrm(fn)
    _same(g, NativeMesh(fn))


@pytest.mark.parametrize('layers,k', [([('hex', 1), ('pyr', 1)], 1),
                                      ([('tet', 1), ('pyt', 1)], 2)])
def test_pyramid_round_trip(tmp_path, layers, k):
    L = 2*np.pi
    blk = [((2, 2, 2), L/2, layers)]
    sides = _sim_sides(SimplexBoxMesh(blk, sorder=k), (L, L, L))
    gen = SimplexBoxMesh(blk, sorder=k, warp=_swarp if k > 1 else None)
    g = GmshMesh(io.StringIO(mesh_to_msh(gen, sides)))
    assert 'pyr' in g.etypes
    fn = str(tmp_path/'pyr.pyfrm')
    g.write_pyfrm(fn)
    _same(g, NativeMesh(fn))


def test_file_contents(tmp_path):
    """Two hexes stacked in z, walls on every side: the datasets and the
    face links by hand."""
    raw = BoxMesh((1, 1, 2), periodic=(False, False, False))
    spec = {s: 'wall' for s in ('x0', 'x1', 'y0', 'y1')}
    spec |= {'z0': 'bottom', 'z1': 'top'}
    g = GmshMesh(io.StringIO(mesh_to_msh(raw, box_sides(raw, spec))))
    fn = str(tmp_path/'two.pyfrm')
    g.write_pyfrm(fn)
    with File(fn) as f:
        assert int(f['version'][()]) == 2
        codec = [c.decode() for c in f['codec']]
        assert codec[:7] == ['eles/hex'] + [f'eles/hex/face/{i}'
                                            for i in range(6)]
        assert set(codec[7:]) == {'bc/wall', 'bc/bottom', 'bc/top',
                                  'tag/fluid'}
        e = f['eles/hex']
        assert e.dtype.names == ('nodes', 'curved', 'faces', 'colour', 'tags')
        assert e['nodes'].shape == (2, 8) and not e['curved'].any()
        assert np.array_equal(f.attr('eles/hex', 'pts'), std_ele('hex', 1))
        nodes = f['nodes']
        assert nodes.dtype.names == ('location', 'valency')
        # vertices of the lower hex lie at z <= the upper hex's
        loc = nodes['location'][e['nodes']]
        lo = int(np.argmin(loc[:, :, 2].mean(axis=1)))
        hi = 1 - lo
        # face 5 (z = +1) of the lower hex links face 0 (z = -1) of the
        # upper one, and back; their other faces are boundaries
        fc = e['faces']
        assert (fc['cidx'][lo, 5], fc['off'][lo, 5]) == (1, hi)
        assert (fc['cidx'][hi, 0], fc['off'][hi, 0]) == (6, lo)
        bot = codec.index('bc/bottom')
        assert (fc['cidx'][lo, 0], fc['off'][lo, 0]) == (bot, -1)
        assert np.all(fc['off'][lo, 1:5] == -1)
        # the middle layer's 4 nodes are shared by both hexes
        assert sorted(np.bincount(nodes['valency'])[1:].tolist()) == [4, 8]
        assert e['colour'][lo] != e['colour'][hi]
        assert np.all(e['tags'] == 1)
        parts = f['partitionings/1/eles']
        assert np.array_equal(parts, [0, 1])
        assert np.array_equal(f.attr('partitionings/1/eles', 'regions'),
                              [[0, 2]])
    m = NativeMesh(fn)
    assert len(m.con[0]) == 1
    assert sorted(m.bcon) == ['bottom', 'top', 'wall']
    assert len(m.bcon['wall']) == 8


def test_partitioning_order(tmp_path):
    """A file whose curved elements are not first in file order loads them
    first (the partitioning's order), interfaces renumbered and oriented by
    the local numbers."""
    raw = BoxMesh(3, periodic=(False, False, False), sorder=2, warp=warp)
    g = GmshMesh(io.StringIO(mesh_to_msh(
        raw, box_sides(raw, {s: 'w' for s in raw.bcon}))))
    r = g.raw
    E = len(r.eles['hex'])
    # reverse the file order of the elements, links included
    rev = r.eles['hex'][::-1].copy()
    off = rev['faces']['off']
    rev['faces']['off'] = np.where(off >= 0, E - 1 - off, off)
    fn = str(tmp_path/'rev.pyfrm')
    RawMesh(r.nodes, {'hex': rev}, r.codec, r.pmap).write(fn)
    m = NativeMesh(fn)
    c = m.spts_curved['hex']
    assert c[:c.sum()].all() and not c[c.sum():].any()
    assert sorted(map(tuple, m.spts['hex'].reshape(-1, 3*E).T.reshape(
        E, -1).tolist())) == sorted(map(tuple, g.spts['hex'].reshape(
            -1, 3*E).T.reshape(E, -1).tolist()))
    for a, b in zip(*m.con):
        ka = (r.codec.index(f'eles/{a[0]}/face/{a[2]}'), a[1])
        kb = (r.codec.index(f'eles/{b[0]}/face/{b[2]}'), b[1])
        assert ka < kb


def test_colouring_neighbours_differ(tmp_path):
    n, nt, npr = 2, 1, 1
    L = 2*np.pi
    flat = SimplexBoxMesh.stacked(n, nt, npr)
    sides = _sim_sides(flat, (L, L, (nt + npr)*L/n))
    g = GmshMesh(io.StringIO(mesh_to_msh(flat, sides)))
    eles = g.raw.eles
    codec = g.raw.codec
    for t, e in eles.items():
        for i in range(len(e)):
            for cidx, off in zip(e['faces']['cidx'][i], e['faces']['off'][i]):
                if off < 0:
                    continue
                t2 = codec[cidx].split('/')[1]
                if (t2, off) != (t, i):
                    assert eles[t2]['colour'][off] != e['colour'][i]
