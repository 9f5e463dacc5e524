"""Gmsh ``.msh`` import (ASCII formats 2.2 and 4.1) into the in-memory mesh
the solver layer consumes (``spts``, ``con``, ``bcon``), following the
reference's reader and nodal assembler (``pyfr/readers/gmsh.py:28-489``,
``pyfr/readers/base.py:107-482``) and its connectivity rules
(``pyfr/readers/native.py:432-495``):

* physical names classify entities: the highest-dimensional ones are
  volumes (merged into one fluid region), ``periodic_<k>_l`` /
  ``periodic_<k>_r`` surface pairs are periodic, every other named surface
  is a boundary ``<name>`` (names lower-cased);
* volume faces pair by their sorted vertex numbers; periodic faces pair by
  the fuzzy sort of their centroids on each side; boundary faces by their
  vertex numbers;
* the left side of an interface is the face with the smaller (face codec
  index, element index) key, where the codec lists element types in Gmsh
  type-number order (tet 4, hex 5, pri 6) and their faces in order;
* elements of a type keep their file order, grouped by (Gmsh type,
  physical entities) sorted, as the reference's volume merge stacks them;
* high-order elements (shape orders 2-4: Gmsh types 11/29/30 tets,
  12/92/93 hexes, 13/90/91 prisms; faces 9/21/23 and 10/36/37) take their
  nodes to PyFR's shape-point order through node maps derived from Gmsh's
  ordering rules (``gmsh_nodemap``; the reference's tables are
  pyfr/readers/gmsh.py:58-150); connectivity uses their corner nodes
  (pyfr/readers/base.py:148-163);
* elements within ``lintol`` of the first-order map of their corners are
  snapped to it and count as linear, the others as curved, which come
  first within their type (pyfr/readers/base.py:22-45, 445-482).  Hexes
  carry the flags in ``spts_curved``; a tet or prism type with any curved
  element keeps every element's shape points (the solver then takes the
  per-point metric of all of them), one without keeps the corners only.

Scope: 3-D meshes of tets, hexes and prisms, one shape order per element
type.  Pyramids and 2-D meshes are declined.  The ``.pyfrm`` (HDF5) writer
is not built: h5py is not available here, and the solver takes the mesh
object directly.
"""
from collections import defaultdict
import re

import numpy as np

from pyfr_amd.mesh import FACE_VERTS, fuzzysort_batched
from pyfr_amd.native import RawMesh, element_colouring, ele_dtype

# Gmsh element type -> (PyFR type, node count): the types the reference
# reads (pyfr/readers/gmsh.py:34-42) for the shapes this solver has
GMSH_TYPES = {2: ('tri', 3), 9: ('tri', 6), 21: ('tri', 10), 23: ('tri', 15),
              3: ('quad', 4), 10: ('quad', 9), 36: ('quad', 16),
              37: ('quad', 25),
              4: ('tet', 4), 11: ('tet', 10), 29: ('tet', 20),
              30: ('tet', 35),
              5: ('hex', 8), 12: ('hex', 27), 92: ('hex', 64),
              93: ('hex', 125),
              6: ('pri', 6), 13: ('pri', 18), 90: ('pri', 40),
              91: ('pri', 75),
              7: ('pyr', 5), 14: ('pyr', 14), 118: ('pyr', 30),
              119: ('pyr', 55)}
NVERTS = {'tri': 3, 'quad': 4, 'tet': 4, 'hex': 8, 'pri': 6, 'pyr': 5}

# Gmsh's reference-element conventions, in its corner numbering: the edges
# (directed: edge nodes run from the first corner to the second) and the
# faces (their interior nodes follow the face's corner order)
_G_EDGES = {
    'tri': [(0, 1), (1, 2), (2, 0)],
    'quad': [(0, 1), (1, 2), (2, 3), (3, 0)],
    'tet': [(0, 1), (1, 2), (2, 0), (3, 0), (3, 2), (3, 1)],
    'hex': [(0, 1), (0, 3), (0, 4), (1, 2), (1, 5), (2, 3), (2, 6), (3, 7),
            (4, 5), (4, 7), (5, 6), (6, 7)],
    'pri': [(0, 1), (0, 2), (0, 3), (1, 2), (1, 4), (2, 5), (3, 4), (3, 5),
            (4, 5)],
    'pyr': [(0, 1), (0, 3), (0, 4), (1, 2), (1, 4), (2, 3), (2, 4), (3, 4)]}
_G_FACES = {
    'tri': [], 'quad': [],
    'tet': [(0, 2, 1), (0, 1, 3), (0, 3, 2), (3, 1, 2)],
    'hex': [(0, 3, 2, 1), (0, 1, 5, 4), (0, 4, 7, 3), (1, 2, 6, 5),
            (2, 3, 7, 6), (4, 5, 6, 7)],
    'pri': [(0, 2, 1), (3, 4, 5), (0, 1, 4, 3), (0, 3, 5, 2), (1, 2, 5, 4)],
    'pyr': [(0, 1, 4), (3, 0, 4), (1, 2, 4), (2, 3, 4), (0, 3, 2, 1)]}
# Gmsh corner j -> PyFR vertex: NODEMAP[kind][i] is the Gmsh corner at
# PyFR vertex i (std_ele(1) order)
NODEMAP = {'tet': [0, 1, 2, 3], 'hex': [0, 1, 3, 2, 4, 5, 7, 6],
           'pri': [0, 1, 2, 3, 4, 5], 'pyr': [0, 1, 3, 2, 4],
           'tri': [0, 1, 2], 'quad': [0, 1, 3, 2]}


def _inset(kind, V, k):
    """Corners of the sub-element holding an order-k element's interior
    nodes: each corner moved 1/k of the way along each of its edges."""
    out = []
    for i, v in enumerate(V):
        nb = [b if a == i else a for a, b in _G_EDGES[kind] if i in (a, b)]
        out.append(v + sum(V[j] - v for j in nb)/k)
    return out


def _gmsh_points(kind, k, V):
    """Reference coordinates of an order-k Gmsh element's nodes in Gmsh
    order, given its corners V in Gmsh corner order.  Gmsh orders nodes
    recursively: corners; the k - 1 nodes of each edge; each face's
    interior as an element of that face's shape and order k - 3 (triangles)
    or k - 2 (quadrilaterals) on the inset face; the volume's interior as an
    order k - 2 hex or order k - 4 tet on the inset corners, or, for prisms,
    the interior of the bottom triangle (outer) times the interior of a
    vertical edge (inner, as an order k - 2 line: both ends, then the
    rest).  Order 0 is the single centre point."""
    if k == 0:
        return [sum(V)/len(V)]
    pts = list(V)
    for a, b in _G_EDGES[kind]:
        pts += [V[a] + (V[b] - V[a])*i/k for i in range(1, k)]
    if kind in ('tri', 'quad'):
        inner = k - 3 if kind == 'tri' else k - 2
        if inner >= 0:
            pts += _gmsh_points(kind, inner, _inset(kind, V, k))
        return pts
    for fv in _G_FACES[kind]:
        fk = 'tri' if len(fv) == 3 else 'quad'
        inner = k - 3 if fk == 'tri' else k - 2
        if inner >= 0:
            pts += _gmsh_points(fk, inner, _inset(fk, [V[i] for i in fv], k))
    if kind == 'hex' and k >= 2:
        pts += _gmsh_points('hex', k - 2, _inset('hex', V, k))
    elif kind == 'tet' and k >= 4:
        pts += _gmsh_points('tet', k - 4, _inset('tet', V, k))
    elif kind == 'pyr' and k >= 3:
        # an order k - 3 pyramid: base corners inset as for the other
        # shapes, the apex moved 2/k of the way towards the base centre
        inner = _inset('pyr', V, k)
        inner[4] = V[4] + (sum(V[:4])/4 - V[4])*2/k
        pts += _gmsh_points('pyr', k - 3, inner)
    elif kind == 'pri' and k >= 3:
        tri = _gmsh_points('tri', k - 3, _inset('tri', V[:3], k))
        zs = [i for i in range(1, k)]
        zs = [zs[0], zs[-1], *zs[1:-1]] if len(zs) > 1 else zs
        z0, z1 = V[0][2], V[3][2]
        pts += [np.array([t[0], t[1], z0 + (z1 - z0)*i/k])
                for t in tri for i in zs]
    return pts


def std_ele(kind, k):
    """PyFR's equispaced order-k shape points of a kind
    (pyfr/shapes.py: TensorProdShape.std_ele, TriShape / TetShape /
    PriShape.std_ele): tensor shapes x fastest; simplices x fastest within
    rows of falling length."""
    t = np.linspace(-1, 1, k + 1)
    if kind == 'quad':
        return np.array([(x, y) for y in t for x in t])
    if kind == 'hex':
        return np.array([(x, y, z) for z in t for y in t for x in t])
    if kind == 'tri':
        return np.array([(x, y) for i, y in enumerate(t)
                         for x in t[:k + 1 - i]])
    if kind == 'tet':
        return np.array([(x, y, z) for i, z in enumerate(t)
                         for j, y in enumerate(t[:k + 1 - i])
                         for x in t[:k + 1 - i - j]])
    if kind == 'pri':
        return np.array([(x, y, z) for z in t for i, y in enumerate(t)
                         for x in t[:k + 1 - i]])
    if kind == 'pyr':
        from pyfr_amd.simplex import PyrShape
        return PyrShape.std_ele(k)
    raise ValueError(f'Unknown element kind {kind}')


def npts_order(kind, n):
    """Shape order k of a kind with n equispaced points."""
    for k in range(1, 32):
        m = len(std_ele(kind, k))
        if m == n:
            return k
        if m > n:
            break
    raise ValueError(f'No {kind} of {n} nodes')


def gmsh_nodemap(kind, k):
    """Node map of an order-k element: entry i is the index, in a Gmsh
    element's node list, of the node at PyFR shape point i
    (``std_ele(k)``), the form of the reference's tables
    (pyfr/readers/gmsh.py:58-150), derived here from Gmsh's ordering
    rules (``_gmsh_points``) by matching reference coordinates."""
    ref = std_ele(kind, 1)
    V = [None]*len(ref)
    for i, j in enumerate(NODEMAP[kind]):
        V[j] = np.asarray(ref[i], dtype=float)
    g = np.array(_gmsh_points(kind, k, V))
    p = std_ele(kind, k)
    if len(g) != len(p):
        raise AssertionError(f'{kind} order {k}: {len(g)} Gmsh nodes')
    d = np.abs(p[:, None, :] - g[None, :, :]).max(axis=2)
    m = d.argmin(axis=1)
    if not (np.all(d[np.arange(len(p)), m] < 1e-9) and
            len(set(m.tolist())) == len(p)):
        raise AssertionError(f'{kind} order {k}: nodes do not match')
    return [int(x) for x in m]


def _vertex_index(kind, k):
    """Positions of the PyFR vertices (std_ele(1) order) in std_ele(k)."""
    p, v = std_ele(kind, k), std_ele(kind, 1)
    return [int(np.abs(p - x).max(axis=1).argmin()) for x in v]


def _linear_weights(kind, k):
    """(nspts, nverts) values of the first-order nodal basis of a kind
    (trilinear hex, affine tet, prism affine in x, y times linear in z) at
    its order-k shape points."""
    def mono(x):
        x, y, z = x[:, 0], x[:, 1], x[:, 2]
        one = np.ones_like(x)
        if kind == 'hex':
            return np.stack([one, x, y, z, x*y, x*z, y*z, x*y*z], 1)
        if kind == 'tet':
            return np.stack([one, x, y, z], 1)
        if kind == 'pyr':
            return np.stack([one, x, y, z, x*y], 1)
        return np.stack([one, x, y, z, x*z, y*z], 1)
    return mono(std_ele(kind, k)) @ np.linalg.inv(mono(std_ele(kind, 1)))


FACES = FACE_VERTS
# codec order of the volume element types (Gmsh type number)
ETYPE_ORDER = ['tet', 'hex', 'pri', 'pyr']


class _Lines:
    """Line iterator over a file with the section helpers of the format."""

    def __init__(self, f):
        self._it = (ln.strip() for ln in f)

    def __iter__(self):
        return self

    def __next__(self):
        return next(self._it)

    def ints(self):
        return [int(x) for x in next(self).split()]

    def expect(self, tag):
        ln = next(self)
        if ln != tag:
            raise ValueError(f'Expected {tag}, got {ln!r}')


class GmshMesh:
    ndims = 3

    def __init__(self, msh, lintol=1e-5):
        self.lintol = lintol
        if isinstance(msh, str):
            with open(msh) as f:
                self._read(_Lines(f))
        else:
            self._read(_Lines(msh))
        self._assemble()

    # -- parsing --------------------------------------------------------------
    def _read(self, it):
        self._version = None
        self._pnames = []
        self._tagpents = {}
        self._nodes = None
        self._elenodes = defaultdict(list)
        handlers = {'$MeshFormat': self._format,
                    '$PhysicalNames': self._phys_names,
                    '$Entities': self._entities, '$Nodes': self._read_nodes,
                    '$Elements': self._read_eles}
        for ln in it:
            if not ln:
                continue
            if not ln.startswith('$'):
                raise ValueError('Expected a mesh section')
            h = handlers.get(ln)
            if h:
                h(it)
            else:
                end = '$End' + ln[1:]
                for el in it:
                    if el == end:
                        break
                else:
                    raise ValueError(f'Expected {end}')
        if self._nodes is None or not self._elenodes:
            raise ValueError('Mesh has no nodes or elements')

    def _format(self, it):
        ver, ftype, dsize = next(it).split()
        if ver not in ('2.2', '4.1'):
            raise ValueError(f'Unsupported mesh version {ver}')
        if ftype != '0':
            raise ValueError('Only ASCII meshes are supported')
        if dsize != '8':
            raise ValueError('Invalid data size')
        self._version = ver
        it.expect('$EndMeshFormat')

    def _phys_names(self, it):
        n = int(next(it))
        for _ in range(n):
            m = re.match(r'(\d+) (\d+) "((?:[^"\\]|\\.)*)"$', next(it))
            if not m:
                raise ValueError('Malformed physical entity')
            self._pnames.append((int(m[1]), int(m[2]), m[3].lower()))
        it.expect('$EndPhysicalNames')

    def _entities(self, it):
        npts, *ents = it.ints()
        for _ in range(npts):
            next(it)
        for ndim, nent in enumerate(ents, start=1):
            for _ in range(nent):
                e = next(it).split()
                tag, nphys = int(e[0]), int(e[7])
                if nphys:
                    self._tagpents[ndim, tag] = tuple(sorted(
                        abs(int(e[8 + i])) for i in range(nphys)))
        it.expect('$EndEntities')

    def _read_nodes(self, it):
        if self._version == '2.2':
            n = int(next(it))
            rows = [next(it).split() for _ in range(n)]
            ids = np.array([int(r[0]) for r in rows])
            xyz = np.array([[float(x) for x in r[1:4]] for r in rows])
        else:
            ne, nn, lo, hi = it.ints()
            ids, xyz = [], []
            for _ in range(ne):
                cnt = int(next(it).split()[-1])
                ids += [int(next(it)) for _ in range(cnt)]
                xyz += [[float(x) for x in next(it).split()[:3]]
                        for _ in range(cnt)]
            ids, xyz = np.array(ids, dtype=np.int64), np.array(xyz)
            if len(ids) != nn:
                raise ValueError('Invalid node count')
        self._nodeoff = int(ids.min())
        nodes = np.full((int(ids.max()) - self._nodeoff + 1, 3), np.nan)
        nodes[ids - self._nodeoff] = xyz
        self._nodes = nodes
        it.expect('$EndNodes')

    def _add_ele(self, gtype, pents, nodes):
        if gtype in (1, 8, 26, 27, 15):
            return              # lines and points: not used in 3-D
        if gtype not in GMSH_TYPES:
            raise ValueError(f'Unsupported element type {gtype}')
        if len(nodes) != GMSH_TYPES[gtype][1]:
            raise ValueError('Invalid element node count')
        self._elenodes[gtype, pents].append(nodes)

    def _read_eles(self, it):
        if self._version == '2.2':
            n = int(next(it))
            for _ in range(n):
                v = it.ints()
                gtype, ntags = v[1], v[2]
                tags, nodes = v[3:3 + ntags], v[3 + ntags:]
                self._add_ele(gtype, (tags[0],), nodes)
        else:
            nb, ne, _, _ = it.ints()
            cnt = 0
            for _ in range(nb):
                edim, etag, gtype, count = it.ints()
                pents = self._tagpents.get((edim, etag), ())
                for _ in range(count):
                    self._add_ele(gtype, pents, it.ints()[1:])
                cnt += count
            if cnt != ne:
                raise ValueError('Invalid element count')
        it.expect('$EndElements')

    # -- assembly -------------------------------------------------------------
    def _assemble(self):
        """The raw mesh of the reference's importer (codec, element records
        with their per-face links, periodic pairs; pyfr/readers/gmsh.py:
        436-487, pyfr/readers/base.py:107-482), then the load path of its
        native reader (``native.RawMesh.load``)."""
        if not self._pnames:
            raise ValueError('Mesh has no physical names')
        voldim = max(d for d, _, _ in self._pnames)
        if voldim != 3:
            raise NotImplementedError('Only 3-D meshes are supported')
        # volume groups by name; boundaries and periodic pairs in the
        # order of their physical names (the first-named side of a pair is
        # its left one, as the reference takes them)
        vol, bnd, per = {}, {}, defaultdict(list)
        for dim, pent, name in self._pnames:
            if name.startswith('periodic'):
                m = re.match(r'periodic[ _-]([a-z0-9]+)[ _-](l|r)$', name)
                if not m:
                    raise ValueError('Invalid periodic boundary condition')
                per[m[1]].append(pent)
            elif dim == voldim:
                vol[name] = pent
            else:
                bnd[pent] = name
        if not vol:
            raise ValueError('No volume elements in mesh')
        if any(len(p) != 2 for p in per.values()):
            raise ValueError('Unpaired periodic boundary in mesh')
        pbits = {p: 1 << i for i, (_, p) in enumerate(sorted(vol.items()))}

        # volume elements per PyFR type, (gmsh type, pents) groups sorted;
        # the types in the order of their Gmsh type numbers (the codec's)
        vnodes, hnodes, tags = defaultdict(list), defaultdict(list), \
            defaultdict(list)
        sorders, surf, etypes = {}, defaultdict(list), []
        for (gtype, pents), rows in sorted(self._elenodes.items()):
            petype, nn = GMSH_TYPES[gtype]
            arr = np.array(rows, dtype=np.int64) - self._nodeoff
            # first-order representation: the corners, PyFR vertex order
            corners = arr[:, :NVERTS[petype]][:, NODEMAP[petype]]
            if petype == 'pyr':
                # the affine pyramid map needs a parallelogram base
                # (pyfr/readers/base.py:135-146)
                v = self._nodes[corners[:, :4]]
                if np.any(np.abs(v[:, 0] - v[:, 1] - v[:, 2] + v[:, 3])
                          > 1e-10):
                    raise ValueError('Pyramids with non-parallelogram bases '
                                     'are currently unsupported')
            mask = sum(pbits[p] for p in pents if p in pbits)
            if mask:
                if petype not in ETYPE_ORDER:
                    raise ValueError(f'{petype} elements in a 3-D volume')
                k = npts_order(petype, nn)
                if sorders.setdefault(petype, k) != k:
                    raise NotImplementedError(
                        f'{petype} elements of more than one shape order')
                if petype not in etypes:
                    etypes.append(petype)
                vnodes[petype].append(corners)
                hnodes[petype].append(arr[:, gmsh_nodemap(petype, k)])
                tags[petype].append(np.full(len(arr), mask, dtype=np.uint64))
            else:
                for p in pents:
                    surf[p].append(corners)
        self.etypes = etypes
        vn = {t: np.vstack(vnodes[t]) for t in etypes}
        hn = {t: np.vstack(hnodes[t]) for t in etypes}
        for nd in hn.values():
            if np.isnan(self._nodes[nd]).any():
                raise ValueError('Element references an undefined node')
        self.sorders = sorders
        nodes, curved = self._linearise(hn)

        # codec (pyfr/readers/base.py:363-398, gmsh.py:481-483)
        codec = []
        for t in etypes:
            codec.append(f'eles/{t}')
            codec += [f'eles/{t}/face/{i}' for i in range(len(FACES[t]))]
        codec += [f'bc/{n}' for n in bnd.values()]
        codec += [f'tag/{n}' for n in sorted(vol)]
        cix = {c: i for i, c in enumerate(codec)}
        links = {t: np.full((len(vn[t]), len(FACES[t]), 2), -2,
                            dtype=np.int64) for t in etypes}

        def link(a, b):
            (ta, ea, fa), (tb, eb, fb) = a, b
            links[ta][ea, fa] = cix[f'eles/{tb}/face/{fb}'], eb
            links[tb][eb, fb] = cix[f'eles/{ta}/face/{fa}'], ea

        # every face of every volume element keyed by its sorted nodes
        owner = {}
        for t in etypes:
            for f, fv in enumerate(FACES[t]):
                keys = np.sort(vn[t][:, list(fv)], axis=1)
                for e, k in enumerate(map(tuple, keys)):
                    other = owner.pop(k, None)
                    if other is None:
                        owner[k] = (t, e, f)
                    else:
                        link(other, (t, e, f))

        # periodic faces: fuzzy-sorted centroids of each side
        # (pyfr/readers/base.py:263-292)
        def by_ftype(rows):
            out = defaultdict(list)
            for r in rows:
                out[r.shape[1]].append(r)
            return {w: np.vstack(v) for w, v in out.items()}

        pdt = [('cidx', np.int16), ('off', np.int64)]
        pmap = {}
        for k, (lp, rp) in per.items():
            lft, rgt = by_ftype(surf.pop(lp, [])), by_ftype(surf.pop(rp, []))
            if set(lft) != set(rgt):
                raise ValueError(f'Periodic boundary {k} is unbalanced')
            plist = []
            for w in sorted(lft):
                sides = []
                for fn in (lft[w], rgt[w]):
                    cen = self._nodes[fn].mean(axis=1)
                    idx = fuzzysort_batched(cen.T[None])[0]
                    sides.append([owner.pop(tuple(sorted(fn[i])))
                                  for i in idx])
                if len(sides[0]) != len(sides[1]):
                    raise ValueError(f'Periodic boundary {k} is unbalanced')
                for a, b in zip(*sides):
                    link(a, b)
                    plist.append([(cix[f'eles/{a[0]}/face/{a[2]}'], a[1]),
                                  (cix[f'eles/{b[0]}/face/{b[2]}'], b[1])])
            pmap[k] = np.array(plist, dtype=pdt).reshape(-1, 2)

        # named boundaries
        for pent, name in bnd.items():
            for fns in surf.get(pent, []):
                for fn in fns:
                    t, e, f = owner.pop(tuple(sorted(fn)))
                    links[t][e, f] = cix[f'bc/{name}'], -1
        if owner:
            raise ValueError('Unpaired faces in mesh')

        # element records (pyfr/readers/base.py:384-396) and colours
        eles = {}
        for t in etypes:
            rec = np.zeros(len(hn[t]), dtype=ele_dtype(hn[t].shape[1],
                                                       len(FACES[t])))
            rec['nodes'] = hn[t]
            rec['curved'] = curved[t]
            rec['faces']['cidx'] = links[t][..., 0]
            rec['faces']['off'] = links[t][..., 1]
            rec['tags'] = np.concatenate(tags[t])
            eles[t] = rec
        element_colouring(eles, codec)
        self.raw = RawMesh(nodes, eles, codec, pmap)
        self.raw.load(self)
        self.codec = codec

    def write_pyfrm(self, fname):
        """Write the mesh as a .pyfrm file (format version 2), as
        ``pyfr import`` would (pyfr/readers/base.py:67-105)."""
        self.raw.write(fname)

    def _linearise(self, hn):
        """Snap the elements within ``lintol`` of the first-order map of
        their corners to it (pyfr/readers/base.py:445-482): per element and
        axis, the largest deviation of a node from that map over the
        element's extent must be below lintol.  Returns the snapped node
        array and the curved flags per type."""
        nodes = self._nodes.copy()
        curved = {}
        for t in self.etypes:
            idx, k = hn[t], self.sorders[t]
            if k == 1:
                curved[t] = np.zeros(len(idx), dtype=bool)
                continue
            e = nodes[idx]                               # (neles, nsp, 3)
            W = _linear_weights(t, k)                    # (nsp, nverts)
            le = np.einsum('pv,evd->epd', W, e[:, _vertex_index(t, k)])
            num = np.abs(e - le).max(axis=1)
            den = np.ptp(e, axis=1)
            with np.errstate(divide='ignore', invalid='ignore'):
                lin = np.all(num/den < self.lintol, axis=1)
            nodes[idx[lin]] = le[lin]
            curved[t] = ~lin
        return nodes, curved
