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
  physical entities) sorted, as the reference's volume merge stacks them.

Scope: first-order elements (Gmsh types 4, 5, 6; surfaces 2 and 3).
Higher-order node orderings (curved elements come from the generated
meshes), pyramids and 2-D meshes are declined.  The ``.pyfrm`` (HDF5) writer is not built: h5py is not
available here, and the solver takes the mesh object directly.
"""
from collections import defaultdict
import re

import numpy as np

from pyfr_amd.mesh import FACE_VERTS, fuzzysort_batched

# Gmsh element type -> (PyFR type, node count); first order only
GMSH_TYPES = {2: ('tri', 3), 3: ('quad', 4), 4: ('tet', 4), 5: ('hex', 8),
              6: ('pri', 6)}
# Gmsh corner-node order -> PyFR vertex order (pyfr/readers/gmsh.py:46-150)
NODEMAP = {'tet': [0, 1, 2, 3], 'hex': [0, 1, 3, 2, 4, 5, 7, 6],
           'pri': [0, 1, 2, 3, 4, 5], 'tri': [0, 1, 2], 'quad': [0, 1, 3, 2]}
# Higher-order / other types the reference reads but this build declines
_DECLINED = {8, 26, 27, 9, 21, 23, 10, 36, 37, 11, 29, 30, 12, 92, 93,
             13, 90, 91, 7, 14, 118, 119}


FACES = FACE_VERTS
# codec order of the volume element types (Gmsh type number)
ETYPE_ORDER = ['tet', 'hex', 'pri']


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

    def __init__(self, msh):
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
        if gtype in (1, 15):
            return              # first-order lines and points: not used
        if gtype in _DECLINED:
            raise NotImplementedError(
                f'Gmsh element type {gtype} is not supported (first-order '
                'tets, hexes and prisms only)')
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
        if not self._pnames:
            raise ValueError('Mesh has no physical names')
        voldim = max(d for d, _, _ in self._pnames)
        if voldim != 3:
            raise NotImplementedError('Only 3-D meshes are supported')
        vol, bnd, per = set(), {}, defaultdict(dict)
        for dim, pent, name in self._pnames:
            if name.startswith('periodic'):
                m = re.match(r'periodic[ _-]([a-z0-9]+)[ _-](l|r)$', name)
                if not m:
                    raise ValueError('Invalid periodic boundary condition')
                per[m[1]][m[2]] = pent
            elif dim == voldim:
                vol.add(pent)
            else:
                bnd[pent] = name
        if not vol:
            raise ValueError('No volume elements in mesh')
        if any(set(p) != {'l', 'r'} for p in per.values()):
            raise ValueError('Unpaired periodic boundary in mesh')

        # volume elements per PyFR type, (gmsh type, pents) groups sorted
        spts, order = {}, []
        vnodes = defaultdict(list)
        surf = defaultdict(list)
        for (gtype, pents), rows in sorted(self._elenodes.items()):
            petype = GMSH_TYPES[gtype][0]
            arr = np.array(rows, dtype=np.int64) - self._nodeoff
            arr = arr[:, NODEMAP[petype]]
            if any(p in vol for p in pents):
                vnodes[petype].append(arr)
            else:
                for p in pents:
                    surf[p].append(arr)
        self.etypes = [t for t in ETYPE_ORDER if t in vnodes]
        if set(vnodes) - set(self.etypes):
            raise NotImplementedError('Unsupported volume element type')
        self._vnodes = {t: np.vstack(vnodes[t]) for t in self.etypes}
        for t, nd in self._vnodes.items():
            pts = self._nodes[nd]                         # (neles, nv, 3)
            if np.isnan(pts).any():
                raise ValueError('Element references an undefined node')
            spts[t] = pts.swapaxes(0, 1).copy()
        self.spts = spts
        self.neles = sum(v.shape[0] for v in self._vnodes.values())

        # every face of every volume element keyed by its sorted nodes
        owner = {}
        pairs = []
        cbase = {}
        c = 0
        for t in self.etypes:
            cbase[t] = c
            c += len(FACES[t])
        for t in self.etypes:
            nd = self._vnodes[t]
            for f, fv in enumerate(FACES[t]):
                keys = np.sort(nd[:, list(fv)], axis=1)
                for e, k in enumerate(map(tuple, keys)):
                    other = owner.pop(k, None)
                    if other is None:
                        owner[k] = (t, e, f)
                    else:
                        pairs.append((other, (t, e, f)))

        # periodic faces: fuzzy-sorted centroids of each side
        # (pyfr/readers/base.py:263-292)
        def by_ftype(rows):
            out = defaultdict(list)
            for r in rows:
                out[r.shape[1]].append(r)
            return {w: np.vstack(v) for w, v in out.items()}

        for k, lr in sorted(per.items()):
            lft, rgt = by_ftype(surf.pop(lr['l'])), by_ftype(surf.pop(lr['r']))
            if set(lft) != set(rgt):
                raise ValueError(f'Periodic boundary {k} is unbalanced')
            for w in sorted(lft):
                sides = []
                for fn in (lft[w], rgt[w]):
                    cen = self._nodes[fn].mean(axis=1)
                    idx = fuzzysort_batched(cen.T[None])[0]
                    sides.append([owner.pop(tuple(sorted(fn[i])))
                                  for i in idx])
                if len(sides[0]) != len(sides[1]):
                    raise ValueError(f'Periodic boundary {k} is unbalanced')
                pairs += list(zip(*sides))

        # named boundaries
        self.bcon = {}
        for pent, rows in sorted(surf.items()):
            if pent not in bnd:
                continue
            faces = []
            for fns in rows:
                for fn in fns:
                    faces.append(owner.pop(tuple(sorted(fn))))
            self.bcon[bnd[pent]] = faces
        if owner:
            raise ValueError('Unpaired faces in mesh')

        def key(q):
            t, e, f = q
            return (cbase[t] + f, e)
        lhs, rhs = [], []
        for a, b in pairs:
            if key(b) < key(a):
                a, b = b, a
            lhs.append(a)
            rhs.append(b)
        srt = sorted(range(len(lhs)), key=lambda i: key(lhs[i]))
        self.con = ([lhs[i] for i in srt], [rhs[i] for i in srt])
