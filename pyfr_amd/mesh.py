"""Periodic box meshes and their interface connectivity (setup, host side).

The reference reads meshes from ``.pyfrm`` files; here the periodic boxes of
the benchmark configurations are generated directly, with the same data the
solver layer consumes:

* ``spts``: {'hex': (nspts=8, neles, 3)} vertex coordinates in ``HexShape``
  vertex order (``pyfr/shapes.py:423-432``);
* ``con``: internal interfaces as (lhs, rhs) lists of (etype, eidx, fidx),
  one entry per face pair.  Following ``_construct_con``
  (``pyfr/readers/native.py:445-495``) each pair appears once, its left side
  being the face with the smaller (face index, element index) key, in
  face-major order.

``fuzzysort_batched`` restates the reference's tolerance-grouped
lexicographic sort of flux points (``pyfr/nputil.py:120-181``) used by
``srtd_face_fpts`` (``pyfr/solvers/base/elements.py:161-172``).
"""
import numpy as np

# (normal axis, side) of each hex face, in the order of HexShape.faces
HEX_FACE_AXES = [(2, -1), (1, -1), (0, +1), (1, +1), (0, -1), (2, +1)]


class BoxMesh:
    """n^3 (or nx*ny*nz) box of hexes on [0, L)^3, periodic along the
    ``periodic`` axes.  The two sides of every other axis are boundaries
    named 'x0'/'x1', 'y0'/'y1', 'z0'/'z1' (low/high side): ``bcon[name]``
    lists their (etype, eidx, fidx) faces, the counterpart of a mesh file's
    ``bcon_<name>`` arrays.

    sorder, warp: elements of shape order k = sorder ((k + 1)^3 equispaced
    shape points, ``HexShape.std_ele``) whose points are moved by
    ``warp(x, y, z) -> (x, y, z)`` (which must respect the periodic axes).
    As the reference's reader does (``pyfr/readers/base.py:445-482``),
    elements whose points lie within ``lintol`` (relative to their extent)
    of the trilinear map of their vertices are snapped to it and count as
    linear; the others are curved.  Curved elements come first
    (``pyfr/readers/base.py:22-45``), ``spts_curved`` flags them."""
    ndims = 3

    def __init__(self, n, length=2*np.pi, origin=0.0,
                 periodic=(True, True, True), sorder=1, warp=None,
                 lintol=1e-5):
        nx, ny, nz = (n, n, n) if np.isscalar(n) else n
        self.shape = (nx, ny, nz)
        self.periodic = tuple(bool(p) for p in periodic)
        self.length = length
        h = np.array([length/nx, length/ny, length/nz])
        E = nx*ny*nz
        e = np.arange(E)
        ijk = np.stack([e % nx, (e // nx) % ny, e // (nx*ny)])
        k1 = sorder + 1
        q = np.arange(k1**3)
        off = np.stack([q % k1, (q // k1) % k1, q // (k1*k1)], 1)/sorder
        spts = origin + (ijk.T[None] + off[:, None])*h      # (nsp, E, 3)
        curved = np.zeros(E, dtype=bool)
        if warp is not None:
            spts = np.stack(warp(spts[..., 0], spts[..., 1], spts[..., 2]),
                            axis=-1)
            if sorder > 1:
                # trilinear map of the vertices at the shape points
                xi = off
                cidx = [((v & 1)*sorder) + k1*((((v >> 1) & 1)*sorder) +
                                               k1*(((v >> 2) & 1)*sorder))
                        for v in range(8)]
                w = np.ones((len(q), 8))
                for v in range(8):
                    for a in range(3):
                        b = (v >> a) & 1
                        w[:, v] *= xi[:, a] if b else 1 - xi[:, a]
                lin = np.einsum('qv,vec->qec', w, spts[cidx])
                dev = np.abs(spts - lin).max(axis=(0, 2))
                ext = np.ptp(spts, axis=0).max(axis=1)
                curved = dev >= lintol*ext
                spts[:, ~curved] = lin[:, ~curved]
        perm = np.argsort(~curved, kind='stable')
        self.spts = {'hex': np.ascontiguousarray(spts[:, perm])}
        self.spts_curved = {'hex': curved[perm]}
        self.etypes = ['hex']
        self.neles = E
        self.con = self._connect(ijk[:, perm])

    def _connect(self, ijk):
        """Pair faces through a structured key: (normal axis, periodic plane
        index, the element's indices along the two other axes)."""
        E = ijk.shape[1]
        nf = len(HEX_FACE_AXES)
        dims = np.array(self.shape)
        keys = np.empty((nf, E, 4), dtype=np.int64)
        bnd = np.zeros((nf, E), dtype=bool)
        self.bcon = {}
        for f, (nax, side) in enumerate(HEX_FACE_AXES):
            oth = [a for a in range(3) if a != nax]
            plane = ijk[nax] + (side > 0)
            if self.periodic[nax]:
                plane = plane % dims[nax]
            else:
                b = (plane == 0) | (plane == dims[nax])
                bnd[f] = b
                name = f'{"xyz"[nax]}{int(side > 0)}'
                self.bcon.setdefault(name, []).extend(
                    ('hex', int(e), f) for e in np.flatnonzero(b))
            keys[f, :, 0] = nax
            keys[f, :, 1] = plane
            keys[f, :, 2] = ijk[oth[0]]
            keys[f, :, 3] = ijk[oth[1]]
        keys = keys.reshape(nf*E, 4)
        fidx = np.repeat(np.arange(nf), E)
        eidx = np.tile(np.arange(E), nf)
        keep = ~bnd.ravel()
        keys, fidx, eidx = keys[keep], fidx[keep], eidx[keep]
        order = np.lexsort(keys.T[::-1])
        sk = keys[order]
        same = np.all(sk[1:] == sk[:-1], axis=1)
        starts = np.flatnonzero(same)
        if 2*len(starts) != len(keys) or np.any(same[1:] & same[:-1]):
            raise ValueError('Faces are not all paired exactly once')
        a, b = order[starts], order[starts + 1]
        # left = smaller (fidx, eidx) key, i.e. smaller flat index
        lft, rgt = np.minimum(a, b), np.maximum(a, b)
        srt = np.argsort(lft, kind='stable')
        lft, rgt = lft[srt], rgt[srt]
        lhs = [('hex', int(eidx[q]), int(fidx[q])) for q in lft]
        rhs = [('hex', int(eidx[q]), int(fidx[q])) for q in rgt]
        return lhs, rhs


def fuzzysort_batched(coords, tol=1e-6):
    """Per-element permutation sorting points lexicographically by x, then y,
    then z, where coordinates within ``tol`` of the first of a run count as
    equal.  coords: (neles, ndims, npts) -> (neles, npts)."""
    neles, ndims, npts = coords.shape
    out = np.empty((neles, npts), dtype=np.int64)
    for e in range(neles):
        out[e] = _fuzzysort(coords[e], list(range(npts)), 0, tol)
    return out


def _fuzzysort(arr, idx, dim, tol):
    """Recursive tolerance-grouped sort (``pyfr/nputil.py:120-136``)."""
    vals = arr[dim]
    srt = sorted(idx, key=vals.__getitem__)
    if dim + 1 >= arr.shape[0] or len(srt) <= 1:
        return srt
    out, start = [], 0
    for j in range(1, len(srt) + 1):
        if j == len(srt) or vals[srt[j]] - vals[srt[start]] >= tol:
            grp = srt[start:j]
            out.extend(_fuzzysort(arr, grp, dim + 1, tol) if len(grp) > 1
                       else grp)
            start = j
    return out


def sorted_face_fpts_structured(plocfpts, facefpts, tol=1e-6):
    """srtd_face_fpts for all faces: list over faces of (neles, nfacefpts)
    flux-point indices.  Elements whose face points have identical relative
    layouts share one sort (all elements of an affine box do); the grouping
    is checked, and any element that differs is sorted on its own."""
    out = []
    for ffpts in facefpts:
        ffpts = np.asarray(ffpts)
        c = plocfpts[ffpts].transpose(1, 2, 0)           # (E, 3, nfp)
        rel = c - c[:, :, :1]
        ref = _fuzzysort(rel[0], list(range(len(ffpts))), 0, tol)
        perm = np.tile(np.asarray(ref), (c.shape[0], 1))
        # elements whose layout differs from element 0 by more than tol
        bad = np.flatnonzero(np.abs(rel - rel[:1]).max(axis=(1, 2)) > 0.1*tol)
        for e in bad:
            perm[e] = _fuzzysort(c[e], list(range(len(ffpts))), 0, tol)
        out.append(ffpts[perm])
    return out


# Vertices of each face in TetShape / PriShape face order
# (pyfr/shapes.py:476-570 with the vertex order of their std_ele(1))
TET_FACE_VERTS = [(0, 1, 2), (0, 1, 3), (0, 2, 3), (1, 2, 3)]
PRI_FACE_VERTS = [(0, 1, 2), (3, 4, 5), (0, 1, 4, 3), (1, 2, 5, 4),
                  (0, 2, 5, 3)]
# hex faces: the vertices on each (normal axis, side) of HEX_FACE_AXES,
# in HexShape (x fastest) order
HEX_FACE_VERTS = [tuple(v for v in range(8) if ((v >> ax) & 1) == (side > 0))
                  for ax, side in HEX_FACE_AXES]
# pyramid: the base (vertices 0, 1, 3, 2 around it), then the triangles
# towards -y, +x, +y and -x (PyrShape, pyfr/shapes.py:573-585)
PYR_FACE_VERTS = [(0, 1, 3, 2), (0, 1, 4), (1, 3, 4), (2, 3, 4), (0, 2, 4)]
FACE_VERTS = {'tet': TET_FACE_VERTS, 'pri': PRI_FACE_VERTS,
              'hex': HEX_FACE_VERTS, 'pyr': PYR_FACE_VERTS}


def _kuhn_offsets():
    """Lattice offsets of the 6 tets of a unit cube (Kuhn split along the
    (0,0,0)-(1,1,1) diagonal), each ordered so that its Jacobian is
    positive: v1 - v0, v2 - v0, v3 - v0 form a right-handed frame."""
    out = []
    for perm in [(0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1),
                 (2, 1, 0)]:
        p = [np.zeros(3, dtype=int)]
        for ax in perm:
            q = p[-1].copy()
            q[ax] += 1
            p.append(q)
        if np.linalg.det(np.array([p[1] - p[0], p[2] - p[0],
                                   p[3] - p[0]])) < 0:
            p[1], p[2] = p[2], p[1]
        out.append(np.array(p))
    return np.array(out)                                # (6, 4, 3)


def _prism_offsets():
    """Two right prisms per cube, split along the (0,0)-(1,1) diagonal of
    the xy faces (conforming with the Kuhn tets' z faces)."""
    tris = [[(0, 0), (1, 0), (1, 1)], [(0, 0), (1, 1), (0, 1)]]
    return np.array([[(x, y, z) for z in (0, 1) for (x, y) in t]
                     for t in tris])                    # (2, 6, 3)


def _prism_x_offsets():
    """Two right prisms per cube extruded along x, split along the
    (0,0)-(1,1) diagonal of the yz faces: each cube's z faces are whole
    quad faces (one per prism), conforming with hexes above and below."""
    tris = [[(0, 0), (1, 0), (1, 1)], [(0, 0), (1, 1), (0, 1)]]
    return np.array([[(x, y, z) for x in (0, 1) for (y, z) in t]
                     for t in tris])                    # (2, 6, 3)


def _hex_offsets():
    """One hex per cube, vertices in HexShape order (x fastest)."""
    return np.array([[(x, y, z) for z in (0, 1) for y in (0, 1)
                      for x in (0, 1)]])                # (1, 8, 3)


def _oriented(verts, name):
    """Vertices in std_ele(1) order with a positive Jacobian: swap the
    two neighbours of vertex 0 that span the reference x and y axes
    (tets: 1, 2; pyramids: 1, 2 with 3 opposite 0 on the base)."""
    v = np.array(verts, dtype=float)
    if name == 'tet':
        a, b, c = v[1] - v[0], v[2] - v[0], v[3] - v[0]
    else:
        a, b, c = v[1] - v[0], v[2] - v[0], v[4] - (v[1] + v[2])/2
    if np.linalg.det(np.array([a, b, c])) < 0:
        v[[1, 2]] = v[[2, 1]]
    return v


def _pyr_offsets(tets=False):
    """The 6 pyramids of a unit cube with their apex at its centre, one on
    each face (bases on the cube faces, so neighbouring cells and hex or
    'prx' layers meet them on whole quad faces).  tets: the z-face
    pyramids split into two tets along the (0,0)-(1,1) face diagonal, so
    the cube's z faces conform with tet and 'pri' layers; returns
    (pyramids (4, 5, 3), tets (4, 4, 3))."""
    c = (0.5, 0.5, 0.5)
    pyrs, tts = [], []
    for ax in (2, 0, 1):
        for side in (0, 1):
            u, w = [a for a in range(3) if a != ax]
            quad = []
            for j in (0, 1):
                for i in (0, 1):
                    p = [0, 0, 0]
                    p[ax], p[u], p[w] = side, i, j
                    quad.append(p)
            if tets and ax == 2:
                # (0,0), (1,0), (1,1) and (0,0), (1,1), (0,1)
                for tri in ((0, 1, 3), (0, 3, 2)):
                    tts.append(_oriented([quad[t] for t in tri] + [c],
                                         'tet'))
            else:
                pyrs.append(_oriented(quad + [c], 'pyr'))
    return np.array(pyrs), np.array(tts)


CELL_SPLITS = {'tet': _kuhn_offsets(), 'pri': _prism_offsets(),
               'prx': _prism_x_offsets(), 'hex': _hex_offsets(),
               'pyr': _pyr_offsets()[0]}
# mixed splits: a cell split into elements of several types
MIXED_SPLITS = {'pyt': {'pyr': _pyr_offsets(True)[0],
                        'tet': _pyr_offsets(True)[1]}}
# element type of each cell split
SPLIT_ETYPE = {'tet': 'tet', 'pri': 'pri', 'prx': 'pri', 'hex': 'hex',
               'pyr': 'pyr'}


class SimplexBoxMesh:
    """Boxes of straight-sided tetrahedra, prisms and hexahedra.

    blocks: list of (shape (nx, ny, nz), spacing h, layers) where layers is
    a list of (split, nlayers) filling the block's z extent bottom to top;
    each block is periodic on its own along the axes where ``periodic`` is
    true, and the faces on its other sides are boundaries named
    'x0'/'x1', 'y0'/'y1', 'z0'/'z1' (low/high side).  Cells split into 6
    tets ('tet'), 2 prisms extruded along z ('pri') or along x ('prx'), 6
    pyramids with their apex at the cell centre ('pyr'), 4 such pyramids on
    the x and y faces and 4 tets on the z faces ('pyt'), or stay one hex
    ('hex') (``CELL_SPLITS``, ``MIXED_SPLITS``).  'pyr' layers meet hex
    and 'prx' layers on quad faces; 'pyt' layers meet tet and 'pri' layers
    on the tets' diagonal.  Tet and 'pri' layers share the
    cube diagonal, so their faces conform; 'prx' and hex layers meet on
    whole quad faces.  A tet or 'pri' layer next to a hex or 'prx' layer
    leaves faces unpaired (an error).

    ``spts``: {etype: (nverts, neles, 3)}; elements are ordered by block,
    then split index, then cell (x fastest).  ``con``: (lhs, rhs) lists of
    (etype, eidx, fidx), each face pair once, left side the smaller
    (etype index, face index, element index) key, sorted by left side.
    ``bcon``: {name: [(etype, eidx, fidx), ...]} in that key order.

    sorder, warp: order-k shape points (``std_ele(k)`` order) of each
    element's affine (hex: trilinear) map, moved by ``warp(x, y, z) -> (x,
    y, z)`` (curved elements; the warp must respect the periods).  Hexes
    of order k > 1 are all flagged curved in ``spts_curved``.
    """
    ndims = 3
    etype_order = ('hex', 'tet', 'pri', 'pyr')

    def __init__(self, blocks, periodic=(True, True, True), sorder=1,
                 warp=None):
        self.periodic = tuple(bool(p) for p in periodic)
        verts = {et: [] for et in self.etype_order}
        nodes = {et: [] for et in self.etype_order}
        noff = 0
        for (nx, ny, nz), h, layers in blocks:
            if sum(n for _, n in layers) != nz:
                raise ValueError('layers must fill the block')
            k0 = 0
            for split, nl in layers:
                kk, jj, ii = np.meshgrid(np.arange(k0, k0 + nl),
                                         np.arange(ny), np.arange(nx),
                                         indexing='ij')
                cell = np.stack([ii.ravel(), jj.ravel(), kk.ravel()], 1)
                parts = MIXED_SPLITS.get(split,
                                         {SPLIT_ETYPE.get(split): None})
                for et, off in parts.items():
                    if off is None:
                        off = CELL_SPLITS[split]        # (ns, nv, 3)
                    # split-major order: consecutive elements have the same
                    # orientation, so their neighbours' flux points sit in
                    # the same rows and interface gathers stay coalesced
                    lat = off[:, None] + cell[None, :, None, :]
                    lat = lat.reshape(-1, off.shape[1], 3)   # (E, nv, 3)
                    verts[et].append(lat*h)
                    nodes[et].append((lat, np.full(len(lat), noff),
                                      np.array([nx, ny, nz])))
                k0 += nl
            noff += 1
        self.spts, self._nodes = {}, {}
        for et in self.etype_order:
            if verts[et]:
                v = np.concatenate(verts[et])           # (E, nv, 3)
                sp = np.ascontiguousarray(v.transpose(1, 0, 2))
                if sorder > 1 or warp is not None:
                    sp = self._shape_points(et, sp, sorder, warp)
                self.spts[et] = sp
                self._nodes[et] = nodes[et]
        self.etypes = list(self.spts)
        self.neles = sum(v.shape[1] for v in self.spts.values())
        # hexes of shape order k > 1 are all curved (their stored metric);
        # tets and prisms are curved by their shape point count alone
        if 'hex' in self.spts and sorder > 1:
            self.spts_curved = {'hex': np.ones(self.spts['hex'].shape[1],
                                               dtype=bool)}
        self.con = self._connect()

    @staticmethod
    def _shape_points(et, verts, k, warp):
        """(nspts, E, 3) order-k shape points of the affine map of verts
        (nverts, E, 3), warped."""
        if et == 'hex':
            # trilinear map of the 8 vertices at the (k + 1)^3 equispaced
            # points of HexShape.std_ele(k) (x fastest)
            t = np.linspace(0, 1, k + 1)
            xi = np.array([(a, b, c) for c in t for b in t for a in t])
            w = np.ones((len(xi), 8))
            for v in range(8):
                for a in range(3):
                    w[:, v] *= xi[:, a] if (v >> a) & 1 else 1 - xi[:, a]
            x = np.einsum('qv,veb->qbe', w, verts)       # (nspts, 3, E)
        else:
            from pyfr_amd.geometry import AffineGeometry
            from pyfr_amd.simplex import SIMPLEX_SHAPES
            ref = SIMPLEX_SHAPES[et].std_ele(k)
            x = AffineGeometry(verts, et).ploc_at(ref)   # (nspts, 3, E)
        if warp is not None:
            x = np.stack(warp(x[:, 0], x[:, 1], x[:, 2]), axis=1)
        return np.ascontiguousarray(x.transpose(0, 2, 1))

    @classmethod
    def stacked(cls, n, ntet, npri, length=2*np.pi,
                periodic=(True, True, True), sorder=1, warp=None):
        """n x n x (ntet + npri) cells of size length/n: tets below, prisms
        above (tet-prism interfaces at both ends of the prism slab when
        periodic in z)."""
        layers = [(et, k) for et, k in (('tet', ntet), ('pri', npri)) if k]
        return cls([((n, n, ntet + npri), length/n, layers)], periodic,
                   sorder, warp)

    @classmethod
    def hex_prism(cls, n, nhex, npri, length=2*np.pi,
                  periodic=(True, True, True), sorder=1, warp=None):
        """n x n x (nhex + npri) cells of size length/n: hexes below,
        prisms extruded along x above (hex-prism interfaces on quad faces at
        both ends of the prism slab when periodic in z)."""
        layers = [(s, k) for s, k in (('hex', nhex), ('prx', npri)) if k]
        return cls([((n, n, nhex + npri), length/n, layers)], periodic,
                   sorder, warp)

    @classmethod
    def disjoint(cls, ntet, npri, length=2*np.pi):
        """A periodic box of ntet^3 tet cells and a separate periodic box of
        npri^3 prism cells, both of edge ``length``."""
        blocks = []
        if ntet:
            blocks.append(((ntet,)*3, length/ntet, [('tet', ntet)]))
        if npri:
            blocks.append(((npri,)*3, length/npri, [('pri', npri)]))
        return cls(blocks)

    def _connect(self):
        """Pair faces through a structured key: (block, face kind, face
        centroid in lattice units, times 12 so that triangle and quad
        centroids are integers, modulo the block's periods along periodic
        axes).  A face left unpaired must lie on a non-periodic side."""
        per = np.array(self.periodic)
        keys, owners, sides = [], [], []
        for ti, et in enumerate(self.etypes):
            E0 = 0
            for lat, blk, dims in self._nodes[et]:
                E = len(lat)
                for f, fv in enumerate(FACE_VERTS[et]):
                    fl = lat[:, list(fv)]                   # (E, nfv, 3)
                    c = fl.sum(axis=1)*(12//len(fv))
                    k = np.empty((E, 5), dtype=np.int64)
                    k[:, 0] = blk
                    k[:, 1] = len(fv)
                    k[:, 2:] = np.where(per, c % (12*dims), c)
                    keys.append(k)
                    owners.append(np.stack([np.full(E, ti), np.full(E, f),
                                            E0 + np.arange(E)], 1))
                    # side code 2*axis + (high side), or -1
                    sd = np.full(E, -1)
                    for ax in np.flatnonzero(~per):
                        sd[np.all(fl[:, :, ax] == 0, axis=1)] = 2*ax
                        sd[np.all(fl[:, :, ax] == dims[ax], axis=1)] = \
                            2*ax + 1
                    sides.append(sd)
                E0 += E
        keys, owners = np.concatenate(keys), np.concatenate(owners)
        sides = np.concatenate(sides)
        order = np.lexsort(keys.T[::-1])
        sk = keys[order]
        first = np.r_[True, np.any(sk[1:] != sk[:-1], axis=1)]
        starts = np.flatnonzero(first)
        cnt = np.diff(np.r_[starts, len(keys)])
        single = order[starts[cnt == 1]]
        if cnt.max() > 2 or np.any(sides[single] < 0):
            raise ValueError('Faces are not all paired exactly once')
        self.bcon = {}
        bo = owners[single]
        bs = sides[single]
        for code in np.unique(bs):
            sel = bo[bs == code]
            sel = sel[np.lexsort(sel.T[::-1])]
            name = f'{"xyz"[code//2]}{code % 2}'
            self.bcon[name] = [(self.etypes[t], int(e), int(f))
                               for t, f, e in sel]
        pst = starts[cnt == 2]
        a, b = owners[order[pst]], owners[order[pst + 1]]
        # left side = smaller (etype index, face, element) key
        less = ((a[:, 0] < b[:, 0]) |
                ((a[:, 0] == b[:, 0]) & ((a[:, 1] < b[:, 1]) |
                 ((a[:, 1] == b[:, 1]) & (a[:, 2] < b[:, 2])))))
        lft = np.where(less[:, None], a, b)
        rgt = np.where(less[:, None], b, a)
        srt = np.lexsort(lft.T[::-1])
        lft, rgt = lft[srt], rgt[srt]
        ets = self.etypes
        lhs = [(ets[t], int(e), int(f)) for t, f, e in lft]
        rhs = [(ets[t], int(e), int(f)) for t, f, e in rgt]
        return lhs, rhs


def sorted_face_fpts(plocfpts, facefpts, tol=1e-6):
    """srtd_face_fpts for elements of any layout: elements whose face points
    have the same relative layout (to 1e-9) share one fuzzy sort."""
    out = []
    for ffpts in facefpts:
        ffpts = np.asarray(ffpts)
        c = plocfpts[ffpts].transpose(1, 2, 0)           # (E, 3, nfp)
        rel = c - c[:, :, :1]
        q = np.round(rel/1e-9).astype(np.int64).reshape(len(c), -1)
        # group identical layouts through a 64-bit hash of the rounded
        # coordinates, then check every element against its group's
        # representative exactly
        w = np.random.default_rng(0).integers(1, 2**62, q.shape[1])
        with np.errstate(over='ignore'):
            h = (q*w).sum(axis=1)
        _, first, inv = np.unique(h, return_index=True, return_inverse=True)
        inv = inv.ravel()
        if not np.array_equal(q, q[first[inv]]):
            raise RuntimeError('flux-point layout hash collision')
        perms = np.array([_fuzzysort(rel[i], list(range(len(ffpts))), 0, tol)
                          for i in first])
        out.append(ffpts[perms[inv]])
    return out
