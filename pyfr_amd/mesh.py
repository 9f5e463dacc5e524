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
    """n^3 (or nx*ny*nz) periodic box of linear hexes on [0, L)^3."""
    ndims = 3

    def __init__(self, n, length=2*np.pi, origin=0.0):
        nx, ny, nz = (n, n, n) if np.isscalar(n) else n
        self.shape = (nx, ny, nz)
        self.length = length
        hx, hy, hz = length/nx, length/ny, length/nz
        E = nx*ny*nz
        e = np.arange(E)
        ii, jj, kk = e % nx, (e // nx) % ny, e // (nx*ny)
        verts = np.empty((8, E, 3))
        for v in range(8):
            bx, by, bz = v & 1, (v >> 1) & 1, (v >> 2) & 1
            verts[v, :, 0] = origin + (ii + bx)*hx
            verts[v, :, 1] = origin + (jj + by)*hy
            verts[v, :, 2] = origin + (kk + bz)*hz
        self.spts = {'hex': verts}
        self.etypes = ['hex']
        self.neles = E
        self.con = self._connect(np.stack([ii, jj, kk]))

    def _connect(self, ijk):
        """Pair faces through a structured key: (normal axis, periodic plane
        index, the element's indices along the two other axes)."""
        E = ijk.shape[1]
        nf = len(HEX_FACE_AXES)
        dims = np.array(self.shape)
        keys = np.empty((nf, E, 4), dtype=np.int64)
        for f, (nax, side) in enumerate(HEX_FACE_AXES):
            oth = [a for a in range(3) if a != nax]
            keys[f, :, 0] = nax
            keys[f, :, 1] = (ijk[nax] + (side > 0)) % dims[nax]
            keys[f, :, 2] = ijk[oth[0]]
            keys[f, :, 3] = ijk[oth[1]]
        keys = keys.reshape(nf*E, 4)
        fidx = np.repeat(np.arange(nf), E)
        eidx = np.tile(np.arange(E), nf)
        order = np.lexsort(keys.T[::-1])
        sk = keys[order]
        same = np.all(sk[1:] == sk[:-1], axis=1)
        starts = np.flatnonzero(same)
        if 2*len(starts) != nf*E or np.any(same[1:] & same[:-1]):
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
