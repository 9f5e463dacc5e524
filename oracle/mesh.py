"""Oracle: periodic hex box meshes and brute-force interface pairing
(test infrastructure).

The mesh itself (vertex coordinates per element) is input data.  The pairing
of faces and of flux points across each interface is computed here by
geometric brute force -- matching physical coordinates modulo the box period
-- and so is independent of the product's sort-based pairing.  It checks the
same facts the reference's ``_construct_con`` (``pyfr/readers/native.py:445-495``)
and ``srtd_face_fpts`` (``pyfr/solvers/base/elements.py:161-172``) establish:
every face is used exactly once and paired flux points coincide.
"""
import numpy as np

from oracle.geom import hex_linspts


def box_hex_verts(n, length=2*np.pi, origin=0.0, warp=None):
    """Vertices (8, 3, E) of an n^3 box of hexes, element index
    e = i + n*(j + n*k) (x fastest); vertex order of ``HexShape``.

    ``warp``: optional callable mapping (x, y, z) arrays to displaced
    coordinates (must be periodic with the box) for non-affine tests.
    """
    h = length/n
    lin = hex_linspts()
    ii, jj, kk = np.meshgrid(np.arange(n), np.arange(n), np.arange(n),
                             indexing='ij')
    # element order: x fastest
    ii = ii.transpose(2, 1, 0).ravel()
    jj = jj.transpose(2, 1, 0).ravel()
    kk = kk.transpose(2, 1, 0).ravel()
    verts = np.empty((8, 3, n**3))
    for v in range(8):
        bx, by, bz = (lin[v] + 1)/2
        verts[v, 0] = origin + (ii + bx)*h
        verts[v, 1] = origin + (jj + by)*h
        verts[v, 2] = origin + (kk + bz)*h
    if warp is not None:
        x, y, z = warp(verts[:, 0], verts[:, 1], verts[:, 2])
        verts = np.stack([x, y, z], axis=1)
    return verts


def _wrap(x, length):
    y = np.mod(x, length)
    y[np.isclose(y, length, atol=1e-9, rtol=0)] = 0.0
    return y


def pair_faces_bruteforce(fpts_phys, nfaces, nfacefpts, length,
                          periodic=(True, True, True), boundaries=None):
    """Pair element faces and their flux points.

    fpts_phys: (nfpts, 3, E) physical flux-point coordinates (faces' points
    contiguous, ``nfacefpts`` per face).  Returns a list of
    (eL, fL, idxL, eR, fR, idxR) where idxL/idxR are flux-point indices
    (into the element's fpts) of matched points, in matched order.

    Coordinates wrap only along the ``periodic`` axes.  Unpaired faces are
    an error unless ``boundaries`` (a dict) is given: it then receives, per
    box side name 'x0', 'x1', 'y0', ... (axis, coordinate 0 or ``length``),
    a list of (e, f, idx) of the faces lying on that side.
    """
    nfpts, _, neles = fpts_phys.shape
    off = np.cumsum([0] + list(nfacefpts))
    per = np.array(periodic, dtype=bool)

    def wrap(x):
        y = x.copy()
        y[..., per] = _wrap(x[..., per], length)
        return y

    keys = {}
    for e in range(neles):
        for f in range(nfaces):
            pts = fpts_phys[off[f]:off[f + 1], :, e]
            c = wrap(pts.mean(axis=0))
            key = tuple(np.round(c, 6))
            keys.setdefault(key, []).append((e, f))

    pairs = []
    for key, faces in keys.items():
        if len(faces) == 1 and boundaries is not None:
            (e, f), = faces
            side = [f'{"xyz"[a]}{int(np.isclose(key[a], length))}'
                    for a in range(3) if not per[a] and
                    (np.isclose(key[a], 0) or np.isclose(key[a], length))]
            if len(side) != 1:
                raise ValueError(f'Unpaired interior face at {key}')
            boundaries.setdefault(side[0], []).append(
                (e, f, off[f] + np.arange(off[f + 1] - off[f])))
            continue
        if len(faces) != 2:
            raise ValueError(f'Unpaired face at {key}: {faces}')
        # left = smaller (face, element) key, as the reference's
        # _construct_con keeps the pair with lkey < rkey, lkey = (face codec
        # index, element index) (pyfr/readers/native.py:404-406, :466-470);
        # HLL is not antisymmetric under a side swap, so this matters
        (eL, fL), (eR, fR) = sorted(faces, key=lambda q: (q[1], q[0]))
        pl = wrap(fpts_phys[off[fL]:off[fL + 1], :, eL])
        pr = wrap(fpts_phys[off[fR]:off[fR + 1], :, eR])
        d = np.abs(pl[:, None, :] - pr[None, :, :])
        d[..., per] = np.minimum(d[..., per], length - d[..., per])
        d = d.sum(axis=2)
        jr = np.argmin(d, axis=1)
        if np.any(d[np.arange(len(pl)), jr] > 1e-8) or \
           len(set(jr.tolist())) != len(pl):
            raise ValueError('Flux points do not coincide')
        pairs.append((eL, fL, off[fL] + np.arange(len(pl)),
                      eR, fR, off[fR] + jr))
    return pairs
