"""Writers of small Gmsh ASCII meshes (formats 2.2 and 4.1) for the reader
tests: the generated box meshes written out node by node (first order, or
the shape points of high-order elements in Gmsh's node order), with
physical groups for the volume, named boundary sides and periodic side
pairs."""
import numpy as np

FACE_TYPE = {3: 'tri', 4: 'quad'}


def _gtype(kind, n):
    from pyfr_amd.gmsh import GMSH_TYPES
    return next(g for g, v in GMSH_TYPES.items() if v == (kind, n))


def _to_gmsh(kind, k):
    """PyFR shape point of each Gmsh node (inverse of the reader's map)."""
    from pyfr_amd.gmsh import gmsh_nodemap
    return np.argsort(gmsh_nodemap(kind, k))


def _face_nodes(kind, k, fv):
    """Shape points (std_ele(k) indices) of the order-k face through the
    PyFR vertices fv, in Gmsh's node order for that face."""
    from pyfr_amd.gmsh import _gmsh_points, _vertex_index, std_ele
    pts = std_ele(kind, k)
    corners = [pts[_vertex_index(kind, k)[v]] for v in fv]
    if len(fv) == 4:                      # PyFR quad order -> Gmsh
        corners = [corners[i] for i in (0, 1, 3, 2)]
    g = np.array(_gmsh_points(FACE_TYPE[len(fv)], k, corners))
    return [int(np.abs(pts - x).max(axis=1).argmin()) for x in g]


def _node_ids(pts, tol=1e-9):
    """Unique node numbers (1-based) of points (n, 3)."""
    key = np.round(pts/tol).astype(np.int64)
    _, first, inv = np.unique(key, axis=0, return_index=True,
                              return_inverse=True)
    return pts[first], inv.ravel() + 1


def mesh_to_msh(mesh, sides, version='4.1', hofaces=False):
    """Text of a .msh file for ``mesh`` (spts in PyFR's std_ele(k) order,
    k from the point count).

    sides: {physical name: list of (etype, eidx, fidx)} surface groups
    (named boundaries and periodic_<k>_l/r pairs), from the mesh's face
    tables.  Boundary faces are written as first-order tri/quad elements,
    or with hofaces as faces of their element's shape order."""
    from pyfr_amd.gmsh import FACES, _vertex_index, npts_order
    etypes = list(mesh.spts)
    order = {t: npts_order(t, mesh.spts[t].shape[0]) for t in etypes}
    allpts = np.concatenate([mesh.spts[t].swapaxes(0, 1).reshape(-1, 3)
                             for t in etypes])
    nodes, ids = _node_ids(allpts)
    vids, off = {}, 0
    for t in etypes:
        nv, ne, _ = mesh.spts[t].shape
        vids[t] = ids[off:off + nv*ne].reshape(ne, nv)
        off += nv*ne
    vols = []
    for t in etypes:
        vols.append((_gtype(t, mesh.spts[t].shape[0]),
                     vids[t][:, _to_gmsh(t, order[t])]))
    surfs = []
    for name, faces in sides.items():
        rows = {}
        for t, e, f in faces:
            k, fv = order[t], FACES[t][f]
            if hofaces:
                fn = vids[t][e, _face_nodes(t, k, fv)]
            else:
                fn = vids[t][e, [_vertex_index(t, k)[v] for v in fv]]
                if len(fv) == 4:          # PyFR quad order -> Gmsh
                    fn = fn[[0, 1, 3, 2]]
            rows.setdefault(_gtype(FACE_TYPE[len(fv)], len(fn)),
                            []).append(fn)
        surfs.append((name, rows))

    L = ['$MeshFormat', f'{version} 0 8', '$EndMeshFormat']
    L += ['$PhysicalNames', str(1 + len(surfs)), '3 1 "fluid"']
    L += [f'2 {i + 2} "{name}"' for i, (name, _) in enumerate(surfs)]
    L += ['$EndPhysicalNames']
    if version == '2.2':
        L += ['$Nodes', str(len(nodes))]
        L += [f'{i + 1} {float(x)!r} {float(y)!r} {float(z)!r}' for i, (x, y, z) in
              enumerate(nodes)]
        L += ['$EndNodes']
        els, k = [], 1
        for i, (name, rows) in enumerate(surfs):
            for gt, fvs in rows.items():
                for fv in fvs:
                    els.append(f'{k} {gt} 2 {i + 2} {i + 2} ' +
                               ' '.join(map(str, fv)))
                    k += 1
        for gt, rows in vols:
            for r in rows:
                els.append(f'{k} {gt} 2 1 1 ' +
                           ' '.join(map(str, r)))
                k += 1
        L += ['$Elements', str(len(els)), *els, '$EndElements']
        return '\n'.join(L) + '\n'

    # 4.1: one surface entity per physical surface, one volume entity
    L += ['$Entities', f'0 0 {len(surfs)} 1']
    for i in range(len(surfs)):
        L.append(f'{i + 1} 0 0 0 1 1 1 1 {i + 2} 0')
    L += ['1 0 0 0 1 1 1 1 1 0', '$EndEntities']
    n = len(nodes)
    L += ['$Nodes', f'1 {n} 1 {n}', f'3 1 0 {n}']
    L += [str(i + 1) for i in range(n)]
    L += [f'{float(x)!r} {float(y)!r} {float(z)!r}' for x, y, z in nodes]
    L += ['$EndNodes']
    blocks, k = [], 1
    for i, (name, rows) in enumerate(surfs):
        for gt, fvs in rows.items():
            blk = [f'2 {i + 1} {gt} {len(fvs)}']
            for fv in fvs:
                blk.append(f'{k} ' + ' '.join(map(str, fv)))
                k += 1
            blocks.append(blk)
    for gt, rows in vols:
        blk = [f'3 1 {gt} {len(rows)}']
        for r in rows:
            blk.append(f'{k} ' + ' '.join(map(str, r)))
            k += 1
        blocks.append(blk)
    nel = k - 1
    L += ['$Elements', f'{len(blocks)} {nel} 1 {nel}']
    for blk in blocks:
        L += blk
    L += ['$EndElements']
    return '\n'.join(L) + '\n'


def box_sides(mesh, spec):
    """Physical surface groups of a non-periodic BoxMesh: spec maps each
    of its sides ('x0', ...) to a physical name; two sides given as
    ('periodic_k_l', 'periodic_k_r') become a periodic pair."""
    out = {}
    for side, name in spec.items():
        out.setdefault(name, []).extend(mesh.bcon[side])
    return out
