"""Writers of small Gmsh ASCII meshes (formats 2.2 and 4.1) for the reader
tests: the generated box meshes written out node by node, with physical
groups for the volume, named boundary sides and periodic side pairs."""
import numpy as np

GMSH_TYPE = {'tet': 4, 'hex': 5, 'pri': 6, 'tri': 2, 'quad': 4 - 1}
# PyFR vertex order -> Gmsh corner order (inverse of the reader's maps)
TO_GMSH = {'hex': [0, 1, 3, 2, 4, 5, 7, 6], 'tet': [0, 1, 2, 3],
           'pri': [0, 1, 2, 3, 4, 5]}
FACE_TYPE = {3: 'tri', 4: 'quad'}


def _node_ids(pts, tol=1e-9):
    """Unique node numbers (1-based) of points (n, 3)."""
    key = np.round(pts/tol).astype(np.int64)
    _, first, inv = np.unique(key, axis=0, return_index=True,
                              return_inverse=True)
    return pts[first], inv.ravel() + 1


def mesh_to_msh(mesh, sides, version='4.1'):
    """Text of a .msh file for ``mesh`` (spts in PyFR vertex order).

    sides: {physical name: list of (etype, eidx, fidx)} surface groups
    (named boundaries and periodic_<k>_l/r pairs), from the mesh's face
    tables.  Boundary faces are written as tri/quad elements."""
    from pyfr_amd.gmsh import FACES
    etypes = list(mesh.spts)
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
        vols.append((t, vids[t][:, TO_GMSH[t]]))
    surfs = []
    for name, faces in sides.items():
        rows = {}
        for t, e, f in faces:
            fv = vids[t][e, list(FACES[t][f])]
            ft = FACE_TYPE[len(fv)]
            if ft == 'quad':          # PyFR quad order -> Gmsh
                fv = fv[[0, 1, 3, 2]]
            rows.setdefault(ft, []).append(fv)
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
            for ft, fvs in rows.items():
                for fv in fvs:
                    els.append(f'{k} {GMSH_TYPE[ft]} 2 {i + 2} {i + 2} ' +
                               ' '.join(map(str, fv)))
                    k += 1
        for t, rows in vols:
            for r in rows:
                els.append(f'{k} {GMSH_TYPE[t]} 2 1 1 ' +
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
        for ft, fvs in rows.items():
            blk = [f'2 {i + 1} {GMSH_TYPE[ft]} {len(fvs)}']
            for fv in fvs:
                blk.append(f'{k} ' + ' '.join(map(str, fv)))
                k += 1
            blocks.append(blk)
    for t, rows in vols:
        blk = [f'3 1 {GMSH_TYPE[t]} {len(rows)}']
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
