"""The reference's native mesh file (``.pyfrm``, format version 2) and the
load path that turns it into the mesh the solver consumes.

A raw mesh is what the reference's importer writes
(``pyfr/readers/base.py:67-105, 363-410``): the node table (location,
valency), one structured array per element type in file order (node
numbers in PyFR shape-point order, a curved flag, per face the codec index
and element number of what lies across it -- ``off`` = -1 for a boundary,
whose codec entry is ``bc/<name>`` -- a colour and a tag bitmask), the
codec, periodic face pairs and one default partitioning with the curved
elements of each type first.

Loading follows ``pyfr/readers/native.py:89-495`` for one partition:
element types in name order, each type's elements in the partitioning's
order, shape points gathered from the node table, and interfaces from the
per-face links, the left side of each the face with the smaller (codec
index, local element number) key, in the order faces are visited (type,
face, element).  The result carries what the solver takes from a mesh:
``spts``, ``spts_curved`` (hexes), ``con``, ``bcon``.  As for meshes read
from Gmsh, a tet or prism type with no curved element keeps its corners
only (the solver then uses its affine metric).

Solution files (``.pyfrs``, version 2; ``pyfr/writers/native.py``,
``pyfr/plugins/writer.py:80-118``): ``write_pyfrs`` stores the state of an
integrator as the reference's writer plugin does -- the configuration and
stats texts, the mesh UUID, and per element type a dataset
``soln/p<order>-<etype>`` of records {soln: one (nupts,) array per
conserved variable; aux: part-id}, in file element order, with the
solution points as its ``pts`` attribute; ``read_pyfrs`` returns the
arrays in the mesh's (local) order for a restart.

Not restated: the mesh UUID is a SHA-256 of the raw arrays rather than of
the reference's pickled tuple, and shared-node tables and multi-partition
files are not built (one GPU).
"""
import hashlib
import re
import uuid as _uuid

import numpy as np

from pyfr_amd.h5 import File

CREATOR = 'pyfr-amd'
NVERTS = {'tet': 4, 'hex': 8, 'pri': 6, 'pyr': 5}


def ele_dtype(nnodes, nfaces):
    """The element record of the format (``pyfr/readers/base.py:386-390``)."""
    return np.dtype([('nodes', np.int64, nnodes), ('curved', bool),
                     ('faces', [('cidx', np.int16), ('off', np.int64)],
                      nfaces),
                     ('colour', np.uint8), ('tags', np.uint64)])


def element_colouring(eles, codec):
    """Greedy element colouring (``pyfr/readers/base.py:308-358``): in
    global order (types in dict order), each element takes the smallest
    colour none of its earlier neighbours has, or a less-used free one."""
    maxc = max(e['faces'].shape[-1] for e in eles.values()) + 1
    edisp, disp = {}, 0
    for t, e in eles.items():
        edisp[t] = disp
        disp += len(e)
    cdisp = np.full(len(codec), -1, dtype=np.int64)
    for t, e in eles.items():
        for i in range(e['faces'].shape[-1]):
            cdisp[codec.index(f'eles/{t}/face/{i}')] = edisp[t]
    colours = np.zeros(disp, dtype=np.int64)
    counts = [0]*maxc
    for t, e in eles.items():
        cid, off = e['faces']['cidx'], e['faces']['off']
        ngl = np.where(off >= 0, cdisp[cid] + off, -1)
        for i in range(len(e)):
            g = edisp[t] + i
            avail = [True]*maxc
            for n in ngl[i]:
                if 0 <= n < g:
                    avail[colours[n]] = False
            c = avail.index(True)
            mn = counts[c]
            for cc, (a, n) in enumerate(zip(avail, counts)):
                if a and n and n < mn:
                    c, mn = cc, n
            colours[g] = c
            counts[c] += 1
        e['colour'] = colours[edisp[t]:edisp[t] + len(e)]


def default_partitioning(eles):
    """Partition 1 of the format (``pyfr/readers/base.py:20-45``): per type
    in name order, the curved elements then the others; regions are the
    running offsets."""
    parts, regions = [], [0]
    for t in sorted(eles):
        c = eles[t]['curved']
        parts += [np.flatnonzero(c), np.flatnonzero(~c)]
        regions.append(regions[-1] + len(c))
    return (np.concatenate(parts).astype(np.int64),
            np.array([regions], dtype=np.int64))


class RawMesh:
    """The mesh as the importer writes it; ``write`` / ``read`` the .pyfrm
    file, ``load`` the solver's view."""

    def __init__(self, nodes, eles, codec, pmap=None, creator=CREATOR,
                 uuid=None):
        self.nodes = np.asarray(nodes, dtype=np.float64)
        self.eles = eles
        self.codec = list(codec)
        self.pmap = pmap or {}
        self.creator = creator
        self.uuid = uuid or self._digest()

    def _digest(self):
        h = hashlib.sha256()
        h.update(np.ascontiguousarray(self.nodes).tobytes())
        for t in sorted(self.eles):
            h.update(t.encode())
            h.update(np.ascontiguousarray(self.eles[t]).tobytes())
        h.update('\n'.join(self.codec).encode())
        return str(_uuid.UUID(h.hexdigest()[:32]))

    def valency(self):
        v = np.zeros(len(self.nodes), dtype=np.uint16)
        for e in self.eles.values():
            k, c = np.unique(e['nodes'], return_counts=True)
            v[k] += c.astype(np.uint16)
        return v

    # -- file ----------------------------------------------------------------
    def write(self, fname):
        """``pyfr/readers/base.py:67-105``."""
        from pyfr_amd.gmsh import npts_order, std_ele
        with File(fname, 'w') as f:
            parts, regions = default_partitioning(self.eles)
            f['codec'] = np.array(self.codec, dtype='S')
            f['creator'] = np.array(self.creator, dtype='S')
            f['mesh-uuid'] = np.array(self.uuid, dtype='S')
            f['version'] = np.int64(2)
            nd = np.zeros(len(self.nodes), dtype=[('location', float, 3),
                                                  ('valency', np.uint16)])
            nd['location'] = self.nodes
            nd['valency'] = self.valency()
            f['nodes'] = nd
            for t, e in self.eles.items():
                f[f'eles/{t}'] = e
                k = npts_order(t, e['nodes'].shape[1])
                f.set_attr(f'eles/{t}', 'pts', std_ele(t, k))
            for name, pidx in self.pmap.items():
                f[f'periodic/{name}'] = pidx
            f['partitionings/1/eles'] = parts
            f.set_attr('partitionings/1/eles', 'regions', regions)

    @classmethod
    def read(cls, fname):
        with File(fname) as f:
            version = int(f['version'][()])
            if version != 2:
                raise ValueError(f'Mesh format version {version} is not 2')
            codec = [c.decode() for c in f['codec']]
            nodes = f['nodes']['location']
            eles = {t: f[f'eles/{t}'] for t in f.keys('eles')}
            pmap = ({k: f[f'periodic/{k}'] for k in f.keys('periodic')}
                    if 'periodic' in f else {})
            raw = cls(nodes, eles, codec, pmap,
                      creator=f['creator'][()].decode(),
                      uuid=f['mesh-uuid'][()].decode())
            parts = f['partitionings/1/eles']
            regions = f.attr('partitionings/1/eles', 'regions')
        if regions.shape[0] != 1:
            raise NotImplementedError('Only single-partition meshes load '
                                      'on one GPU (partitioning 1)')
        raw.order = {t: parts[regions[0, i]:regions[0, i + 1]]
                     for i, t in enumerate(sorted(eles))}
        return raw

    # -- load ----------------------------------------------------------------
    def load(self, mesh):
        """Fill mesh's solver-facing attributes (``pyfr/readers/native.py:
        339-463``, one partition)."""
        order = getattr(self, 'order', None)
        if order is None:
            parts, regions = default_partitioning(self.eles)
            order = {t: parts[regions[0, i]:regions[0, i + 1]]
                     for i, t in enumerate(sorted(self.eles))}
        etypes = sorted(self.eles)
        for t in etypes:
            if t not in NVERTS:
                raise NotImplementedError(f'{t} elements are not supported')
        mesh.etypes = etypes
        mesh.spts, mesh.spts_curved = {}, {}
        local = {}
        for t in etypes:
            e = self.eles[t][order[t]]
            local[t] = e
            sp = self.nodes[e['nodes']].swapaxes(0, 1)   # (nspts, neles, 3)
            curved = e['curved'].astype(bool)
            if not curved.any():
                from pyfr_amd.gmsh import _vertex_index, npts_order
                sp = sp[_vertex_index(t, npts_order(t, sp.shape[0]))]
            elif t == 'hex':
                mesh.spts_curved['hex'] = curved
            mesh.spts[t] = np.ascontiguousarray(sp)
        if not mesh.spts_curved:
            del mesh.spts_curved
        mesh.neles = sum(len(order[t]) for t in etypes)

        # codec index -> (etype, face) and the face entries per type
        cidxmap = {}
        for c, name in enumerate(self.codec):
            m = re.match(r'eles/(\w+)/face/(\d+)$', name)
            if m:
                cidxmap[c] = (m[1], int(m[2]))
        g2l = {}
        for t in etypes:
            inv = np.full(len(self.eles[t]), -1, dtype=np.int64)
            inv[order[t]] = np.arange(len(order[t]))
            g2l[t] = inv
        lc, le, rc, rg = [], [], [], []
        for t in etypes:
            fc = local[t]['faces']
            for f in range(fc.shape[1]):
                n = len(fc)
                lc.append(np.full(n, self.codec.index(f'eles/{t}/face/{f}')))
                le.append(np.arange(n))
                rc.append(fc['cidx'][:, f].astype(np.int64))
                rg.append(fc['off'][:, f])
        lc, le, rc, rg = map(np.concatenate, (lc, le, rc, rg))
        re_ = np.full(len(lc), -1, dtype=np.int64)
        for c, (t, _) in cidxmap.items():
            m = (rc == c) & (rg >= 0)
            re_[m] = g2l[t][rg[m]]
        bnd = rg < 0
        if np.any(~bnd & (re_ < 0)):
            raise ValueError('Face links an element outside the mesh')
        stride = max(int(le.max(initial=-1)), int(re_.max(initial=-1))) + 1
        lkey, rkey = lc*stride + le, rc*stride + re_
        ii = np.flatnonzero(~bnd & (lkey < rkey))
        face = (lambda c, e: (cidxmap[c][0], int(e), cidxmap[c][1]))
        mesh.con = ([face(c, e) for c, e in zip(lc[ii], le[ii])],
                    [face(c, e) for c, e in zip(rc[ii], re_[ii])])
        mesh.bcon = {}
        for bc in np.unique(rc[bnd]):
            name = self.codec[bc]
            if not name.startswith('bc/'):
                raise ValueError(f'Boundary face linked to {name}')
            bi = np.flatnonzero(bnd & (rc == bc))
            mesh.bcon[name[3:]] = [face(c, e) for c, e in zip(lc[bi], le[bi])]
        return mesh


class NativeMesh:
    """A mesh read from a .pyfrm file (format version 2): ``spts``,
    ``spts_curved`` (hexes), ``con``, ``bcon`` as the solver takes them."""
    ndims = 3

    def __init__(self, fname):
        self.raw = RawMesh.read(fname)
        self.raw.load(self)
        self.codec, self.uuid = self.raw.codec, self.raw.uuid


def _local_order(mesh):
    """Per element type, the file index of each local element."""
    raw = mesh.raw
    order = getattr(raw, 'order', None)
    if order is None:
        parts, regions = default_partitioning(raw.eles)
        order = {t: parts[regions[0, i]:regions[0, i + 1]]
                 for i, t in enumerate(sorted(raw.eles))}
    return order


def write_pyfrs(fname, mesh, integ, prefix='soln', fpdtype=None):
    """Write integ's current solution on mesh (a NativeMesh or GmshMesh)
    as a .pyfrs file (format version 2)."""
    from pyfr_amd.inifile import Inifile
    from pyfr_amd.gmsh import npts_order
    sysm = integ.system
    order = _local_order(mesh)
    p = integ.cfg.getint('solver', 'order')
    stats = Inifile()
    stats.set('data', 'prefix', prefix)
    stats.set('solver-time-integrator', 'tcurr', repr(float(integ.tcurr)))
    stats.set('solver-time-integrator', 'nsteps', str(integ.nsteps))
    stats.set('solver-time-integrator', 'nrhsevals', str(integ.nrhsevals))
    cfgtxt = integ.cfg.tostr()
    with File(fname, 'w') as f:
        f['config'] = np.array(cfgtxt, dtype='S')
        f['config-0'] = np.array(cfgtxt, dtype='S')
        f['stats'] = np.array(stats.tostr(), dtype='S')
        f['mesh-uuid'] = np.array(mesh.raw.uuid, dtype='S')
        f['creator'] = np.array(CREATOR, dtype='S')
        f['version'] = np.int64(2)
        for t, u in zip(sysm.ele_types, sysm.soln_upts(integ.idxcurr)):
            e = sysm.ele_map[t]
            nupts, nvars, neles = u.shape
            dt = np.dtype(fpdtype or u.dtype)
            rec = np.zeros(neles, dtype=[
                ('soln', [(v, dt, (nupts,)) for v in e.convars]),
                ('aux', [('part-id', np.int64)])])
            fidx = order[t]
            for j, v in enumerate(e.convars):
                rec['soln'][v][fidx] = u[:, j, :].T
            f[f'{prefix}/p{p}-{t}'] = rec
            f.set_attr(f'{prefix}/p{p}-{t}', 'pts',
                       np.asarray(e.basis.upts, dtype=np.float64))


def read_pyfrs(fname, mesh, prefix=None):
    """(solution arrays {etype: (nupts, nvars, neles)} in the mesh's local
    order, config text, stats text) of a .pyfrs file written for mesh."""
    order = _local_order(mesh)
    with File(fname) as f:
        if int(f['version'][()]) != 2:
            raise ValueError('Solution file must be format v2')
        if f['mesh-uuid'][()].decode() != mesh.raw.uuid:
            raise ValueError('Invalid solution for mesh')
        cfg = f['config'][()].decode()
        stats = f['stats'][()].decode()
        if prefix is None:
            from pyfr_amd.inifile import Inifile
            prefix = Inifile(stats).get('data', 'prefix')
        out = {}
        for name in f.keys(prefix):
            m = re.match(r'p(\d+)-(\w+)$', name)
            if not m:
                continue
            t = m[2]
            rec = f[f'{prefix}/{name}']
            names = rec.dtype['soln'].names
            a = np.stack([rec['soln'][v] for v in names], axis=1)
            out[t] = np.ascontiguousarray(a[order[t]].transpose(2, 1, 0))
    return out, cfg, stats
