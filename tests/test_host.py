"""Host-side logic of the product (CPU): operators, mesh connectivity,
geometry, storage layout, views, configuration, the C-ABI symbol table and
the bench contract.  No GPU calls."""
import ctypes as ct
import os
import re
import subprocess
import sys

import numpy as np
import pytest

from oracle.basis import HexShapeOracle
from oracle.geom import pnorm_at, rcpdjac_at, smats_at, trilinear_ploc
from oracle.mesh import box_hex_verts, pair_faces_bruteforce
from pyfr_amd.geometry import HexGeometry
from pyfr_amd.inifile import Inifile
from pyfr_amd.mesh import BoxMesh, fuzzysort_batched
from pyfr_amd.shapes import HexShape, gauss_legendre

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
OPS = ['M0', 'M1 - M3*M2', 'M3', 'M4 - M6*M0', 'M6']


@pytest.mark.parametrize('order', [1, 2, 3, 4, 5])
def test_operators_match_oracle(order):
    h, o = HexShape(order), HexShapeOracle(order)
    for op in OPS:
        a, b = h.opmat(op), o.opmat(op)
        assert np.array_equal(a != 0, b != 0), op
        assert np.abs(a - b).max() < 3e-14*max(1, np.abs(b).max()), op


def test_operators_match_fixture():
    ref = np.load(os.path.join(ROOT, 'tests', 'golden', 'hex_gleg_ord3.npz'))
    h = HexShape(3)
    for k in ('m0', 'm1', 'm2', 'm3'):
        assert np.abs(getattr(h, k) - ref[k]).max() < 1e-13*max(
            1, np.abs(ref[k]).max())


def test_gauss_legendre_tables():
    g = np.load(os.path.join(ROOT, 'tests', 'golden', 'gl_rules.npz'))
    for n in range(1, 7):
        x, w = gauss_legendre(n)
        assert np.abs(x - g[f'line_n{n}_pts'][:, 0]).max() <= 2.3e-16
        assert np.abs(w - g[f'line_n{n}_wts']).max() <= 1e-15


def test_generated_kernels_match_operators():
    """The values baked into the generated kernels are bit-identical to the
    operators the host builds at run time (plans match bit for bit)."""
    bdir = os.path.join(ROOT, 'pyfr_amd', 'csrc', '_build')
    gens = [os.path.join(bdir, f'gen_ops_p{p}.hip') for p in (1, 2, 3, 4)]
    if not all(os.path.exists(g) for g in gens):
        pytest.skip('library not built')
    src = ''.join(open(g).read() for g in gens)
    for p in (1, 2, 3, 4):
        h = HexShape(p)
        for i, op in enumerate(OPS):
            A = h.opmat(op)
            m = re.search(rf'val_opk_hex{p}_{i}\[\d+\] = \{{([^}}]*)\}}', src)
            vals = np.array([float.fromhex(v) for v in m.group(1).split(',')])
            assert np.array_equal(vals, A[A != 0])


@pytest.mark.parametrize('n', [1, 2, 3, 4])
def test_connectivity_against_bruteforce(n):
    from pyfr_amd.solvers import FluidElements
    p = 2
    m = BoxMesh(n)
    cfg = Inifile.from_dict({'solver': {'order': p},
                             'constants': {'gamma': 1.4}})
    e = FluidElements(HexShape, m.spts['hex'], cfg)
    sf = e.srtd_face_fpts
    lhs, rhs = m.con
    assert len(lhs) == 3*n**3
    ours = set()
    for (_, el, fl), (_, er, fr) in zip(lhs, rhs):
        for a, b in zip(sf[fl][el], sf[fr][er]):
            ours.add(tuple(sorted(((el, a), (er, b)))))
    o = HexShapeOracle(p)
    pl = trilinear_ploc(box_hex_verts(n), o.fpts)
    ref = set()
    for eL, fL, iL, eR, fR, iR in pair_faces_bruteforce(pl, 6, o.nfacefpts,
                                                        2*np.pi):
        for a, b in zip(iL, iR):
            ref.add(tuple(sorted(((eL, a), (eR, b)))))
    assert ours == ref
    # left side = smaller (face, element) key, face-major order
    keys = [(f, e) for (_, e, f) in lhs]
    assert keys == sorted(keys)
    assert all((fl, el) < (fr, er)
               for (_, el, fl), (_, er, fr) in zip(lhs, rhs))


def test_fuzzysort():
    c = np.array([[[0.0, 1.0, 0.0, 1.0], [1.0, 0.0, 0.0, 1.0],
                   [5.0, 5.0, 5.0 + 1e-9, 5.0]]])
    assert list(fuzzysort_batched(c)[0]) == [2, 0, 1, 3]


def test_geometry_against_oracle_warped():
    n, p = 3, 2
    m = BoxMesh(n)
    rng = np.random.default_rng(1)
    v = m.spts['hex'] + 0.05*rng.standard_normal(m.spts['hex'].shape)
    sh = HexShape(p)
    g = HexGeometry(v, p)
    ov = v.transpose(0, 2, 1)
    assert np.abs(g.smat_at(sh.upts) - smats_at(ov, p, sh.upts)).max() < 1e-13
    assert np.abs(g.rcpdjac_at(sh.upts) -
                  rcpdjac_at(ov, p, sh.upts)).max() < 1e-13
    assert np.abs(g.pnorm_at(sh.fpts, sh.norm_fpts) -
                  pnorm_at(ov, p, sh.fpts, sh.norm_fpts)).max() < 1e-13


class _FakeBackend:
    soasz = csubsz = 64
    blocks = False

    def __init__(self):
        self.mats = {}
        self._n = 0

    def malloc(self, obj, extent):
        obj.basedata, obj.offset = 0, self._n
        obj._storage_root = self
        obj.data = self._n
        self._n += obj.nbytes

    def const_matrix(self, m, dtype=None, tags=()):
        class C:
            pass
        c = C()
        c.data = None
        c.host = m
        return c


def test_aosoa_layout_roundtrip():
    from pyfr_amd.backend import Matrix
    be = _FakeBackend()
    a = np.random.default_rng(0).standard_normal((7, 5, 130))
    m = Matrix(be, np.float64, a.shape, None, None, {'align'})
    assert m.leaddim == 5*192 and m.nrow == 7
    buf = m._pack(a)
    assert np.array_equal(m._unpack(buf.ravel()), a)
    # element e / variable v of row r at r*ld + ((e/64)*nvar + v)*64 + e%64
    flat = buf.ravel()
    for r, v, e in [(0, 0, 0), (3, 2, 65), (6, 4, 129), (2, 1, 63)]:
        assert flat[r*m.leaddim + ((e//64)*5 + v)*64 + e % 64] == a[r, v, e]
    a4 = np.random.default_rng(1).standard_normal((3, 4, 5, 70))
    m4 = Matrix(be, np.float64, a4.shape, None, None, {'align'})
    assert m4.nrow == 12
    assert np.array_equal(m4._unpack(m4._pack(a4).ravel()), a4)


def test_aosoa_row_padding():
    """[backend-hip] row-pad: rows of AoSoA arrays padded by a multiple of
    64 elements; the layout formula holds with the padded leading
    dimension and the pad never reaches the unpacked array."""
    from pyfr_amd.backend import Matrix
    be = _FakeBackend()
    be.row_pad = 128
    a = np.random.default_rng(2).standard_normal((7, 5, 130))
    m = Matrix(be, np.float64, a.shape, None, None, {'align'})
    assert m.ncol == 5*192 and m.leaddim == 5*192 + 128
    assert m.nbytes == 7*m.leaddim*8
    buf = m._pack(a)
    assert buf.shape == (7, m.leaddim) and not buf[:, m.ncol:].any()
    flat = buf.ravel()
    for r, v, e in [(0, 0, 0), (3, 2, 65), (6, 4, 129), (2, 1, 63)]:
        assert flat[r*m.leaddim + ((e//64)*5 + v)*64 + e % 64] == a[r, v, e]
    assert np.array_equal(m._unpack(flat), a)
    a4 = np.random.default_rng(3).standard_normal((3, 4, 5, 70))
    m4 = Matrix(be, np.float64, a4.shape, None, None, {'align'})
    assert np.array_equal(m4._unpack(m4._pack(a4).ravel()), a4)


def test_view_mapping_formula():
    from pyfr_amd.backend import Matrix, View
    be = _FakeBackend()
    m = Matrix(be, np.float64, (6, 5, 100), None, None, {'align'})
    m.mid = 0
    be.mats[0] = m
    rmap = np.array([0, 3, 5, 2])
    cmap = np.array([0, 64, 99, 17])
    v = View(be, np.zeros(4, int), rmap, cmap, 1, (5,), set())
    exp = m.offset//8 + rmap*m.leaddim + (cmap//64)*64*5 + cmap % 64
    assert np.array_equal(v.mapping_np, exp)


def test_inifile_expressions():
    c = Inifile.from_dict({'a': {'p': '100/1.4 + cos(2*x)', 'q': '1e-3*x2 + 10'}})
    assert c.getexpr('a', 'p') == '100.0/1.4 + cos(2.0*x)'
    assert c.getexpr('a', 'q') == '1e-3*x2 + 10.0'
    assert c.getint('a', 'missing', 3) == 3


def test_library_exports_header_symbols():
    """libfrhip.so loads without a GPU and exports every function that
    include/frhip.h declares (no compute calls)."""
    lib = os.path.join(ROOT, 'pyfr_amd', 'libfrhip.so')
    if not os.path.exists(lib):
        pytest.skip('library not built')
    hdr = open(os.path.join(ROOT, 'include', 'frhip.h')).read()
    names = set(re.findall(r'\b(frhip_\w+)\s*\(', hdr))
    assert len(names) > 30
    so = ct.CDLL(lib)
    for nm in sorted(names):
        assert hasattr(so, nm), nm
    from pyfr_amd._lib import _FUNCS
    assert set(_FUNCS) | {'frhip_last_error'} == names


def test_bench_refuses_multi_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, 'bench.py'),
                        '--gpus', '2'], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 2 and 'one GPU' in r.stderr


def test_planted_oracle_mistake_is_caught():
    """A sign error planted in the oracle's correction operator is caught by
    the density-wave convergence check used to pin the RHS."""
    from oracle.rhs import OracleHexSystem
    U = np.array([0.5, 0.3, -0.2])
    errs = []
    for n in (4, 8):
        o = OracleHexSystem(n, 2)
        o.M3 = -o.M3
        x, y, z = trilinear_ploc(o.verts, o.shape.upts).transpose(1, 0, 2)
        rho = 1 + 0.2*np.sin(x + y + z)
        u = np.stack([rho, rho*U[0], rho*U[1], rho*U[2],
                      1/0.4 + 0.5*rho*(U**2).sum()], axis=1)
        exact = -U.sum()*0.2*np.cos(x + y + z)
        errs.append(np.sqrt(np.mean((o.rhs(u)[:, 0] - exact)**2)))
    assert np.log2(errs[0]/errs[1]) < 2 - 0.7
