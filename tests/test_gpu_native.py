"""GPU runs on native files: a mesh imported from Gmsh, written as .pyfrm
and read back runs the boundary-condition RHS against the oracle (as the
Gmsh-read mesh does, tests/test_gpu_bcs.py); a solution written as .pyfrs
and read back restarts a run bit for bit (tet/prism, pyramid/tet and hex
meshes).
Tolerances: RHS max(1e-12, 10 x the oracle's input-noise floor); restart
exact on hexes, 1e-13 on tets and prisms (a second backend's tuner may pick
another dense-kernel variant)."""
import io

import numpy as np
import pytest

from tests.fr_cases import make_simplex_cfg, mode_ics, rel_l2

pytestmark = pytest.mark.gpu


def _gmsh_box(n, spec, sorder=1, warp=None):
    from pyfr_amd.gmsh import GmshMesh
    from pyfr_amd.mesh import BoxMesh
    from tests.msh_files import box_sides, mesh_to_msh
    raw = BoxMesh(n, periodic=(False, False, False), sorder=sorder,
                  warp=warp)
    return GmshMesh(io.StringIO(mesh_to_msh(raw, box_sides(raw, spec))))


def test_bc_rhs_parity_pyfrm(backend_factory, tmp_path):
    from pyfr_amd.native import NativeMesh
    from tests.test_gpu_bcs import _floor, _setup
    n = 3
    spec = {'x0': 'x0', 'x1': 'x1', 'y0': 'y0', 'y1': 'y1',
            'z0': 'periodic_z_l', 'z1': 'periodic_z_r'}
    fn = str(tmp_path/'box.pyfrm')
    _gmsh_box(n, spec).write_pyfrm(fn)
    mesh = NativeMesh(fn)
    be = backend_factory()
    print('\n[bc rhs pyfrm]', flush=True)
    integ, orc = _setup(be, 'ns-walls', n, 3, 'rusanov', 0.1, mesh=mesh)
    u0 = orc.ics()
    integ.system.rhs(0.0, 0, 1)
    got = integ.system.ele_scal_upts(1)[0]
    ref = orc.rhs(u0)
    tol = max(1e-12, 10*_floor(orc, u0))
    err = rel_l2(got, ref)
    print(f'[bc rhs pyfrm] rel L2 {err:.3e} (tol {tol:.3e})', flush=True)
    assert err < tol


def _restart(backend_factory, tmp_path, mesh, cfg_fn, dt, exact):
    """Run 10 steps; and 5, write, read into a fresh integrator, 5 more.
    exact: bit for bit (hex kernels have no tuned variants); else the tuner
    of the second backend may pick another dense-kernel variant (another
    summation order), so 1e-13."""
    from pyfr_amd.native import read_pyfrs, write_pyfrs
    from pyfr_amd.solvers import get_integrator
    be = backend_factory()
    a = get_integrator(be, mesh, cfg_fn(be))
    a.run_steps(5, dt)
    fn = str(tmp_path/'soln.pyfrs')
    write_pyfrs(fn, mesh, a)
    a.run_steps(5, dt)
    want = a.soln
    data, cfgtxt, stats = read_pyfrs(fn, mesh)
    assert '[data]' in stats and 'prefix = soln' in stats
    b = get_integrator(backend_factory(), mesh, cfg_fn(be))
    names = b.system.ele_types
    b.system.set_soln_upts(b.idxcurr, [data[t] for t in names])
    b.tcurr = 5*dt
    b.run_steps(5, dt)
    for x, y in zip(want, b.soln):
        if exact:
            assert np.array_equal(x, y)
        else:
            assert rel_l2(y, x) < 1e-13


def test_pyfrs_restart_hex(backend_factory, tmp_path):
    from pyfr_amd.native import NativeMesh
    from tests.fr_cases import make_cfg
    from tests.test_curved import warp
    spec = {s: f'periodic_{s[0]}_{"lr"[int(s[1])]}'
            for s in ('x0', 'x1', 'y0', 'y1', 'z0', 'z1')}
    fn = str(tmp_path/'hex.pyfrm')
    _gmsh_box(4, spec, sorder=2, warp=warp).write_pyfrm(fn)
    mesh = NativeMesh(fn)
    assert 0 < mesh.spts_curved['hex'].sum() < 64
    print('\n[pyfrs restart] hex', flush=True)
    _restart(backend_factory, tmp_path, mesh,
             lambda be: make_cfg('euler', 2, 'rusanov'), 1e-3, True)


def test_pyfrs_restart_tet_prism(backend_factory, tmp_path):
    from pyfr_amd.gmsh import GmshMesh
    from pyfr_amd.mesh import SimplexBoxMesh
    from pyfr_amd.native import NativeMesh
    from tests.msh_files import mesh_to_msh
    from tests.test_gmsh import _sim_sides
    n, nt, npr = 2, 1, 1
    L = 2*np.pi
    gen = SimplexBoxMesh.stacked(n, nt, npr)
    sides = _sim_sides(gen, (L, L, (nt + npr)*L/n))
    fn = str(tmp_path/'tp.pyfrm')
    GmshMesh(io.StringIO(mesh_to_msh(gen, sides))).write_pyfrm(fn)
    mesh = NativeMesh(fn)
    print('\n[pyfrs restart] tet/prism', flush=True)
    _restart(backend_factory, tmp_path, mesh,
             lambda be: make_simplex_cfg(3, be.precision, 'rk4', 2e-3,
                                         mode_ics(seed=3)), 2e-3, False)


def test_pyfrs_restart_pyramid_tet(backend_factory, tmp_path):
    """Pyramids with tets (Gauss-Legendre-Lobatto pyramid points) through
    Gmsh, .pyfrm and a .pyfrs restart."""
    from pyfr_amd.gmsh import GmshMesh
    from pyfr_amd.mesh import SimplexBoxMesh
    from pyfr_amd.native import NativeMesh
    from tests.msh_files import mesh_to_msh
    from tests.test_gmsh import _sim_sides
    L = 2*np.pi
    gen = SimplexBoxMesh([((2, 2, 2), L/2, [('tet', 1), ('pyt', 1)])])
    fn = str(tmp_path/'pyt.pyfrm')
    GmshMesh(io.StringIO(mesh_to_msh(gen, _sim_sides(gen, (L, L, L))))) \
        .write_pyfrm(fn)
    mesh = NativeMesh(fn)

    def cfg(be):
        c = make_simplex_cfg(3, be.precision, 'rk4', 2e-3, mode_ics(seed=3))
        c.set('solver-elements-pyr', 'soln-pts', 'gauss-legendre-lobatto')
        return c
    print('\n[pyfrs restart] pyramid/tet', flush=True)
    _restart(backend_factory, tmp_path, mesh, cfg, 2e-3, False)
