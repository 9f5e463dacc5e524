"""GPU parity on curved hex meshes (shape order 2-3, smooth warp): one RHS
and an RK4 trajectory against the oracle, with the curved region on the
stored-metric tflux kernels and the linear region on the in-kernel vertex
metric.  Tolerances as for the periodic hex RHS."""
import numpy as np
import pytest

from tests.fr_cases import make_cfg, oracle_consts, rel_l2
from tests.test_curved import warp

pytestmark = pytest.mark.gpu


def _setup(be, system, n, order, k, scheme='rk4', mesh=None):
    from oracle.rhs import OracleHexSystem
    from pyfr_amd.mesh import BoxMesh
    from pyfr_amd.solvers import get_integrator
    cfg = make_cfg(system, order, scheme=scheme)
    cfg.set('backend-hip', 'fused', str(be.fused))
    if mesh is None:
        mesh = BoxMesh(n, sorder=k, warp=warp)
    integ = get_integrator(be, mesh, cfg)
    e = integ.system.ele_map['hex']
    # the product's curved region: the curved elements, rounded up to a
    # whole 64-element block (pyfr/solvers/base/elements.py:224-227)
    curved = np.arange(mesh.neles) < e.linoff
    orc = OracleHexSystem(n, order, system, constants=oracle_consts(system),
                          verts=mesh.spts['hex'].transpose(0, 2, 1),
                          curved=curved)
    return integ, orc, e


@pytest.mark.parametrize('fused', [True, False])
@pytest.mark.parametrize('system,n,order,k', [('euler', 6, 2, 2),
                                              ('euler', 6, 3, 3),
                                              ('navier-stokes', 6, 3, 2),
                                              ('euler', 4, 1, 3)])
def test_curved_rhs_parity(backend_factory, system, n, order, k, fused):
    be = backend_factory(fused=fused)
    print(f'\n[curved rhs] {system} n={n} p={order} k={k} fused={fused}',
          flush=True)
    integ, orc, e = _setup(be, system, n, order, k)
    regions = e.mesh_regions
    print(f'[curved rhs] regions {regions}', flush=True)
    if order >= 2:
        assert set(regions) == {'curved', 'linear'}
    u0 = orc.ics()
    assert rel_l2(integ.system.ele_scal_upts(0)[0], u0) < 1e-14
    integ.system.rhs(0.0, 0, 1)
    got = integ.system.ele_scal_upts(1)[0]
    ref = orc.rhs(u0)
    rng = np.random.default_rng(0)
    floor = rel_l2(orc.rhs(u0*(1 + 1e-16*rng.standard_normal(u0.shape))),
                   ref)
    tol = max(1e-12, 10*floor)
    err = rel_l2(got, ref)
    print(f'[curved rhs] rel L2 {err:.3e} (tol {tol:.3e})', flush=True)
    assert err < tol


def test_curved_rk4_trajectory(backend_factory):
    from oracle.rhs import rk4_step
    be = backend_factory()
    integ, orc, _ = _setup(be, 'euler', 6, 3, 3)
    integ.run_steps(10, 1e-3)
    u = orc.ics()
    for _ in range(10):
        u = rk4_step(orc.rhs, u, 1e-3)
    err = rel_l2(integ.soln[0], u)
    print(f'\n[curved rk4] 10 steps rel L2 {err:.3e}', flush=True)
    assert err < 1e-12


@pytest.mark.parametrize('k', [2, 3])
def test_curved_rhs_gmsh_import(backend_factory, k):
    """The curved mesh read from a high-order .msh file (Gmsh type 12 /
    92 hexes, periodic surface pairs): one RHS against the oracle on the
    shape points and curved region as read."""
    import io
    from pyfr_amd.gmsh import GmshMesh
    from pyfr_amd.mesh import BoxMesh
    from tests.msh_files import box_sides, mesh_to_msh
    from tests.test_gmsh import _hex_spec
    n, order = 6, 2
    raw = BoxMesh(n, periodic=(False, False, False), sorder=k, warp=warp)
    mesh = GmshMesh(io.StringIO(mesh_to_msh(
        raw, box_sides(raw, _hex_spec((True, True, True))), '4.1')))
    assert 0 < mesh.spts_curved['hex'].sum() < n**3
    be = backend_factory()
    print(f'\n[curved gmsh] k={k}', flush=True)
    integ, orc, e = _setup(be, 'euler', n, order, k, mesh=mesh)
    assert set(e.mesh_regions) == {'curved', 'linear'}
    u0 = orc.ics()
    integ.system.rhs(0.0, 0, 1)
    got = integ.system.ele_scal_upts(1)[0]
    ref = orc.rhs(u0)
    rng = np.random.default_rng(0)
    floor = rel_l2(orc.rhs(u0*(1 + 1e-16*rng.standard_normal(u0.shape))),
                   ref)
    tol = max(1e-12, 10*floor)
    err = rel_l2(got, ref)
    print(f'[curved gmsh] rel L2 {err:.3e} (tol {tol:.3e})', flush=True)
    assert err < tol
