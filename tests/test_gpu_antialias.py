"""GPU parity of anti-aliasing on hexes against the oracle, on warped and
curved meshes, fused and unfused.  Flux anti-aliasing: the flux at the
element quadrature points (M7), the projected divergence ((M1 - M3 M2) M9)
and, for NS, the unfused gradient path (gradcoru, gradcoru_qpts, unfused
tflux).  Surface-flux anti-aliasing: flux points at the face quadrature
rule (dense M0 / M3 operators, interfaces on the new points), alone and
with flux anti-aliasing.  Tolerances as for the periodic hex RHS."""
import numpy as np
import pytest

from tests.fr_cases import make_cfg, oracle_consts, rel_l2
from tests.test_curved import warp

pytestmark = pytest.mark.gpu


def _setup(be, system, n, order, k, nq=None, aa=('flux',), nf=None):
    from oracle.rhs import OracleHexSystem
    from pyfr_amd.mesh import BoxMesh
    from pyfr_amd.solvers import get_integrator
    cfg = make_cfg(system, order)
    cfg.set('backend-hip', 'fused', str(be.fused))
    cfg.set('solver', 'anti-alias', ', '.join(aa))
    if nq:
        cfg.set('solver-elements-hex', 'quad-npts', nq**3)
    if nf:
        cfg.set('solver-interfaces-quad', 'quad-npts', nf**2)
    mesh = BoxMesh(n, sorder=k, warp=warp)
    integ = get_integrator(be, mesh, cfg)
    e = integ.system.ele_map['hex']
    curved = np.arange(mesh.neles) < e.linoff
    orc = OracleHexSystem(n, order, system, constants=oracle_consts(system),
                          verts=mesh.spts['hex'].transpose(0, 2, 1),
                          curved=curved, antialias=aa, nq1=nq, nf1=nf)
    return integ, orc


SF = ('surf-flux',)
BOTH = ('flux', 'surf-flux')


@pytest.mark.parametrize('fused', [True, False])
@pytest.mark.parametrize('system,n,order,k,nq,aa,nf', [
    ('euler', 4, 2, 1, None, ('flux',), None),
    ('euler', 6, 3, 2, None, ('flux',), None),
    ('euler', 4, 2, 1, 5, ('flux',), None),
    ('navier-stokes', 4, 2, 1, None, ('flux',), None),
    ('navier-stokes', 6, 2, 2, None, ('flux',), None),
    ('navier-stokes', 4, 1, 1, None, ('flux',), None),
    ('euler', 4, 2, 1, None, SF, None), ('euler', 6, 3, 2, None, SF, 6),
    ('navier-stokes', 4, 2, 1, None, SF, None),
    ('navier-stokes', 6, 3, 2, None, SF, None),
    ('euler', 4, 2, 1, None, BOTH, None),
    ('navier-stokes', 4, 2, 2, None, BOTH, None)])
def test_antialias_rhs_parity(backend_factory, system, n, order, k, nq, aa,
                              nf, fused):
    be = backend_factory(fused=fused)
    print(f'\n[aa rhs] {system} n={n} p={order} k={k} nq={nq} aa={aa} '
          f'nf={nf} fused={fused}', flush=True)
    integ, orc = _setup(be, system, n, order, k, nq, aa, nf)
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
    print(f'[aa rhs] rel L2 {err:.3e} (tol {tol:.3e})', flush=True)
    assert err < tol
