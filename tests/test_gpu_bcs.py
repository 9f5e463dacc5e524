"""GPU parity of the boundary-condition kernels (bccflux, bcconu): one RHS
on hex boxes with non-periodic sides against the oracle, for every Euler
and NS boundary type, with boundary values given as expressions of x, y, z.
Tolerance as for the periodic RHS: max(1e-12, 10 x the oracle's own
1e-16-input-noise floor) in relative L2; RK trajectories 1e-12."""
import numpy as np
import pytest

from tests.fr_cases import eval_ics, make_cfg, oracle_consts, rel_l2

pytestmark = pytest.mark.gpu

# side -> (type, {option: expression}); values exercise the ploc path
FREE = {'rho': '1 + 0.05*sin(z)', 'u': '0.3', 'v': '0.1*cos(z)',
        'w': '-0.05', 'p': '100/1.4 + 0.2*cos(y)'}
CASES = {
    'euler-walls': ('euler', (False, False, True), {
        'x0': ('slp-adia-wall', {}), 'x1': ('slp-adia-wall', {}),
        'y0': ('char-riem-inv', FREE), 'y1': ('char-riem-inv', FREE)}),
    'euler-supersonic': ('euler', (False, True, True), {
        'x0': ('sup-in-fa', FREE), 'x1': ('sup-out-fn', {})}),
    'ns-walls': ('navier-stokes', (False, False, True), {
        'x0': ('no-slp-adia-wall', {}),
        'x1': ('no-slp-isot-wall', {'u': '0.1*sin(z)', 'v': '0', 'w': '0',
                                    'cpTw': '250'}),
        'y0': ('char-riem-inv', FREE), 'y1': ('sub-out-fp', {'p': '71'})}),
    'ns-inflow': ('navier-stokes', (False, False, True), {
        'x0': ('sub-in-frv', {'rho': '1', 'u': '0.2', 'v': '0.05*sin(z)'}),
        'x1': ('sup-out-fn', {}),
        'y0': ('sub-in-ftpttang', {'cpTt': '251', 'pt': '72',
                                   'theta': '20', 'phi': '80'}),
        'y1': ('sup-in-fa', FREE)}),
    'ns-slip': ('navier-stokes', (True, False, False), {
        'y0': ('slp-adia-wall', {}), 'y1': ('slp-adia-wall', {}),
        'z0': ('char-riem-inv', FREE), 'z1': ('sub-out-fp', {'p': '71.5'})}),
}
SCALARS = ('cpTw', 'cpTt', 'pt', 'theta', 'phi')


def _setup(be, case, n, order, rsolver, tau=0.0, scheme='rk4', mesh=None):
    from oracle.rhs import OracleHexSystem
    from pyfr_amd.mesh import BoxMesh
    from pyfr_amd.solvers import get_integrator
    system, periodic, bcs = CASES[case]
    cfg = make_cfg(system, order, rsolver, scheme=scheme, tau=tau)
    cfg.set('backend-hip', 'fused', str(be.fused))
    oc = oracle_consts(system, tau=tau)
    obcs = {}
    for side, (bctype, opts) in bcs.items():
        sect = f'soln-bcs-{side}'
        cfg.set(sect, 'type', bctype)
        vals = {}
        for k, v in opts.items():
            cfg.set(sect, k, v)
            if k in SCALARS:
                oc[k] = float(eval(v))
            else:
                vals[k] = eval_ics({k: v})[k]
        # defaults of the no-slip isothermal wall and sub-in-frv velocity
        if bctype in ('no-slp-isot-wall', 'sub-in-frv'):
            for k in 'uvw':
                vals.setdefault(k, 0.0)
        obcs[side] = (bctype, vals)
    mesh = mesh or BoxMesh(n, periodic=periodic)
    integ = get_integrator(be, mesh, cfg)
    orc = OracleHexSystem(n, order, system, rsolver, constants=oc,
                          periodic=periodic, bcs=obcs)
    return integ, orc


def _floor(orc, u, seed=0):
    rng = np.random.default_rng(seed)
    up = u*(1 + 1e-16*rng.standard_normal(u.shape))
    return rel_l2(orc.rhs(up), orc.rhs(u))


@pytest.mark.parametrize('fused', [True, False])
@pytest.mark.parametrize('case,n,order,rsolver,tau', [
    ('euler-walls', 4, 2, 'rusanov', 0.0),
    ('euler-walls', 3, 3, 'hllc', 0.0),
    ('euler-supersonic', 4, 2, 'roe', 0.0),
    ('ns-walls', 3, 3, 'rusanov', 0.0),
    ('ns-walls', 3, 2, 'hllc', 0.1),
    ('ns-inflow', 3, 2, 'rusanov', 0.1),
    ('ns-inflow', 3, 3, 'hll', 0.0),
    ('ns-slip', 3, 2, 'rusanov', 0.0),
    ('euler-walls', 3, 2, 'exact', 0.0),
    ('ns-walls', 3, 2, 'exact', 0.1),
])
def test_bc_rhs_parity(backend_factory, case, n, order, rsolver, tau, fused):
    be = backend_factory(fused=fused)
    print(f'\n[bc rhs] {case} n={n} p={order} {rsolver} tau={tau} '
          f'fused={fused}', flush=True)
    integ, orc = _setup(be, case, n, order, rsolver, tau)
    assert len(integ.system._bc_inters) == len(orc.bcs)
    u0 = orc.ics()
    assert rel_l2(integ.system.ele_scal_upts(0)[0], u0) < 1e-14
    integ.system.rhs(0.0, 0, 1)
    got = integ.system.ele_scal_upts(1)[0]
    ref = orc.rhs(u0)
    tol = max(1e-12, 10*_floor(orc, u0))
    err = rel_l2(got, ref)
    print(f'[bc rhs] rel L2 {err:.3e} (tol {tol:.3e})', flush=True)
    assert err < tol


def test_bc_rk45_trajectory(backend_factory):
    from oracle.rhs import advance
    be = backend_factory()
    integ, orc = _setup(be, 'ns-walls', 3, 2, 'rusanov', scheme='rk45')
    integ.dt = 2e-4
    integ.advance_to(10*2e-4)
    u_ref, nsteps = advance(orc.rhs, orc.ics(), 2e-4, 10*2e-4, 'rk45')
    assert integ.nsteps == nsteps
    err = rel_l2(integ.soln[0], u_ref)
    print(f'\n[bc rk45] {nsteps} steps rel L2 {err:.3e}', flush=True)
    assert err < 1e-12


def test_bc_declines_time_dependent_values(backend_factory):
    from pyfr_amd.backend import NotSuitableError
    from pyfr_amd.mesh import BoxMesh
    from pyfr_amd.solvers import get_integrator
    be = backend_factory()
    cfg = make_cfg('euler', 2)
    for side in ('x0', 'x1'):
        cfg.set(f'soln-bcs-{side}', 'type', 'sup-in-fa')
        for k, v in FREE.items():
            cfg.set(f'soln-bcs-{side}', k, v)
    cfg.set('soln-bcs-x0', 'u', '0.3 + 0.1*t')
    with pytest.raises(NotSuitableError):
        get_integrator(be, BoxMesh(3, periodic=(False, True, True)), cfg)


@pytest.mark.parametrize('version', ['2.2', '4.1'])
def test_bc_rhs_parity_gmsh_mesh(backend_factory, version):
    """The same boundaries on a mesh read from a .msh file (periodic z as a
    periodic_z_l/r surface pair, the x and y sides as named surfaces)."""
    import io
    from pyfr_amd.gmsh import GmshMesh
    from pyfr_amd.mesh import BoxMesh
    from tests.msh_files import box_sides, mesh_to_msh
    n = 3
    raw = BoxMesh(n, periodic=(False, False, False))
    spec = {'x0': 'x0', 'x1': 'x1', 'y0': 'y0', 'y1': 'y1',
            'z0': 'periodic_z_l', 'z1': 'periodic_z_r'}
    mesh = GmshMesh(io.StringIO(mesh_to_msh(raw, box_sides(raw, spec),
                                            version)))
    be = backend_factory()
    print(f'\n[bc rhs gmsh] {version}', flush=True)
    integ, orc = _setup(be, 'ns-walls', n, 3, 'rusanov', 0.1, mesh=mesh)
    u0 = orc.ics()
    integ.system.rhs(0.0, 0, 1)
    got = integ.system.ele_scal_upts(1)[0]
    ref = orc.rhs(u0)
    tol = max(1e-12, 10*_floor(orc, u0))
    err = rel_l2(got, ref)
    print(f'[bc rhs gmsh] rel L2 {err:.3e} (tol {tol:.3e})', flush=True)
    assert err < tol
