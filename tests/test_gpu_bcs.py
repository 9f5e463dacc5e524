"""GPU parity of the boundary-condition kernels (bccflux, bcconu): one RHS
on hex boxes with non-periodic sides against the oracle, for every Euler
and NS boundary type, with boundary values given as expressions of x, y, z
and, evaluated on the device at each RHS, of t.
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


# side -> (type, {option: expression of x, y, z, t}); the oracle's values
# are the same expressions as callables of (x, y, z, t)
TIMED = {
    'euler': ('euler', (False, True, True), {
        'x0': ('sup-in-fa', {'rho': '1 + 0.05*sin(z)*cos(3*t)',
                             'u': '0.3 + 0.1*sin(2*t)', 'v': '0.1*cos(z)',
                             'w': '-0.05*t', 'p': '100/1.4 + 0.2*cos(y + t)'}),
        'x1': ('char-riem-inv', {'rho': '1', 'u': '0.3*exp(-t)', 'v': '0',
                                 'w': '0', 'p': 'pow(1 + t, 0.5)*100/1.4'})}),
    'navier-stokes': ('navier-stokes', (False, False, True), {
        'x0': ('no-slp-isot-wall', {'u': '0.1*sin(z)*min(t, 0.5)',
                                    'v': '0', 'w': '0', 'cpTw': '250'}),
        'x1': ('sup-out-fn', {}),
        'y0': ('sub-in-frv', {'rho': '1', 'u': '0.2*tanh(4*t)',
                              'v': '0.05*sin(z - t)'}),
        'y1': ('sub-out-fp', {'p': '71 + atan(t)'})}),
}
NPSYMS = {'__builtins__': {}, 'exp': np.exp, 'sin': np.sin, 'cos': np.cos,
          'tanh': np.tanh, 'pow': np.power, 'min': np.minimum,
          'atan': np.arctan}


def _timed_setup(be, system, n, order, scheme='rk4'):
    from oracle.rhs import OracleHexSystem
    from pyfr_amd.mesh import BoxMesh
    from pyfr_amd.solvers import get_integrator
    sysname, periodic, bcs = TIMED[system]
    cfg = make_cfg(sysname, order, scheme=scheme, dt=2e-4)
    cfg.set('backend-hip', 'fused', str(be.fused))
    oc = oracle_consts(sysname)
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
                vals[k] = (lambda e: lambda x, y, z, t: eval(
                    e, NPSYMS, {'x': x, 'y': y, 'z': z, 't': t}) + 0*x)(v)
        if bctype in ('no-slp-isot-wall', 'sub-in-frv'):
            for k in 'uvw':
                vals.setdefault(k, 0.0)
        obcs[side] = (bctype, vals)
    integ = get_integrator(be, BoxMesh(n, periodic=periodic), cfg)
    orc = OracleHexSystem(n, order, sysname, constants=oc,
                          periodic=periodic, bcs=obcs)
    return integ, orc


@pytest.mark.parametrize('fused', [True, False])
@pytest.mark.parametrize('system', ['euler', 'navier-stokes'])
def test_bc_time_dependent_rhs(backend_factory, system, fused):
    """Boundary values that depend on t: compiled expression programs
    evaluated on the device at each RHS, t set before the (captured) RHS;
    two RHS at different t through the same graph."""
    be = backend_factory(fused=fused)
    print(f'\n[bc timed] {system} fused={fused}', flush=True)
    integ, orc = _timed_setup(be, system, 3, 2)
    b = integ.system._bc_inters
    assert any(i.time_dependent for i in b)
    u0 = orc.ics()
    for t in (0.37, 1.25):
        integ.system.rhs(t, 0, 1)
        got = integ.system.ele_scal_upts(1)[0]
        ref = orc.rhs(u0, t)
        tol = max(1e-12, 10*_floor(orc, u0))
        err = rel_l2(got, ref)
        print(f'[bc timed] t={t} rel L2 {err:.3e} (tol {tol:.3e})',
              flush=True)
        assert err < tol
        assert rel_l2(orc.rhs(u0, 0.0), ref) > 1e-6   # t matters


@pytest.mark.parametrize('scheme', ['rk4', 'rk45'])
def test_bc_time_dependent_trajectory(backend_factory, scheme):
    """RK stage times reach the boundary values: 10 steps against the
    oracle stepping with the reference's stage times."""
    from oracle.rhs import advance
    be = backend_factory()
    integ, orc = _timed_setup(be, 'navier-stokes', 3, 2, scheme=scheme)
    integ.dt = 2e-3
    integ.advance_to(10*2e-3)
    u_ref, nsteps = advance(orc.rhs, orc.ics(), 2e-3, 10*2e-3, scheme,
                            timed=True)
    u_untimed, _ = advance(orc.rhs, orc.ics(), 2e-3, 10*2e-3, scheme)
    assert integ.nsteps == nsteps
    err = rel_l2(integ.soln[0], u_ref)
    print(f'\n[bc timed {scheme}] {nsteps} steps rel L2 {err:.3e} '
          f'(vs t = 0 values: {rel_l2(u_untimed, u_ref):.3e})', flush=True)
    assert err < 1e-12
    assert rel_l2(u_untimed, u_ref) > 1e3*max(err, 1e-15)


def test_bc_program_validation(backend_factory):
    """Malformed programs are refused on the host before any device
    work: bad opcode, stack underflow, constant index, depth, leftovers."""
    import ctypes as ct
    from pyfr_amd._lib import FrHipError
    be = backend_factory()
    pool = np.array([1.0, 2.0])

    def make(code, offs):
        code = np.asarray(code, dtype=np.int32).reshape(-1, 2)
        offs = np.asarray(offs, dtype=np.int32)
        p = ct.c_void_p()
        be.lib.bc_program_create(ct.byref(p), len(offs) - 1,
                                 code.ctypes.data, offs.ctypes.data,
                                 pool.ctypes.data, len(pool))
        be.lib.bc_program_destroy(p.value)

    make([(0, 0), (1, 3), (2, 0)], [0, 3])          # 1.0 + t: accepted
    for code, offs in [([(0, 0), (99, 0)], [0, 2]),  # bad opcode
                       ([(0, 0), (2, 0)], [0, 2]),   # underflow
                       ([(0, 5)], [0, 1]),           # constant index
                       ([(1, 4)], [0, 1]),           # variable index
                       ([(0, 0), (0, 1)], [0, 2]),   # two results
                       ([(0, 0)]*17 + [(2, 0)]*16, [0, 33])]:  # depth 17
        with pytest.raises(FrHipError):
            make(code, offs)


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


def _controlled_setup(be, system, scheme):
    from oracle.rhs import OracleHexSystem
    from pyfr_amd.mesh import BoxMesh
    from pyfr_amd.solvers import get_integrator
    cfg = make_cfg(system, 2, scheme=scheme, dt=2e-4)
    cfg.set('backend-hip', 'fused', str(be.fused))
    sides = {
        'x0': ('char-riem-inv-mass-flow', 'mass-flow-rate', -5.0,
               {'rho': '1', 'u': '0.3', 'v': '0', 'w': '0.05',
                'p': '71.4'}, 5, 2.0),
        'x1': ('char-riem-inv-pressure', 'pressure', 71.0,
               {'rho': '1', 'u': '0.3', 'v': '0', 'w': '0'}, 4, 1.0)}
    obcs = {}
    for side, (bctype, topt, target, vals, qdeg, eta) in sides.items():
        sect = f'soln-bcs-{side}'
        cfg.set(sect, 'type', bctype)
        for k, v in vals.items():
            cfg.set(sect, k, v)
        for k, v in ((topt, target), ('alpha', 0.5), ('eta', eta),
                     ('nsteps', 3), ('quad-deg', qdeg)):
            cfg.set(sect, k, str(v))
        ov = {k: float(eval(v)) for k, v in vals.items() if k != 'p'}
        ov['ctrl'] = {'target': target, 'alpha': 0.5, 'eta': eta,
                      'nsteps': 3, 'qdeg': qdeg,
                      'p': float(eval(vals.get('p', '0')))}
        obcs[side] = (bctype, ov)
    periodic = (False, True, True)
    integ = get_integrator(be, BoxMesh(3, periodic=periodic), cfg)
    orc = OracleHexSystem(3, 2, system, constants=oracle_consts(system),
                          periodic=periodic, bcs=obcs)
    return integ, orc


@pytest.mark.parametrize('scheme', ['rk4', 'rk45'])
@pytest.mark.parametrize('system', ['euler', 'navier-stokes'])
def test_bc_controlled_trajectory(backend_factory, system, scheme):
    """char-riem-inv-mass-flow and char-riem-inv-pressure: the controller
    measures the boundary's mass flow / mean pressure by a surface
    integral of each RHS's input state every nsteps (3) RHS evaluations
    and moves the boundary pressure; 10 steps against the oracle stepping
    with the reference's stage times.  (With nsteps = 2, RK4 would update
    at its third stage, whose time equals the second's: the reference's
    controller then divides by a zero time step.)"""
    from oracle.rhs import advance
    be = backend_factory()
    integ, orc = _controlled_setup(be, system, scheme)
    integ.dt = 2e-4
    integ.advance_to(10*2e-4)
    u_ref, nsteps = advance(orc.rhs, orc.ics(), 2e-4, 10*2e-4, scheme,
                            timed=True)
    assert integ.nsteps == nsteps
    err = rel_l2(integ.soln[0], u_ref)
    ctrls = [b._ctrl for b in integ.system._bc_inters]
    print(f'\n[bc controlled {system} {scheme}] rel L2 {err:.3e}; '
          + ', '.join(f'{c.kind} p = {c.interp_c:.6f} + {c.interp_m:.4f} t'
                      for c in ctrls), flush=True)
    assert err < 1e-12
    for c, side in zip(ctrls, ('x0', 'x1')):
        oc = orc._ctrls[side]
        assert c.interp_m != 0.0
        assert abs(c.interp_m - oc.interp_m) <= 1e-9*abs(oc.interp_m)
        assert abs(c.interp_c - oc.interp_c) <= 1e-12*abs(oc.interp_c)


def test_bc_controlled_surface_integral(backend_factory):
    """The product's surface integral of a uniform flow: mass flow
    -rho u A through the x0 side, mean pressure p."""
    be = backend_factory()
    integ, _ = _controlled_setup(be, 'euler', 'rk4')
    L = 2*np.pi
    sysm = integ.system
    u = {}
    for et, e in sysm.ele_map.items():
        a = np.empty((e.nupts, 5, e.neles))
        a[:, 0], a[:, 1], a[:, 2:4] = 1.2, 1.2*0.3, 0.0
        a[:, 4] = 0.9/0.4 + 0.5*1.2*0.09
        u[et] = a
    mf, pr = (b._ctrl for b in sysm._bc_inters)
    assert abs(mf._measure(u) + 1.2*0.3*L*L) < 1e-12*L*L
    assert abs(pr._measure(u) - 0.9) < 1e-13
    assert abs(pr.area - L*L) < 1e-12*L*L
