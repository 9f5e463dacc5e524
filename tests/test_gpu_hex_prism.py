"""GPU parity on meshes of hexes and prisms together: hexes below a slab of
prisms extruded along x, meeting on whole quad faces (a hex core with
prism layers).  The hexes run the tensor-product kernels, the prisms the
dense matrix-core kernels; the interface kernels pair flux points across
the two types.  Euler and NS RHS (periodic, and with boundaries on hex and
prism faces), flux and surface-flux anti-aliasing, curved (warped)
meshes of both types, RK4 and RK45
trajectories, FP64 and FP32, against the oracle's mixed-mesh system with
its hex element type.

Tolerances as for the tet/prism tests: FP64 RHS max(1e-12, 10 x the
oracle's 1e-16-input-noise floor), trajectories 1e-12; FP32 1e-4."""
import numpy as np
import pytest

from tests.fr_cases import eval_ics, make_simplex_cfg, mode_ics, rel_l2
from tests.test_gpu_simplex import MU, SIMPLEX_BCS, _cat, _noise_floor, \
    _rk4_dict
from tests.test_hex_prism import _tables, warp

pytestmark = pytest.mark.gpu


def _hex_prism(be, n, nh, npr, order, system='euler', bcs=False,
               rsolver='rusanov', beta=0.0, tau=0.0, scheme='rk4', dt=1e-3,
               aa=(), sorder=1, warp=None):
    from oracle.simplex import OracleMixedSystem
    from pyfr_amd.mesh import SimplexBoxMesh
    from pyfr_amd.solvers import get_integrator
    ics = mode_ics(seed=6)
    cfg = make_simplex_cfg(order, be.precision, scheme, dt, ics=ics,
                           rsolver=rsolver, system=system, beta=beta,
                           tau=tau)
    cfg.set('backend-hip', 'fused', str(be.fused))
    if aa:
        cfg.set('solver', 'anti-alias', ', '.join(aa))
    c = {'gamma': 1.4}
    if system == 'navier-stokes':
        cfg.set('constants', 'mu', repr(MU))
        c |= {'mu': MU, 'Pr': 0.71, 'ldg-beta': beta, 'ldg-tau': tau}
    periodic = (False, True, False) if bcs else (True, True, True)
    obcs = {}
    if bcs:
        for side, (bctype, opts) in SIMPLEX_BCS[system].items():
            sect = f'soln-bcs-{side}'
            cfg.set(sect, 'type', bctype)
            vals = {}
            for k, v in opts.items():
                cfg.set(sect, k, v)
                if k == 'cpTw':
                    c[k] = float(eval(v))
                else:
                    vals[k] = eval_ics({k: v})[k]
            obcs[side] = (bctype, vals)
    mesh = SimplexBoxMesh.hex_prism(n, nh, npr, periodic=periodic,
                                    sorder=sorder, warp=warp)
    integ = get_integrator(be, mesh, cfg)
    L = 2*np.pi
    verts = {et: v.transpose(0, 2, 1) for et, v in mesh.spts.items()}
    idx = {et: np.arange(v.shape[1]) for et, v in mesh.spts.items()}
    orc = OracleMixedSystem(verts, order, {et: _tables(et, order, aa)
                                           for et in verts},
                            [((L, L, (nh + npr)*L/n), idx)],
                            rsolver=rsolver, system=system, constants=c,
                            periodic=periodic, bcs=obcs,
                            antialias=tuple(a for a in aa if a == 'flux'))
    return integ, orc, orc.ics(eval_ics(ics))


def _rhs_err(integ, orc, u0):
    sysm = integ.system
    names = sysm.ele_types
    assert names == ['hex', 'pri']
    got0 = dict(zip(names, sysm.ele_scal_upts(0)))
    assert rel_l2(_cat(got0, names), _cat(u0, names)) < 1e-7
    sysm.rhs(0.0, 0, 1)
    got = dict(zip(names, sysm.ele_scal_upts(1)))
    ref = orc.rhs(u0)
    return (rel_l2(_cat(got, names), _cat(ref, names)),
            max(1e-12, 10*_noise_floor(orc, u0)))


@pytest.mark.parametrize('fused', ['flux', 'div', 'none'])
@pytest.mark.parametrize('order', [1, 2, 3, 4])
def test_euler_rhs_parity(backend_factory, order, fused):
    be = backend_factory(fused=fused != 'none', fuse_flux=fused == 'flux')
    print(f'\n[hex-prism euler] p={order} fused={fused}', flush=True)
    integ, orc, u0 = _hex_prism(be, 3, 2, 1, order)
    err, tol = _rhs_err(integ, orc, u0)
    print(f'[hex-prism euler] rel L2 {err:.3e} (tol {tol:.3e})', flush=True)
    assert err < tol


@pytest.mark.parametrize('fused', [True, False])
@pytest.mark.parametrize('order,beta,tau', [(2, 0.0, 0.0), (3, 0.5, 0.1),
                                            (4, -0.5, 0.0)])
def test_ns_rhs_parity(backend_factory, order, beta, tau, fused):
    be = backend_factory(fused=fused)
    print(f'\n[hex-prism ns] p={order} beta={beta} tau={tau} '
          f'fused={fused}', flush=True)
    integ, orc, u0 = _hex_prism(be, 2, 1, 1, order, 'navier-stokes',
                                beta=beta, tau=tau)
    err, tol = _rhs_err(integ, orc, u0)
    print(f'[hex-prism ns] rel L2 {err:.3e} (tol {tol:.3e})', flush=True)
    assert err < tol


@pytest.mark.parametrize('system,order,rsolver', [
    ('euler', 2, 'hllc'), ('euler', 3, 'roe'),
    ('navier-stokes', 2, 'rusanov'), ('navier-stokes', 3, 'hll')])
def test_bc_rhs_parity(backend_factory, system, order, rsolver):
    """Periodic in y only: the x sides cut both hexes and prisms (quad
    faces of both types), the bottom is hex faces, the top prism faces."""
    be = backend_factory()
    print(f'\n[hex-prism bc] {system} p={order} {rsolver}', flush=True)
    integ, orc, u0 = _hex_prism(be, 3, 2, 1, order, system, bcs=True,
                                rsolver=rsolver)
    assert len(integ.system._bc_inters) == 4
    err, tol = _rhs_err(integ, orc, u0)
    print(f'[hex-prism bc] rel L2 {err:.3e} (tol {tol:.3e})', flush=True)
    assert err < tol


@pytest.mark.parametrize('fused', [True, False])
@pytest.mark.parametrize('system,order,aa,bcs', [
    ('euler', 2, ('flux',), False), ('euler', 3, ('surf-flux',), False),
    ('euler', 2, ('flux', 'surf-flux'), True),
    ('navier-stokes', 2, ('flux',), False),
    ('navier-stokes', 3, ('surf-flux',), True),
    ('navier-stokes', 2, ('flux', 'surf-flux'), False)])
def test_antialias_rhs_parity(backend_factory, system, order, aa, bcs,
                              fused):
    """Each type on its own default rules: Gauss-Legendre (p + 2)^3 in the
    hexes, the reference's default prism rule, and (p + 2)^2 GL points on
    every quad face, hex and prism alike, so the faces still pair."""
    be = backend_factory(fused=fused)
    print(f'\n[hex-prism aa] {system} p={order} aa={aa} bcs={bcs} '
          f'fused={fused}', flush=True)
    integ, orc, u0 = _hex_prism(be, 3, 2, 1, order, system, bcs=bcs, aa=aa)
    err, tol = _rhs_err(integ, orc, u0)
    print(f'[hex-prism aa] rel L2 {err:.3e} (tol {tol:.3e})', flush=True)
    assert err < tol


def _warp_bc(x, y, z):
    """Moves no point off the non-periodic x and z sides."""
    return (x, y + 0.08*np.sin(x)*np.sin(z), z)


@pytest.mark.parametrize('fused', [True, False])
@pytest.mark.parametrize('system,order,sorder,extra', [
    ('euler', 1, 2, None), ('euler', 2, 2, None), ('euler', 3, 3, None),
    ('navier-stokes', 2, 2, None), ('navier-stokes', 3, 2, None),
    ('euler', 2, 2, 'bcs'), ('navier-stokes', 2, 2, 'bcs'),
    ('euler', 2, 2, 'flux')])
def test_curved_rhs_parity(backend_factory, system, order, sorder, extra,
                           fused):
    """A warped box of order-k hexes and prisms, all curved: the stored
    invariant-form metric in both types, curved normals on both sides of
    every hex-prism face; with boundaries and with flux anti-aliasing."""
    be = backend_factory(fused=fused)
    print(f'\n[hex-prism curved] {system} p={order} k={sorder} {extra} '
          f'fused={fused}', flush=True)
    bcs = extra == 'bcs'
    integ, orc, u0 = _hex_prism(be, 3, 2, 1, order, system, bcs=bcs,
                                aa=('flux',) if extra == 'flux' else (),
                                sorder=sorder, warp=_warp_bc if bcs else warp)
    e = integ.system.ele_map
    # the linear region starts at or past the end (rounded up to a block)
    assert e['hex'].linoff >= e['hex'].neles and e['pri'].curved_simplex
    err, tol = _rhs_err(integ, orc, u0)
    print(f'[hex-prism curved] rel L2 {err:.3e} (tol {tol:.3e})', flush=True)
    assert err < tol


@pytest.mark.parametrize('scheme', ['rk4', 'rk45'])
def test_trajectory(backend_factory, scheme):
    from oracle.rhs import rk4_step, rk45_step
    be = backend_factory()
    integ, orc, u0 = _hex_prism(be, 3, 2, 1, 3, scheme=scheme, dt=2e-3)
    names = integ.system.ele_types
    print(f'\n[hex-prism {scheme}] 10 steps', flush=True)
    integ.run_steps(10, 2e-3)
    step = rk4_step if scheme == 'rk4' else rk45_step
    u = u0
    for _ in range(10):
        u = _rk4_dict(orc.rhs, u, 2e-3, step)
    err = rel_l2(_cat(dict(zip(names, integ.soln)), names), _cat(u, names))
    print(f'[hex-prism {scheme}] rel L2 {err:.3e}', flush=True)
    assert err < 1e-12


@pytest.mark.parametrize('precision', ['double', 'single'])
@pytest.mark.parametrize('npr,order', [(0, 2), (0, 3), (1, 3)])
def test_rk4_fused_stages(backend_factory, precision, npr, order):
    """RK4 with each stage's updates in the divergence epilogues: the
    generated hex kernel (and the prisms' flux-fused kernel) against the
    reference kernel sequence with separate axnpby updates, 10 steps on an
    all-hex box and on the hex/prism block."""
    be = backend_factory(precision=precision)
    print(f'\n[hex rk4 fused] {precision} prism layers={npr} p={order}',
          flush=True)
    integ, orc, u0 = _hex_prism(be, 3, 3 - npr, npr, order, dt=2e-3)
    assert integ.system.fused_rk4
    integ.run_steps(10, 2e-3)
    ref_be = backend_factory(precision=precision, fused=False)
    ref, _, _ = _hex_prism(ref_be, 3, 3 - npr, npr, order, dt=2e-3)
    assert not ref.system.fused_rk4
    ref.run_steps(10, 2e-3)
    names = integ.system.ele_types
    got = _cat(dict(zip(names, integ.soln)), names)
    err = rel_l2(got, _cat(dict(zip(names, ref.soln)), names))
    print(f'[hex rk4 fused] vs unfused rel L2 {err:.3e}', flush=True)
    assert err < (1e-14 if precision == 'double' else 1e-5)
    assert integ.nrhsevals == ref.nrhsevals == 40
    if precision == 'double':
        from oracle.rhs import rk4_step
        u = u0
        for _ in range(10):
            u = _rk4_dict(orc.rhs, u, 2e-3, rk4_step)
        assert rel_l2(got, _cat(u, names)) < 1e-12


@pytest.mark.parametrize('precision', ['double', 'single'])
def test_mul2_rk4_epilogue_needs_structured_plan(backend_factory, precision):
    """A dense two-input plan refuses the RK4 epilogues on the host,
    before any launch, and so does the provider (NotSuitableError)."""
    from pyfr_amd._lib import FrHipError
    from pyfr_amd.backend import NotSuitableError
    from tests.test_gpu_simplex import _operands, _Plan2
    be = backend_factory(precision=precision)
    A0, A1, b0, b1, c, rng = _operands(be, 35, 105, 60, 320, 5)
    p = _Plan2(be, A0, A1)
    assert p.kind == 2
    rcp = be.const_matrix(np.ones((1, 320)), tags={'align'})
    print(f'\n[mul2 rk4 dense] {precision}', flush=True)
    with pytest.raises(FrHipError, match='structured'):
        be.lib.mul2(p.plan, c.ncol, b0.data, b0.leaddim, b1.data, b1.leaddim,
                    c.data, c.leaddim, 10, rcp.data, 0, c.data, 1e-3, 0.0, 0,
                    be.stream)
    a0, a1 = be.const_matrix(A0), be.const_matrix(A1)
    with pytest.raises(NotSuitableError):
        be.kernel('tdivf', a0, a1, b0, b1, c, rcp, xa=(c, c), stage='rk4-1')


def test_fp32_rhs(backend_factory):
    be = backend_factory(precision='single')
    print('\n[hex-prism fp32] p=4', flush=True)
    integ, orc, u0 = _hex_prism(be, 3, 2, 1, 4)
    err, _ = _rhs_err(integ, orc, u0)
    print(f'[hex-prism fp32] rel L2 {err:.3e}', flush=True)
    assert err < 1e-4


@pytest.mark.parametrize('system,scheme', [('euler', 'rk4'), ('euler', 'rk45'),
                                           ('navier-stokes', 'rk4'),
                                           ('navier-stokes', 'rk45')])
def test_concurrent_graph_branches(backend_factory, system, scheme):
    """The fused RHS graphs hold each element type's kernels of a phase as
    parallel branches (captured on side streams).  The graphs, replayed
    over 10 steps with host reads between, give the eager one-stream run
    bit for bit."""
    from pyfr_amd.backend import Graph
    print(f'\n[concurrent graph] {system} {scheme}', flush=True)
    runs = []
    for graphs in (True, False):
        be = backend_factory(graphs=graphs)
        integ, _, _ = _hex_prism(be, 3, 2, 1, 3, system=system,
                                 scheme=scheme, dt=2e-3)
        out = []
        for _ in range(2):
            integ.run_steps(5, 2e-3)
            out.append(np.concatenate([a.ravel() for a in integ.soln]))
        if graphs:
            gs = [g for v in integ.system._kcache.values()
                  for g in (v if isinstance(v, list) else [v])
                  if isinstance(g, Graph)]
            assert gs and all(g._groups for g in gs)
        runs.append(out)
    for a, b in zip(*runs):
        assert np.array_equal(a, b)
