"""GPU parity tests: the HIP path through libfrhip.so against the oracle.

Tolerances (FP64):
* states after 100 RK4 / 20 RK45 steps: relative L2 <= 1e-12 (the stated bar);
* single operator applications: relative L2 <= 1e-13;
* one RHS evaluation: relative L2 <= max(1e-12, 10 * floor), where floor is
  the oracle's own change in RHS when its input is perturbed by 1e-16
  relative noise.  The RHS is a cancellation of O(100) fluxes into O(0.1-1)
  values, so one ulp of input moves it by ~3e-12 relative L2; a bar below
  that would test rounding order, not correctness.
"""
import numpy as np
import pytest

from oracle.basis import HexShapeOracle
from oracle.rhs import OracleHexSystem, advance
from tests.fr_cases import make_cfg, oracle_consts, rel_l2

pytestmark = pytest.mark.gpu

OPS = ['M0', 'M1 - M3*M2', 'M3', 'M4 - M6*M0', 'M6']


@pytest.mark.parametrize('order', [1, 2, 3, 4])
@pytest.mark.parametrize('op', OPS)
@pytest.mark.parametrize('beta', [0.0, 1.0])
def test_mul_structured(backend_factory, order, op, beta):
    from pyfr_amd.shapes import HexShape
    be = backend_factory()
    A = HexShape(order).opmat(op)
    Aref = HexShapeOracle(order).opmat(op)
    m, k = A.shape
    neles = 300
    rng = np.random.default_rng(order*10 + OPS.index(op))
    bnp = rng.standard_normal((k, 5, neles))
    cnp = rng.standard_normal((m, 5, neles))
    a = be.const_matrix(A, tags={op, 'align'})
    b = be.matrix((k, 5, neles), bnp, tags={'align'})
    c = be.matrix((m, 5, neles), cnp, tags={'align'})
    kern = be.kernel('mul', a, b, out=c, beta=beta)
    assert kern.kind == 'structured'
    be.run_kernels([kern], wait=True)
    ref = (Aref @ bnp.reshape(k, -1)).reshape(m, 5, neles) + beta*cnp
    assert rel_l2(c.get(), ref) < 1e-13


def test_mul_dense_fallback(backend_factory):
    be = backend_factory()
    rng = np.random.default_rng(7)
    A = rng.standard_normal((13, 9))
    bnp = rng.standard_normal((9, 5, 130))
    a = be.const_matrix(A)
    b = be.matrix((9, 5, 130), bnp, tags={'align'})
    c = be.matrix((13, 5, 130), tags={'align'})
    kern = be.kernel('mul', a, b, out=c)
    assert kern.kind == 'dense'
    be.run_kernels([kern], wait=True)
    ref = (A @ bnp.reshape(9, -1)).reshape(13, 5, 130)
    assert rel_l2(c.get(), ref) < 1e-14


def rhs_noise_floor(orc, u, seed=0):
    """Change of the oracle RHS under 1e-16 relative input noise."""
    rng = np.random.default_rng(seed)
    up = u*(1 + 1e-16*rng.standard_normal(u.shape))
    return rel_l2(orc.rhs(up), orc.rhs(u))


def _systems(be, system, n, order, rsolver='rusanov', scheme='rk4'):
    from pyfr_amd.mesh import BoxMesh
    from pyfr_amd.solvers import get_integrator
    cfg = make_cfg(system, order, rsolver, scheme=scheme)
    cfg.set('backend-hip', 'fused', str(be.fused))
    integ = get_integrator(be, BoxMesh(n), cfg)
    orc = OracleHexSystem(n, order, system, rsolver,
                          constants=oracle_consts(system))
    return integ, orc


@pytest.mark.parametrize('fused', [True, False])
@pytest.mark.parametrize('system,n,order,rsolver', [
    ('euler', 8, 2, 'rusanov'),
    ('euler', 5, 3, 'hllc'),
    ('euler', 4, 1, 'rusanov'),
    ('euler', 3, 4, 'rusanov'),
    ('navier-stokes', 4, 3, 'rusanov'),
    ('navier-stokes', 4, 2, 'hllc'),
    ('navier-stokes', 3, 1, 'rusanov'),
    ('euler', 4, 2, 'hll'),
    ('euler', 3, 3, 'hll'),
    ('navier-stokes', 3, 3, 'hll'),
    ('euler', 4, 2, 'exact'),
    ('navier-stokes', 3, 3, 'exact'),
    ('euler', 4, 2, 'roe'),
    ('euler', 4, 3, 'roem'),
    ('navier-stokes', 3, 2, 'roe'),
])
def test_rhs_parity(backend_factory, system, n, order, rsolver, fused):
    be = backend_factory(fused=fused)
    print(f'\n[rhs] {system} n={n} p={order} {rsolver} fused={fused}',
          flush=True)
    integ, orc = _systems(be, system, n, order, rsolver)
    u0 = orc.ics()
    # the product's ICs must equal the oracle's
    assert rel_l2(integ.system.ele_scal_upts(0)[0], u0) < 1e-14
    integ.system.rhs(0.0, 0, 1)
    got = integ.system.ele_scal_upts(1)[0]
    ref = orc.rhs(u0)
    tol = max(1e-12, 10*rhs_noise_floor(orc, u0))
    err = rel_l2(got, ref)
    print(f'[rhs] rel L2 {err:.3e} (tol {tol:.3e})', flush=True)
    assert err < tol


@pytest.mark.parametrize('fused', [True, False])
@pytest.mark.parametrize('beta,tau', [(-0.5, 0.0), (0.5, 0.0), (0.25, 0.1),
                                      (0.5, 0.2)])
def test_ns_ldg_parity(backend_factory, beta, tau, fused):
    """LDG beta and tau (navstokes/kernels/intconu.mako, intcflux.mako);
    at |beta| = 1/2 intconu writes one side only and the other keeps the
    element's own trace (the copy_fpts kernel)."""
    from pyfr_amd.mesh import BoxMesh
    from pyfr_amd.solvers import get_integrator
    be = backend_factory(fused=fused)
    print(f'\n[ldg] beta={beta} tau={tau} fused={fused}', flush=True)
    cfg = make_cfg('navier-stokes', 3, beta=beta, tau=tau)
    cfg.set('backend-hip', 'fused', str(be.fused))
    cfg.set('constants', 'mu', '0.02')
    integ = get_integrator(be, BoxMesh(3), cfg)
    c = oracle_consts('navier-stokes', beta, tau) | {'mu': 0.02}
    orc = OracleHexSystem(3, 3, 'navier-stokes', constants=c)
    u0 = orc.ics()
    integ.system.rhs(0.0, 0, 1)
    ref = orc.rhs(u0)
    tol = max(1e-12, 10*rhs_noise_floor(orc, u0))
    err = rel_l2(integ.system.ele_scal_upts(1)[0], ref)
    print(f'[ldg] rel L2 {err:.3e} (tol {tol:.3e})', flush=True)
    assert err < tol


@pytest.mark.parametrize('fused', [True, False])
def test_cfg0_rk4_100_steps(backend_factory, fused):
    """Plain method, CPU-runnable config: 8^3 hex p=2 Euler, Rusanov, RK4,
    dt=1e-3 to t=0.1.  The reference's fixed-dt controller takes 101 steps
    here (the last ten at dt=9e-4; see oracle.rhs.clamp_dt)."""
    be = backend_factory(fused=fused)
    integ, orc = _systems(be, 'euler', 8, 2)
    integ.advance_to(0.1)
    u_ref, nsteps = advance(orc.rhs, orc.ics(), 1e-3, 0.1, 'rk4')
    assert integ.nsteps == nsteps == 101
    err = rel_l2(integ.soln[0], u_ref)
    print(f'\n[cfg0] 100 RK4 steps rel L2 = {err:.3e}', flush=True)
    assert err < 1e-12


@pytest.mark.parametrize('fused', [True, False])
def test_ns_rk45_steps(backend_factory, fused):
    be = backend_factory(fused=fused)
    integ, orc = _systems(be, 'navier-stokes', 4, 3, scheme='rk45')
    assert integ.system.fused_stages == fused
    integ.dt = 2e-4
    integ.advance_to(20*2e-4)
    u_ref, nsteps = advance(orc.rhs, orc.ics(), 2e-4, 20*2e-4, 'rk45')
    assert integ.nsteps == nsteps
    err = rel_l2(integ.soln[0], u_ref)
    print(f'\n[ns rk45] {nsteps} steps rel L2 = {err:.3e}', flush=True)
    assert err < 1e-12


@pytest.mark.parametrize('fused', [True, False])
@pytest.mark.parametrize('scheme,system', [('euler', 'euler'),
                                           ('tvd-rk3', 'euler'),
                                           ('rk34', 'euler'),
                                           ('rk34', 'navier-stokes'),
                                           ('euler', 'navier-stokes')])
def test_other_steppers(backend_factory, scheme, system, fused):
    """Forward Euler, TVD-RK3 and RK34 (``steppers.py:11-57``, ``:193-215``)
    against the oracle's steppers over 20 steps; fused, Euler's and RK34's
    updates ride in the divergence epilogue as RK45's do."""
    be = backend_factory(fused=fused)
    integ, orc = _systems(be, system, 4, 3, scheme=scheme)
    if scheme != 'tvd-rk3':
        assert integ.system.fused_stages == fused
    dt = 2e-4 if system == 'navier-stokes' else 1e-3
    integ.dt = dt
    integ.advance_to(20*dt)
    u_ref, nsteps = advance(orc.rhs, orc.ics(), dt, 20*dt, scheme)
    assert integ.nsteps == nsteps
    assert integ.nrhsevals == nsteps*integ.nrhs_per_step
    err = rel_l2(integ.soln[0], u_ref)
    print(f'\n[{scheme}] {system} {nsteps} steps rel L2 = {err:.3e}',
          flush=True)
    assert err < 1e-12


def test_graph_replay_matches_eager(backend_factory):
    """The captured-graph path equals the eager launch sequence on every
    replay, with eager work and synchronisation between replays."""
    from pyfr_amd.mesh import BoxMesh
    from pyfr_amd.solvers import get_integrator
    outs = []
    for graphs in (False, True):
        be = backend_factory(graphs=graphs)
        integ = get_integrator(be, BoxMesh(4), make_cfg('navier-stokes', 3))
        res = []
        for rep in range(3):
            integ.system.rhs(0.0, 0, 1)
            be.wait()
            res.append(integ.system.ele_scal_upts(1)[0])
        outs.append(res)
    for a, b in zip(*outs):
        assert np.array_equal(a, b)


@pytest.mark.parametrize('system', ['euler', 'navier-stokes'])
def test_fused_rk4_advance_to_output_times(backend_factory, system):
    """RK4 with the stage updates fused (Euler and NS hexes), advanced to
    output times that are not multiples of dt, so the last steps before
    each re-split dt: equals the unfused stepping to the FP64 bar, and the
    graphs are captured once per (stage, dt), not once per step."""
    from pyfr_amd.mesh import BoxMesh
    from pyfr_amd.solvers import get_integrator
    sols, caps = [], []
    for fused in (True, False):
        be = backend_factory(fused=fused)
        integ = get_integrator(be, BoxMesh(4), make_cfg(system, 3, dt=1e-3))
        assert integ.system.fused_rk4 == fused
        c0 = be.ncaptures
        for t in (0.0105, 0.0213, 0.03):
            integ.advance_to(t)
        sols.append(integ.soln[0])
        caps.append(be.ncaptures - c0)
    err = rel_l2(sols[0], sols[1])
    print(f'\n[rk4 fused] {system} advance_to rel L2 {err:.2e}, captures '
          f'{caps[0]} for {integ.nsteps} steps', flush=True)
    assert err < 1e-12
    # the epilogues read dt from device memory: the stage graphs are
    # captured once per stage and register parity (8), although the last
    # steps before each output time take 16 distinct values of dt
    assert caps[0] <= 8
