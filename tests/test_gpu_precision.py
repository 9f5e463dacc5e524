"""The precision form on the GPU: FP32 right-hand side with FP64 RK state
(RK45 and RK4), the base-state form it uses on periodic hex meshes, and its
FP64 reductions.

Bars:
* base-state form in FP64 against the oracle's RHS: relative L2
  <= max(1e-12, 10 x the oracle's 1-ulp floor), the bar of the FP64 path
  (the base state's flux has zero discrete divergence, so the RHS is the
  same up to rounding);
* precision form against the FP64 run of the same steps, per field:
  relative L2 <= 1e-5 for rho and E, and for each momentum component
  relative to the L2 norm of the momentum vector (the w-momentum of the
  Taylor-Green field starts at zero, so its own norm is no scale);
* fused (epilogue) and unfused (separate stage kernel) forms of the same
  scheme: relative L2 <= 1e-6 of each other (the same FP64 updates; the
  FP32 divergence kernels round differently).
"""
import numpy as np
import pytest

from oracle.rhs import OracleHexSystem
from tests.fr_cases import make_cfg, oracle_consts, rel_l2

pytestmark = pytest.mark.gpu


def _integ(backend_factory, system, order, n, precision, scheme, dt,
           fused=True, base='auto'):
    from pyfr_amd.mesh import BoxMesh
    from pyfr_amd.solvers import get_integrator
    cfg = make_cfg(system, order, scheme=scheme, dt=dt, precision=precision)
    cfg.set('backend-hip', 'fused', str(fused))
    cfg.set('backend-hip', 'base-state', base)
    be = backend_factory(precision=precision, fused=fused)
    return get_integrator(be, BoxMesh(n), cfg)


def field_errors(a, b):
    """Per-field relative L2 errors of a against b ((nupts, 5, neles)):
    rho, E own norms; momenta against the momentum vector's norm."""
    mom = np.linalg.norm(b[:, 1:4])
    out = {'rho': rel_l2(a[:, 0], b[:, 0]), 'E': rel_l2(a[:, 4], b[:, 4])}
    for i, nm in enumerate(('rhou', 'rhov', 'rhow')):
        out[nm] = float(np.linalg.norm(a[:, 1 + i] - b[:, 1 + i])/mom)
    return out


@pytest.mark.parametrize('system', ['euler', 'navier-stokes'])
@pytest.mark.parametrize('fused', [True, False])
def test_base_state_rhs_fp64(backend_factory, system, fused):
    """FP64 with base-state 'on': the RHS equals the oracle's."""
    integ = _integ(backend_factory, system, 3, 4, 'double', 'rk4', 1e-3,
                   fused=fused, base='on')
    s = integ.system
    assert s.base is not None and s.base[1] == 0 and s.base[4] > 100
    orc = OracleHexSystem(4, 3, system, constants=oracle_consts(system))
    u0 = orc.ics()
    # the register holds u - u0; the readout adds u0 back
    assert np.max(np.abs(s.soln_upts(0)[0] - u0)) < 1e-12
    s.rhs(0.0, 0, 1)
    ref = orc.rhs(u0)
    rng = np.random.default_rng(0)
    floor = rel_l2(orc.rhs(u0*(1 + 1e-16*rng.standard_normal(u0.shape))),
                   ref)
    err = rel_l2(s.ele_scal_upts(1)[0], ref)
    print(f'\n[base-state fp64] {system} fused={fused} RHS rel L2 '
          f'{err:.2e} (floor {floor:.2e})', flush=True)
    assert err < max(1e-12, 10*floor)


def test_base_state_refusals(backend_factory):
    from pyfr_amd.mesh import BoxMesh
    from pyfr_amd.solvers import get_integrator
    cfg = make_cfg('euler', 2, rsolver='hllc', precision='mixed')
    cfg.set('backend-hip', 'base-state', 'on')
    be = backend_factory(precision='mixed')
    with pytest.raises(ValueError, match='Rusanov'):
        get_integrator(be, BoxMesh(4), cfg)
    # auto: the full-state precision form where the base state does not
    # apply
    cfg.set('backend-hip', 'base-state', 'auto')
    assert get_integrator(be, BoxMesh(4), cfg).system.base is None


@pytest.mark.parametrize('scheme,nsteps', [('rk45', 20), ('rk4', 20),
                                           ('rk34', 20)])
def test_precision_form_per_field(backend_factory, scheme, nsteps):
    """NS hex p=3, 4^3, dt = 2e-4: the precision form (fused and unfused)
    against FP64, per field; the full-state form (base-state off) for
    comparison; FP64 conservation sums of the FP64 state."""
    from pyfr_amd.solvers import MIXED_INTEGRATORS as mixed_cls
    sols = {}
    for prec, fused, base in (('double', True, 'auto'),
                              ('mixed', True, 'auto'),
                              ('mixed', False, 'auto'),
                              ('mixed', True, 'off')):
        integ = _integ(backend_factory, 'navier-stokes', 3, 4, prec, scheme,
                       2e-4, fused=fused, base=base)
        assert isinstance(integ, mixed_cls[scheme]) == (prec == 'mixed')
        if prec == 'mixed':
            assert (integ.system.base is not None) == (base == 'auto')
        if prec == 'mixed' and scheme == 'rk4':
            assert integ.system.fused_rk4_mixed == fused
        integ.run_steps(nsteps)
        sols[prec, fused, base] = sol = integ.soln[0]
        if prec == 'mixed':
            cs = integ.system.conservation_sums(integ.idxcurr)
            w = integ.system.ele_map['hex'].basis.upts_wts
            rj = integ.system.ele_map['hex'].rcpdjac_at_np('upts')
            ref = np.einsum('u,uve->v', w, sol/rj[:, None, :])
            assert np.allclose(cs, ref, rtol=1e-12, atol=1e-9)
            l2 = integ.system.l2_norms(integ.idxcurr)
            refl2 = np.sqrt(np.einsum('u,uve->v', w, sol**2/rj[:, None, :]))
            assert np.allclose(l2, refl2, rtol=1e-12)
    b = sols['double', True, 'auto']
    a = sols['mixed', True, 'auto']
    assert rel_l2(a, sols['mixed', False, 'auto']) < 1e-6
    errs = field_errors(a, b)
    full = field_errors(sols['mixed', True, 'off'], b)
    for k in errs:
        print(f'\n[precision {scheme}] {k}: base-state {errs[k]:.2e}, '
              f'full-state {full[k]:.2e}', flush=True)
    assert max(errs.values()) < 1e-5, errs


def test_precision_form_euler_rk4(backend_factory):
    """cfg0's scheme in the precision form: Euler hex p=2, 8^3, 100 RK4
    steps of dt = 1e-3 against FP64, per field."""
    sols = []
    for prec in ('double', 'mixed'):
        integ = _integ(backend_factory, 'euler', 2, 8, prec, 'rk4', 1e-3)
        integ.advance_to(0.1)
        assert integ.nsteps == 101
        sols.append(integ.soln[0])
    errs = field_errors(sols[1], sols[0])
    print(f'\n[precision euler rk4] {errs}', flush=True)
    assert max(errs.values()) < 1e-5, errs
