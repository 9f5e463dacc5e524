"""Property checks at the benchmark sizes, where the oracle cannot follow:
the 64^3 hex NS p=3 FP64 workload, the 48^3 x 6 tet + 16^3 x 2 prism
Euler p=4 FP32 workload and the 32^3 x 6 pyramid Euler p=4 FP32 box.  Large-grid indexing, ordering or race faults
(an element block or interface pair computed twice or skipped) break them.

* Discrete conservation on the periodic box: the quadrature integrals of
  mass, momentum and energy (FP64 sums) are invariant under the RHS.  FP64:
  relative change <= 1e-13 of the integral of |u| after 5 RK45 steps; FP32
  (states rounded to FP32 every stage): <= 2e-6.  (Pyramids conserve to
  truncation error from p = 2, which at 32^3 cells and these smooth modes
  is far below the FP32 bar.)
* Free-stream preservation: a uniform moving state has a zero RHS up to
  rounding.  The metric terms are differences of absolute vertex
  coordinates, so the residual grows with the box's extent over the element
  size: the oracle gives max |RHS| = 2.8e-13, 3.4e-12 and 3.0e-11 on 4^3,
  8^3 and 16^3 boxes (p = 3, this state), about n^3, which extrapolates to
  ~3e-9 at 64^3.  Bars: max |RHS| <= 1e-9 (FP64) and 1e-3 (FP32) of the
  largest flux component times the largest 1/|J|^(1/3) (~2e-8 and ~2e-2
  here); a wrong index or metric term gives O(1e-2) or more.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _uniform(system, rho=1.2, vel=(0.3, -0.2, 0.1), p=0.9):
    v = np.array(vel)
    return np.array([rho, *(rho*v), p/0.4 + 0.5*rho*v @ v])


def _flux_scale(u):
    rho, m, E = u[0], u[1:4], u[4]
    p = 0.4*(E - 0.5*m @ m/rho)
    return max(np.abs(m).max(), np.abs(np.outer(m, m)/rho).max() + p,
               np.abs((E + p)*m/rho).max())


def _integ(c, precision):
    import bench
    from pyfr_amd.backend import AMDBackend
    from pyfr_amd.solvers import get_integrator
    cfg = bench.make_cfg(c, precision)
    be = AMDBackend(cfg)
    return be, get_integrator(be, bench.make_mesh(c), cfg)


def _check_conservation(integ, nsteps, dt, tol):
    s = integ.system
    before = s.conservation_sums(integ.idxcurr)
    mags = np.sum(s._reduce(integ.idxcurr, 'sumsq'), axis=0)
    integ.run_steps(nsteps, dt)
    after = s.conservation_sums(integ.idxcurr)
    # scale of each integral: the integral of |u| is bounded by
    # sqrt(volume * integral of u^2); the momenta share the momentum
    # vector's (a component may vanish, as w does at first here)
    vol = (2*np.pi)**3*len(s.ele_types)
    mags[1:4] = mags[1:4].sum()
    scale = np.sqrt(vol*mags)
    rel = np.abs(after - before)/scale
    print(f'\n[conservation] before {before}\n  after {after}\n  rel {rel}',
          flush=True)
    assert np.all(rel < tol), rel


def _check_uniform(integ, tol):
    s = integ.system
    u0 = _uniform(s.name)
    s.set_soln_upts(0, [np.broadcast_to(u0[None, :, None], sh).copy()
                        for sh in s.ele_shapes.values()])
    wk = len(s.ele_banks[0]) - 1
    s.rhs(0.0, 0, wk)
    rmax = max(np.abs(a).max() for a in s.ele_scal_upts(wk))
    rcp = max(e.rcpdjac_at_np('upts').max()**(1/3)
              for e in s.ele_map.values())
    bound = tol*_flux_scale(u0)*rcp
    print(f'\n[free stream] max |RHS| {rmax:.3e} (bound {bound:.3e})',
          flush=True)
    assert rmax < bound


def test_ns_hex64_fp64_conservation():
    import bench
    c = bench.CONFIGS['cfg1']
    be, integ = _integ(c, 'double')
    assert integ.system.gndofs == 83886080
    _check_conservation(integ, 5, c['dt'], 1e-13)


def test_ns_hex64_fp64_free_stream():
    import bench
    be, integ = _integ(bench.CONFIGS['cfg1'], 'double')
    _check_uniform(integ, 1e-9)


def test_tet_pri_p4_fp32_conservation():
    import bench
    c = bench.CONFIGS['cfg3']
    be, integ = _integ(c, 'single')
    assert integ.system.gndofs == 119193600
    _check_conservation(integ, 2, c['dt'], 2e-6)


def test_tet_pri_p4_fp32_free_stream():
    import bench
    be, integ = _integ(bench.CONFIGS['cfg3'], 'single')
    _check_uniform(integ, 1e-3)


def test_pyr_p4_fp32_conservation():
    import bench
    c = bench.CONFIGS['pyr']
    be, integ = _integ(c, 'single')
    assert integ.system.gndofs == 54067200
    _check_conservation(integ, 2, c['dt'], 2e-6)


def test_pyr_p4_fp32_free_stream():
    import bench
    be, integ = _integ(bench.CONFIGS['pyr'], 'single')
    _check_uniform(integ, 1e-3)
