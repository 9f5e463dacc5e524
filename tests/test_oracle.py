"""Pinning the oracle (CPU).

What pins each oracle function:
* operators M0..M3 (oracle.basis)   -- the reference's golden fixture
  hex-gleg-ord3.npz (its only test data for the path), plus polynomial
  exactness and the quadrature identity of the correction matrix at p=1..5;
* GL points                         -- values parsed from the reference's rule
  tables (tests/golden/gl_rules.npz);
* clean()                           -- a crafted array with known output;
* point physics (oracle.physics)    -- closed forms (flux of a state, simple
  shear stress), Riemann consistency and conservation, HLLC upwinding;
* the RHS (oracle.rhs)              -- free-stream preservation, discrete
  conservation and the convergence order of an exact travelling density
  wave (catches sign, index, metric and correction-scaling mistakes);
* RK steppers                       -- order of accuracy on y' = lambda*y;
* controller                        -- step count / final time for cfg0;
* pairing (oracle.mesh)             -- every face used exactly once and
  paired flux points coincide, on 1^3..4^3 boxes.
"""
import os

import numpy as np
import pytest

from oracle.basis import (HexShapeOracle, clean, gauss_legendre_hex,
                          gauss_legendre_line, gauss_legendre_quad)
from oracle.geom import trilinear_ploc
from oracle.mesh import box_hex_verts, pair_faces_bruteforce
from oracle.physics import hllc, inviscid_flux, rusanov, viscous_flux_add
from oracle.rhs import (RK45_A, RK45_B, OracleHexSystem, advance, clamp_dt,
                        euler_step, rk4_step, rk34_step, rk45_step,
                        rkvdh2_step, tvdrk3_step)

GOLD = os.path.join(os.path.dirname(__file__), 'golden')


def test_fixture_hex_gleg_ord3():
    ref = np.load(os.path.join(GOLD, 'hex_gleg_ord3.npz'))
    h = HexShapeOracle(3)
    for k in ('m0', 'm1', 'm2', 'm3'):
        a, b = getattr(h, k), ref[k]
        assert a.shape == b.shape
        assert np.abs(a - b).max() < 1e-13*max(1, np.abs(b).max())


def test_gl_rules_against_reference_tables():
    g = np.load(os.path.join(GOLD, 'gl_rules.npz'))
    for n in range(1, 7):
        x, w = gauss_legendre_line(n)
        assert np.abs(x - g[f'line_n{n}_pts'][:, 0]).max() <= 4e-16
        assert np.abs(w - g[f'line_n{n}_wts']).max() <= 1e-15
        assert np.abs(gauss_legendre_quad(n)[0] -
                      g[f'quad_n{n}_pts']).max() <= 4e-16
        assert np.abs(gauss_legendre_hex(n)[0] -
                      g[f'hex_n{n}_pts']).max() <= 4e-16


def test_clean():
    a = np.array([1e-11, -0.5, 0.5 + 1e-12, 0.5 - 2e-12, 0.7, -1e-12])
    c = clean(a)
    assert c[0] == 0 and c[5] == 0
    assert c[1] == -c[2] == -c[3] == -np.median([0.5, 0.5 + 1e-12,
                                                  0.5 - 2e-12])
    assert c[4] == 0.7


def _poly(pts, ex):
    x, y, z = pts.T
    return x**ex[0]*y**ex[1]*z**ex[2]


@pytest.mark.parametrize('order', [1, 2, 3, 4, 5])
def test_operator_exactness(order):
    h = HexShapeOracle(order)
    rng = np.random.default_rng(order)
    for _ in range(6):
        ex = rng.integers(0, order + 1, 3)
        u = _poly(h.upts, ex)
        # M0 interpolates degree-p polynomials exactly
        assert np.allclose(h.m0 @ u, _poly(h.fpts, ex), atol=1e-12)
        # M1 (through its re-layout M4) differentiates them exactly
        g = h.m4 @ u
        for d in range(3):
            exd = ex.copy()
            c = exd[d]
            exd[d] = max(c - 1, 0)
            ref = c*_poly(h.upts, exd) if c > 0 else 0*u
            assert np.allclose(g[d*h.nupts:(d + 1)*h.nupts], ref, atol=1e-11)
    # Integral of each correction function equals its face weight:
    # sum_u w_u M3[u, f] = w_f  (g_f lifts the face delta; shapes.py:231-259)
    _, fw = gauss_legendre_quad(order + 1)
    fwts = np.tile(fw, 6)
    assert np.allclose(h.upts_wts @ h.m3, fwts, atol=1e-13)


def test_inviscid_flux_closed_form():
    gamma = 1.4
    rho, u, v, w, p = 1.3, 0.4, -0.2, 0.7, 2.1
    E = p/(gamma - 1) + 0.5*rho*(u*u + v*v + w*w)
    s = np.array([rho, rho*u, rho*v, rho*w, E])
    f, pp, vv = inviscid_flux(s, gamma)
    vel = np.array([u, v, w])
    assert np.isclose(pp, p)
    for i in range(3):
        assert np.isclose(f[i, 0], rho*vel[i])
        for j in range(3):
            assert np.isclose(f[i, j + 1], rho*vel[i]*vel[j] + p*(i == j))
        assert np.isclose(f[i, 4], (E + p)*vel[i])


def _rand_states(rng, n):
    rho = rng.uniform(0.5, 2, n)
    vel = rng.uniform(-1, 1, (3, n))
    p = rng.uniform(0.5, 3, n)
    E = p/0.4 + 0.5*rho*(vel**2).sum(0)
    return np.vstack([rho, rho*vel, E])


@pytest.mark.parametrize('rs', [rusanov, hllc])
def test_riemann_consistency_and_conservation(rs):
    rng = np.random.default_rng(3)
    ul, ur = _rand_states(rng, 50), _rand_states(rng, 50)
    n = rng.standard_normal((3, 50))
    n /= np.linalg.norm(n, axis=0)
    f, _, _ = inviscid_flux(ul, 1.4)
    fn = np.einsum('dvn,dn->vn', f, n)
    assert np.allclose(rs(ul, ul, n, 1.4), fn, rtol=1e-13, atol=1e-13)
    assert np.allclose(rs(ul, ur, n, 1.4), -rs(ur, ul, -n, 1.4),
                       rtol=1e-12, atol=1e-12)


def test_hllc_supersonic_upwind():
    # Both states moving at Mach ~3 along +n: the flux is the left flux
    ul = np.array([[1.0], [3.5], [0.0], [0.0], [1/0.4 + 0.5*3.5**2]])
    ur = np.array([[1.1], [3.8], [0.1], [0.0], [1.1/0.4 + 0.5*(3.8**2 + 0.01)/1.1]])
    n = np.array([[1.0], [0.0], [0.0]])
    f, _, _ = inviscid_flux(ul, 1.4)
    assert np.allclose(hllc(ul, ur, n, 1.4), f[0], rtol=1e-14)


def test_viscous_flux_simple_shear():
    mu, s = 0.01, 0.3
    c = {'gamma': 1.4, 'mu': mu, 'Pr': 0.71}
    u = np.array([1.0, 0.2, 0.0, 0.0, 2.5])
    g = np.zeros((3, 5))
    g[1, 1] = s                     # d(rho u)/dy with rho = 1: du/dy = s
    f = viscous_flux_add(u, g, np.zeros((3, 5)), c)
    # negated stress: f_y(x-mom) = f_x(y-mom) = -mu du/dy.  Energy: work
    # u*t_xy plus heat flux -mu*gamma/Pr*dT/dy where, with E and rho fixed,
    # c_v dT/dy = -u du/dy.
    assert np.isclose(f[1, 1], -mu*s) and np.isclose(f[0, 2], -mu*s)
    assert np.isclose(f[1, 4], 0.2*(-mu*s) + mu*(1.4/0.71)*(0.2*s))
    assert np.allclose(f[:, 0], 0)


UNIF = (1.0, 0.3, -0.2, 0.1, 2.5/0.4 + 0.5*(0.09 + 0.04 + 0.01))


@pytest.mark.parametrize('system,order', [('euler', 1), ('euler', 2),
                                          ('euler', 3),
                                          ('navier-stokes', 2)])
def test_freestream_and_conservation(system, order):
    c = {'gamma': 1.4} if system == 'euler' else \
        {'gamma': 1.4, 'mu': 0.01, 'Pr': 0.71, 'ldg-beta': 0.0,
         'ldg-tau': 0.0}
    warp = lambda x, y, z: (x + 0.1*np.sin(y)*np.sin(z),
                            y + 0.1*np.sin(x)*np.sin(z), z)
    verts = box_hex_verts(3, warp=warp)
    o = OracleHexSystem(3, order, system, constants=c, verts=verts)
    u = np.empty((o.nupts, 5, o.neles))
    u[:] = np.array(UNIF)[None, :, None]
    assert np.abs(o.rhs(u)).max() < 1e-11
    r = o.rhs(o.ics())
    w = o.shape.upts_wts
    tot = [(w[:, None]*r[:, v]/o.rcpdjac).sum() for v in range(5)]
    assert np.abs(tot).max() < 1e-10


@pytest.mark.parametrize('order', [1, 2, 3])
def test_rhs_convergence_density_wave(order):
    """Euler exact solution rho = 1 + 0.2 sin(x + y + z), velocity (U, V, W)
    and pressure uniform: d(rho)/dt = -(U + V + W) rho'(.)"""
    U = np.array([0.5, 0.3, -0.2])
    errs = []
    for n in (4, 8):
        o = OracleHexSystem(n, order)
        x, y, z = trilinear_ploc(o.verts, o.shape.upts).transpose(1, 0, 2)
        rho = 1 + 0.2*np.sin(x + y + z)
        p = 1.0
        u = np.stack([rho, rho*U[0], rho*U[1], rho*U[2],
                      p/0.4 + 0.5*rho*(U**2).sum()], axis=1)
        r = o.rhs(u)
        exact = -U.sum()*0.2*np.cos(x + y + z)
        errs.append(np.sqrt(np.mean((r[:, 0] - exact)**2)))
    # The RHS is a derivative of a degree-p interpolant: error O(h^p)
    # (measured 1.4 / 2.0 / 3.6 for p = 1 / 2 / 3).
    rate = np.log2(errs[0]/errs[1])
    assert rate > order - 0.7, (errs, rate)


def test_rk_orders():
    lam = -1.0 + 0.5j

    def f(y):
        return lam*y
    for step, order in ((rk4_step, 4), (rk45_step, 4), (rk34_step, 3),
                        (tvdrk3_step, 3), (euler_step, 1)):
        errs = []
        for dt in (0.1, 0.05):
            y = np.array([1.0 + 0j])
            for _ in range(int(round(1/dt))):
                y = step(f, y, dt)
            errs.append(abs(y[0] - np.exp(lam)))
        assert np.log2(errs[0]/errs[1]) > order - 0.3


def test_rk_one_step_closed_forms():
    """On y' = lam y one step multiplies y by the scheme's stability
    polynomial: 1 + z (Euler); the cubic Taylor polynomial (TVD-RK3, a
    3-stage third-order scheme); for the low-storage schemes the Taylor
    terms up to their order, z = lam dt.  A dropped or misplaced stage
    term changes a coefficient."""
    from math import factorial
    lam = -0.7 + 0.4j
    for step, order, exact in ((euler_step, 1, True), (tvdrk3_step, 3, True),
                               (rk34_step, 3, False), (rk45_step, 4, False),
                               (rk4_step, 4, True)):
        # fit the polynomial from several z (degree <= stages)
        zs = np.array([0.05, 0.1, 0.15, 0.2, 0.25, 0.3])*lam
        vals = np.array([step(lambda y, l=z: l*y, np.array([1.0 + 0j]),
                              1.0)[0] for z in zs])
        V = np.vander(zs, 6, increasing=True)
        coef = np.linalg.solve(V, vals)
        for k in range(order + 1):
            assert abs(coef[k] - 1/factorial(k)) < 1e-9, (step, k, coef)
        if exact:
            assert np.allclose(coef[order + 1:], 0, atol=1e-9)
    # the shared low-storage step with RK45's table is rk45_step
    y = np.array([1.0, -2.0])
    f = (lambda v: np.array([-v[1], v[0]]))
    assert np.array_equal(rkvdh2_step(f, y, 0.1, RK45_A, RK45_B),
                          rk45_step(f, y, 0.1))


def test_controller_cfg0_steps():
    # The reference's clamp (-(-r // dt) and a 10-step look-ahead) takes
    # 101 steps to reach t=0.1 with dt=1e-3; the last ten are 9e-4 each.
    t, comp, n, dts = 0.0, 0.0, 0, []
    while t < 0.1:
        h = clamp_dt(1e-3, 0.1, t)
        y = h - comp
        tt = t + y
        comp = (tt - t) - y
        t = tt
        n += 1
        dts.append(h)
    assert n == 101 and t == 0.1
    assert np.allclose(dts[-10:], 9e-4, rtol=1e-12)
    _, ns = advance(lambda u: 0*u, np.zeros(1), 1e-3, 0.1)
    assert ns == 101


@pytest.mark.parametrize('n', [1, 2, 3, 4])
def test_pairing_bruteforce(n):
    h = HexShapeOracle(2)
    v = box_hex_verts(n)
    pl = trilinear_ploc(v, h.fpts)
    pairs = pair_faces_bruteforce(pl, 6, h.nfacefpts, 2*np.pi)
    assert len(pairs) == 3*n**3
    used = [(p[0], p[1]) for p in pairs] + [(p[3], p[4]) for p in pairs]
    assert len(set(used)) == 6*n**3


@pytest.mark.parametrize('system,n,order', [('euler', 4, 2),
                                            ('navier-stokes', 4, 3),
                                            ('navier-stokes', 3, 1)])
def test_c_restatement_matches_numpy(system, n, order):
    """The oracle's C/OpenMP RHS (the CPU baseline) agrees with the NumPy
    oracle to the RHS rounding floor."""
    import subprocess
    from oracle.crhs import LIB, CRhs
    from tests.fr_cases import oracle_consts
    if not os.path.exists(LIB):
        subprocess.run(['make', '-s', '-C', os.path.dirname(LIB) + '/..'],
                       check=True)
    orc = OracleHexSystem(n, order, system, constants=oracle_consts(system))
    u = orc.ics()
    ref = orc.rhs(u)
    got = CRhs(orc).rhs(u)
    assert np.linalg.norm(got - ref)/np.linalg.norm(ref) < 2e-12


def _random_states(rng, n, mach=0.5):
    rho = rng.uniform(0.5, 2.0, n)
    p = rng.uniform(0.5, 2.0, n)
    v = rng.normal(0, mach, (3, n))
    return np.stack([rho, *(rho*v), p/0.4 + 0.5*rho*(v**2).sum(0)])


@pytest.mark.parametrize('name', ['rusanov', 'hllc', 'hll', 'roe', 'roem',
                                  'exact'])
def test_riemann_solver_properties(name):
    """Consistency F(u, u, n) = f(u).n for every solver; conservation
    F(ul, ur, n) = -F(ur, ul, -n) where the scheme has it (the reference's
    HLL estimates the left wave speed with the right sound speed, so it is
    symmetric only for states of equal sound speed); exact upwinding for
    supersonic flow along n."""
    from oracle.physics import RSOLVERS, inviscid_flux
    rs = RSOLVERS[name]
    rng = np.random.default_rng(11)
    N = 300
    ul, ur = _random_states(rng, N), _random_states(rng, N)
    n = rng.normal(size=(3, N))
    n /= np.linalg.norm(n, axis=0)
    f, _, _ = inviscid_flux(ul, 1.4)
    fn = (n[:, None, :]*f).sum(0)
    assert np.abs(rs(ul, ul, n, 1.4) - fn).max() < 1e-12
    if name == 'hll':
        # equal density and pressure on both sides: al == ar
        ur2 = ul.copy()
        v2 = rng.normal(0, 0.5, (3, N))
        ur2[1:4] = ul[0]*v2
        pl = 0.4*(ul[4] - 0.5*(ul[1:4]**2).sum(0)/ul[0])
        ur2[4] = pl/0.4 + 0.5*ul[0]*(v2**2).sum(0)
        assert np.abs(rs(ul, ur2, n, 1.4) + rs(ur2, ul, -n, 1.4)).max() < 1e-12
    else:
        assert np.abs(rs(ul, ur, n, 1.4) + rs(ur, ul, -n, 1.4)).max() < 1e-12
    # supersonic along n: both states moving at Mach > 3 along n
    sl, sr = _random_states(rng, N, 0.05), _random_states(rng, N, 0.05)
    for s in (sl, sr):
        s[4] -= 0.5*(s[1:4]**2).sum(0)/s[0]
        s[1:4] += 4.0*s[0]*n
        s[4] += 0.5*(s[1:4]**2).sum(0)/s[0]
    f, _, _ = inviscid_flux(sl, 1.4)
    fn = (n[:, None, :]*f).sum(0)
    fr, _, _ = inviscid_flux(sr, 1.4)
    fnr = (n[:, None, :]*fr).sum(0)
    got = rs(sl, sr, n, 1.4)
    if name == 'roe':
        # the reference's Roe flux is 1/2 (fl + fr) - sum |l| alpha r, with
        # no 1/2 on the dissipation; with every eigenvalue positive and the
        # Roe property sum l alpha r = fr - fl it reduces to (3 fl - fr)/2
        assert np.abs(got - (1.5*fn - 0.5*fnr)).max() < \
            1e-10*np.abs(fn).max()
    elif name != 'rusanov':
        assert np.abs(got - fn).max() < 1e-10*np.abs(fn).max()


@pytest.mark.parametrize('rho_l,u_l,p_l,rho_r,u_r,p_r,w', [
    # Toro (2009) Table 4.3, tests 1 (Sod), 3 and 5; w = (rho, u, p) of the
    # exact solution at x/t = 0: the left star state (rho*L, u*, p*) for
    # tests 1 and 3; for test 5 the left shock moves right (S_L = +0.79),
    # so x/t = 0 sees the unperturbed left state
    (1.0, 0.0, 1.0, 0.125, 0.0, 0.1, (0.42632, 0.92745, 0.30313)),
    (1.0, 0.0, 1000.0, 1.0, 0.0, 0.01, (0.57506, 19.5975, 460.894)),
    (5.99924, 19.5975, 460.894, 5.99242, -6.19633, 46.0950,
     (5.99924, 19.5975, 460.894)),
])
def test_exact_solver_star_states(rho_l, u_l, p_l, rho_r, u_r, p_r, w):
    """Known answers: the exact solver's interface flux equals the flux of
    the state Toro's exact solution has at x/t = 0 (to the table's 5-6
    digits); the face may have any orientation, and a tangential velocity
    is carried through."""
    from oracle.physics import exact
    g = 1.4
    rs, us, ps = w
    want = np.array([rs*us, rs*us*us + ps, 0, 0,
                     (g/(g - 1)*ps + 0.5*rs*us*us)*us])
    rng = np.random.default_rng(5)
    n = rng.normal(size=3)
    n /= np.linalg.norm(n)
    t = np.cross(n, rng.normal(size=3))
    t /= np.linalg.norm(t)
    vt = 0.3*t

    def con(r, u, p):
        v = u*n + vt
        return np.array([r, *(r*v), p/(g - 1) + 0.5*r*(v @ v)])[:, None]
    f = exact(con(rho_l, u_l, p_l), con(rho_r, u_r, p_r), n[:, None], g)[:, 0]
    # back to the face frame: normal and tangential momentum flux
    got = np.array([f[0], f[1:4] @ n, 0, 0, f[4] - 0.5*(vt @ vt)*f[0]])
    assert np.allclose((f[1:4] - (f[1:4] @ n)*n), f[0]*vt, rtol=1e-10,
                       atol=1e-10*abs(f[0]))
    assert np.allclose(got, want, rtol=2e-5, atol=2e-5*np.abs(want).max())


def _contact_states(rng, N, U):
    """Left/right states of an isolated contact moving at normal speed U:
    equal pressure and normal velocity, different densities and tangential
    velocities.  Returns (ul, ur, n, p)."""
    n = rng.normal(size=(3, N))
    n /= np.linalg.norm(n, axis=0)
    p = rng.uniform(0.5, 2.0, N)
    out = []
    for _ in range(2):
        rho = rng.uniform(0.3, 3.0, N)
        t = rng.normal(0, 0.5, (3, N))
        t -= (t*n).sum(0)*n                  # tangential part only
        v = U*n + t
        out.append(np.stack([rho, *(rho*v),
                             p/0.4 + 0.5*rho*(v**2).sum(0)]))
    return out[0], out[1], n, p


@pytest.mark.parametrize('name', ['hllc', 'roe', 'roem', 'exact'])
def test_stationary_contact_is_exact(name):
    """A stationary contact (u = 0, equal p, any density jump) is a steady
    solution of the Euler equations: the exact interface flux is
    (0, p n, 0).  HLLC, Roe and exact resolve it exactly; RoeM to its
    1e-15 Mach offset."""
    from oracle.physics import RSOLVERS
    rng = np.random.default_rng(21)
    ul, ur, n, p = _contact_states(rng, 200, 0.0)
    ul[1:4] = ur[1:4] = 0.0
    ul[4] = ur[4] = p/0.4
    got = RSOLVERS[name](ul, ur, n, 1.4)
    ref = np.concatenate([np.zeros((1, 200)), p*n, np.zeros((1, 200))])
    assert np.abs(got - ref).max() < 1e-13


def test_rusanov_stationary_contact_closed_form():
    """Rusanov on a stationary contact: momentum p n, energy 0 and mass
    a (rho_l - rho_r) with the reference's wave speed
    a = sqrt(gamma (p_l + p_r) / (4 (rho_l + rho_r))) (no velocity term)."""
    from oracle.physics import rusanov
    rng = np.random.default_rng(22)
    ul, ur, n, p = _contact_states(rng, 200, 0.0)
    ul[1:4] = ur[1:4] = 0.0
    ul[4] = ur[4] = p/0.4
    got = rusanov(ul, ur, n, 1.4)
    a = np.sqrt(1.4*2*p/(4*(ul[0] + ur[0])))
    ref = np.concatenate([(a*(ul[0] - ur[0]))[None], p*n,
                          np.zeros((1, 200))])
    assert np.abs(got - ref).max() < 1e-13


@pytest.mark.parametrize('U', [0.3, -0.3, 0.7, -0.05])
@pytest.mark.parametrize('name', ['hllc', 'exact'])
def test_moving_contact_is_upwinded(name, U):
    """An isolated subsonic contact (with a shear jump) moving at normal
    speed U: the exact solution at x/t = 0 is the left state for U > 0 and
    the right state for U < 0, so the flux is f(ul).n or f(ur).n.  HLLC's
    star states reproduce it exactly (s* = U)."""
    from oracle.physics import RSOLVERS, inviscid_flux
    rng = np.random.default_rng(23)
    ul, ur, n, _ = _contact_states(rng, 200, U)
    got = RSOLVERS[name](ul, ur, n, 1.4)
    f, _, _ = inviscid_flux(ul if U > 0 else ur, 1.4)
    ref = (n[:, None, :]*f).sum(0)
    assert np.abs(got - ref).max() < 1e-12*(1 + np.abs(ref).max())


@pytest.mark.parametrize('old,new', [
    # plausible mistakes in the star-state formulas
    ('ul[0]*nvl*(sl - nvl)', 'ul[0]*nvl*(sl + nvl)'),
    ('(sl - nvl)/(sl - sstar)', '(sl - nvl)/(sl + sstar)'),
    ('pl/(sl - nvl)', 'pr/(sl - nvl)'),
    ('usl[0]*(vl[i] + (sstar - nvl)*n[i])', 'usl[0]*(vl[i] + sstar*n[i])'),
])
def test_planted_hllc_mistake_is_caught(old, new):
    """Each mistake planted in a copy of the oracle's HLLC fails one of the
    pins: the contact checks above, consistency F(u, u, n) = f(u).n or
    conservation F(ul, ur, n) = -F(ur, ul, -n) on random states."""
    import inspect

    import oracle.physics as ph
    src = inspect.getsource(ph.hllc)
    assert old in src
    ns = dict(vars(ph))
    exec(src.replace(old, new), ns)
    bad = ns['hllc']
    rng = np.random.default_rng(24)
    caught = False
    for U in (0.0, 0.3, -0.3):
        ul, ur, n, p = _contact_states(rng, 200, U)
        f, _, _ = ph.inviscid_flux(ul if U >= 0 else ur, 1.4)
        ref = (n[:, None, :]*f).sum(0)
        if U == 0.0:
            ul[1:4] = ur[1:4] = 0.0
            ul[4] = ur[4] = p/0.4
            ref = np.concatenate([np.zeros((1, 200)), p*n,
                                  np.zeros((1, 200))])
        caught |= np.abs(bad(ul, ur, n, 1.4) - ref).max() > 1e-8
    ul, ur = _random_states(rng, 300), _random_states(rng, 300)
    n = rng.normal(size=(3, 300))
    n /= np.linalg.norm(n, axis=0)
    f, _, _ = ph.inviscid_flux(ul, 1.4)
    caught |= np.abs(bad(ul, ul, n, 1.4) -
                     (n[:, None, :]*f).sum(0)).max() > 1e-8
    caught |= np.abs(bad(ul, ur, n, 1.4) +
                     bad(ur, ul, -n, 1.4)).max() > 1e-8
    assert caught
