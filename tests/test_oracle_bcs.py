"""Pinning the oracle's boundary conditions (CPU).

Each ghost state is checked against what it is built to impose, in closed
form: a characteristic boundary whose interior state equals its free stream
returns that state on every branch (sub/supersonic, in/outflow); walls
reflect or zero the velocity; an isothermal wall state has the wall
temperature; an adiabatic wall's gradient state has no normal temperature
gradient; the subsonic inflow/outflow states keep the quantities the
templates hold fixed.  On whole meshes: free-stream preservation with
every Euler and NS boundary type, and conservation of mass, energy and
tangential momentum in a box with slip walls.
"""
import numpy as np
import pytest

from oracle.bcs import (NS_TYPES, bc_ldg_grad_state, bc_ldg_state, bc_state,
                        bccflux_euler, bccflux_ns, ftpttang_consts)
from oracle.mesh import box_hex_verts
from oracle.physics import inviscid_flux, rusanov
from oracle.rhs import OracleHexSystem

G = 1.4
C = {'gamma': G, 'mu': 0.02, 'Pr': 0.72, 'ldg-beta': 0.0, 'ldg-tau': 0.1,
     'cpTw': 3.5, 'cpTt': 4.2, 'pt': 1.3, 'theta': 20.0, 'phi': 70.0}


def _con(rho, vel, p):
    vel = np.asarray(vel, dtype=float)
    rho, p = rho + 0*vel[0], p + 0*vel[0]
    return np.array([rho, *(rho*vel), p/(G - 1) + 0.5*rho*(vel**2).sum(0)])


def _states(n=50, seed=0):
    rng = np.random.default_rng(seed)
    rho = rng.uniform(0.5, 2.0, n)
    vel = rng.uniform(-0.6, 0.6, (3, n))
    p = rng.uniform(0.5, 2.0, n)
    nrm = rng.standard_normal((3, n))
    nrm /= np.linalg.norm(nrm, axis=0)
    return _con(rho, vel, p), nrm, rng


def _prim(u):
    v = u[1:4]/u[0]
    return u[0], v, (G - 1)*(u[4] - 0.5*u[0]*(v*v).sum(0))


@pytest.mark.parametrize('vn,mach', [(-1, 0.5), (1, 0.5), (-1, 1.6),
                                     (1, 1.6)])
def test_char_riem_inv_free_stream_fixed_point(vn, mach):
    _, n, rng = _states()
    rho, p = 1.2, 0.9
    a = np.sqrt(G*p/rho)
    t = np.cross(n.T, rng.standard_normal((n.shape[1], 3))).T
    t /= np.linalg.norm(t, axis=0)
    vel = vn*mach*a*n + 0.2*a*t
    ul = _con(rho, vel, p)
    vals = {'rho': rho, 'p': p, 'u': vel[0], 'v': vel[1], 'w': vel[2]}
    ur = bc_state('char-riem-inv', ul, n, vals, C)
    assert np.abs(ur - ul).max() < 1e-13*np.abs(ul).max()


def test_wall_and_copy_states():
    ul, n, _ = _states()
    # slip wall: reflect the normal momentum, keep the rest
    ur = bc_state('slp-adia-wall', ul, n, {}, C)
    mn = (ul[1:4]*n).sum(0)
    assert np.allclose((ur[1:4]*n).sum(0), -mn, atol=1e-14)
    assert np.allclose(ur[1:4] - ul[1:4], -2*mn*n, atol=1e-14)
    assert np.array_equal(ur[[0, 4]], ul[[0, 4]])
    # no-slip: opposite velocity for the Riemann state, zero for LDG, same
    # internal energy
    ur = bc_state('no-slp-adia-wall', ul, n, {}, C)
    assert np.array_equal(ur[1:4], -ul[1:4])
    ug = bc_ldg_state('no-slp-adia-wall', ul, n, {}, C)
    assert np.all(ug[1:4] == 0)
    assert np.allclose(_prim(ug)[2], _prim(ul)[2], rtol=1e-13)
    # supersonic outflow: the interior state; its flux is f(ul).n
    assert np.array_equal(bc_state('sup-out-fn', ul, n, {}, C), ul)
    f, _, _ = inviscid_flux(ul, G)
    fn = bccflux_euler('sup-out-fn', rusanov, ul, 2.5*n, {}, C)
    assert np.allclose(fn, 2.5*np.einsum('dvn,dn->vn', f, n), rtol=1e-13,
                       atol=1e-14)


def test_slip_wall_flux_has_no_mass_or_energy_flux():
    ul, n, _ = _states()
    fn = bccflux_euler('slp-adia-wall', rusanov, ul, n, {}, C)
    assert np.abs(fn[[0, 4]]).max() < 1e-14
    # momentum flux is normal to the wall
    t = fn[1:4] - (fn[1:4]*n).sum(0)*n
    assert np.abs(t).max() < 1e-13


def test_isothermal_wall_temperature_and_velocity():
    ul, n, _ = _states()
    vals = {'u': 0.1, 'v': -0.2, 'w': 0.05}
    for f in (bc_state, bc_ldg_state):
        ur = f('no-slp-isot-wall', ul, n, vals, C)
        rho, v, p = _prim(ur)
        assert np.allclose(G/(G - 1)*p/rho, C['cpTw'], rtol=1e-13)
        assert np.array_equal(ur[0], ul[0])
    ug = bc_ldg_state('no-slp-isot-wall', ul, n, vals, C)
    assert np.allclose(_prim(ug)[1], np.array([0.1, -0.2, 0.05])[:, None],
                       rtol=1e-13)
    # the Riemann state's velocity mirrors the interior's about the wall's
    assert np.allclose(_prim(bc_state('no-slp-isot-wall', ul, n, vals,
                                      C))[1] + _prim(ul)[1],
                       2*np.array([0.1, -0.2, 0.05])[:, None], rtol=1e-12)


def test_adiabatic_wall_gradient_has_no_normal_heat_flux():
    ul, n, rng = _states()
    gradul = rng.standard_normal((3, 5, ul.shape[1]))
    ur = bc_ldg_state('no-slp-adia-wall', ul, n, {}, C)
    gr = bc_ldg_grad_state('no-slp-adia-wall', ur, n, gradul)
    # c_v rho grad T from a state's gradient (flux.mako:74-76 form), with
    # the wall state's zero velocity: E_x - rho_x E / rho
    Tg = gr[:, 4] - gr[:, 0]*ur[4]/ur[0]
    assert np.abs((Tg*n).sum(0)).max() < 1e-12*np.abs(gradul).max()
    # tangential components unchanged; other variables copied
    Tl = gradul[:, 4] - gradul[:, 0]*ur[4]/ur[0]
    tang = Tl - (Tl*n).sum(0)*n
    assert np.allclose(Tg, tang, atol=1e-12)
    assert np.array_equal(gr[:, :4], gradul[:, :4])


def test_subsonic_inflow_outflow_states():
    ul, n, _ = _states()
    rl, vl, pl = _prim(ul)
    # sub-in-frv: prescribed density and velocity, interior pressure
    vals = {'rho': 1.1, 'u': 0.3, 'v': 0.0, 'w': -0.1}
    r, v, p = _prim(bc_state('sub-in-frv', ul, n, vals, C))
    assert np.allclose(r, 1.1) and np.allclose(p, pl, rtol=1e-12)
    assert np.allclose(v, np.array([0.3, 0.0, -0.1])[:, None])
    # sub-out-fp: prescribed pressure, interior density and velocity
    r, v, p = _prim(bc_state('sub-out-fp', ul, n, {'p': 0.8}, C))
    assert np.allclose(p, 0.8, rtol=1e-12)
    assert np.array_equal(r, rl) and np.allclose(v, vl, rtol=1e-13)
    # sub-in-ftpttang: interior pressure, total enthalpy cpTt, isentropic
    # relation to the total pressure, prescribed flow direction
    ul = _con(rl, 0.1*vl, 1.0 + 0.2*(pl - 1))
    pl = _prim(ul)[2]
    r, v, p = _prim(bc_state('sub-in-ftpttang', ul, n, {}, C))
    Rdcp, vc = ftpttang_consts(C)
    q = (v*v).sum(0)
    assert np.allclose(p, pl, rtol=1e-12)
    assert np.allclose(p/(Rdcp*r) + 0.5*q, C['cpTt'], rtol=1e-12)
    assert np.allclose(pl/C['pt'], (1 - 0.5*q/C['cpTt'])**(1/Rdcp),
                       rtol=1e-12)
    assert np.allclose(v/np.sqrt(q), vc[:, None], rtol=1e-12)


UNIF_PRI = (1.0, 0.3, -0.2, 0.1, 2.5)


def _uniform(o, pri=UNIF_PRI):
    u = np.empty((o.nupts, 5, o.neles))
    u[:] = _con(pri[0], np.array(pri[1:4]), pri[4])[None, :, None]
    return u


def _fs_vals(pri=UNIF_PRI):
    return {'rho': pri[0], 'u': pri[1], 'v': pri[2], 'w': pri[3],
            'p': pri[4]}


@pytest.mark.parametrize('system,types', [
    ('euler', ('char-riem-inv', 'sup-in-fa', 'sup-out-fn')),
    ('navier-stokes', ('char-riem-inv', 'sup-in-fa', 'sup-out-fn')),
    ('navier-stokes', ('sub-in-frv', 'sub-out-fp', 'char-riem-inv'))])
def test_free_stream_preservation_with_boundaries(system, types):
    c = {'gamma': G} if system == 'euler' else dict(C)
    warp = lambda x, y, z: (x + 0.1*np.sin(y)*np.sin(z) *
                            np.sin(x/2)**2, y, z)
    verts = box_hex_verts(3, warp=warp)
    fs = _fs_vals()
    bcs = {'x0': (types[0], fs), 'x1': (types[1], fs),
           'y0': (types[2], fs), 'y1': (types[2], fs)}
    o = OracleHexSystem(3, 2, system, constants=c, verts=verts,
                        periodic=(False, False, True), bcs=bcs)
    assert len(o.bcs) == 4
    assert np.abs(o.rhs(_uniform(o))).max() < 1e-11


@pytest.mark.parametrize('bctype', ['slp-adia-wall', 'no-slp-adia-wall',
                                    'no-slp-isot-wall'])
def test_fluid_at_rest_between_walls(bctype):
    pri = (1.3, 0.0, 0.0, 0.0, 2.0)
    c = dict(C, cpTw=G/(G - 1)*pri[4]/pri[0])
    bcs = {s: (bctype, {'u': 0.0, 'v': 0.0, 'w': 0.0}) for s in ('x0', 'x1')}
    o = OracleHexSystem(3, 2, 'navier-stokes', constants=c,
                        periodic=(False, True, True), bcs=bcs)
    assert np.abs(o.rhs(_uniform(o, pri))).max() < 1e-12


@pytest.mark.parametrize('system', ['euler', 'navier-stokes'])
def test_slip_wall_box_conservation(system):
    c = {'gamma': G} if system == 'euler' else dict(C, **{'ldg-tau': 0.0})
    bcs = {s: ('slp-adia-wall', {}) for s in ('x0', 'x1')}
    o = OracleHexSystem(3, 3, system, constants=c,
                        periodic=(False, True, True), bcs=bcs)
    from oracle.rhs import TGV_ICS
    ics = dict(TGV_ICS, p=lambda x, y, z: TGV_ICS['p'](x, y, z) + 0.1*x)
    r = o.rhs(o.ics(ics, vscale=0.3))
    w = o.shape.upts_wts
    tot = np.array([(w[:, None]*r[:, v]/o.rcpdjac).sum() for v in range(5)])
    if system == 'euler':
        # mass, tangential momentum and energy cannot cross a slip wall
        assert np.abs(tot[[0, 2, 3, 4]]).max() < 1e-10
    else:
        # the viscous wall flux (ghost-free slip wall: none) likewise
        assert np.abs(tot[[0, 4]]).max() < 1e-10
    # the wall pressure force (higher pressure at x1) does act on the
    # normal momentum
    assert abs(tot[1]) > 1e-2


def test_ns_bc_flux_types_run():
    ul, n, rng = _states(20)
    gl = 0.1*rng.standard_normal((3, 5, 20))
    for t in NS_TYPES:
        vals = _fs_vals()
        f = bccflux_ns(t, rusanov, ul, gl, 1.7*n, vals, C)
        assert f.shape == ul.shape and np.all(np.isfinite(f))


def test_controlled_surface_integral_and_update():
    """The controlled char-riem-inv oracle: its surface integral of a
    uniform flow (mass flow -rho u A through the x0 side, mean pressure p,
    area A) and one controller update by hand
    (pyfr/solvers/euler/inters.py:165-186)."""
    from oracle.rhs import OracleHexSystem
    L = 2*np.pi
    ctrl = {'target': -1.0, 'alpha': 0.5, 'eta': 2.0, 'nsteps': 1,
            'qdeg': 5, 'p': 0.8}
    bcs = {'x0': ('char-riem-inv-mass-flow',
                  {'rho': 1.0, 'u': 0.3, 'v': 0.0, 'w': 0.0, 'ctrl': ctrl}),
           'x1': ('char-riem-inv-pressure',
                  {'rho': 1.0, 'u': 0.3, 'v': 0.0, 'w': 0.0,
                   'ctrl': dict(ctrl, target=0.7)})}
    o = OracleHexSystem(3, 2, 'euler', periodic=(False, True, True),
                        bcs=bcs)
    u = np.empty((o.nupts, 5, o.neles))
    u[:, 0], u[:, 1], u[:, 2:4] = 1.2, 1.2*0.3, 0.0
    u[:, 4] = 0.9/0.4 + 0.5*1.2*0.09
    mf, pr = o._ctrls['x0'], o._ctrls['x1']
    assert abs(mf._measure(u) + 1.2*0.3*L*L) < 1e-12*L*L
    assert abs(pr._measure(u) - 0.9) < 1e-13
    assert abs(pr.area - L*L) < 1e-12*L*L
    # first measurements at t = 0.1 (tprev set), the update at t = 0.3
    mf.prepare(u, 0.1)
    mf.prepare(u, 0.3)
    meas = -1.2*0.3*L*L
    p1 = 0.8 + 0.2*2.0*(1 + 1.0/meas)
    # the new line passes through (t, p0) and reaches p1 a step later
    assert abs(mf.interp_m - (p1 - 0.8)/0.2) < 1e-12
    assert abs(mf.interp_c + mf.interp_m*0.3 - 0.8) < 1e-12
    assert abs(mf.interp_c + mf.interp_m*0.5 - p1) < 1e-12
