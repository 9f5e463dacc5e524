"""Oracle: boundary-condition ghost states and boundary common fluxes
(test infrastructure).

Arrays follow ``oracle.physics``: first axis variable (or direction), then
points.  ``n`` is the unit outward normal (3, npts); ``vals`` holds the
boundary's prescribed values ('rho', 'p', 'u', 'v', 'w' as the type needs),
scalars or per-point arrays; ``c`` the constants ('gamma', and for NS 'mu',
'Pr', 'ldg-tau', plus the type's own: 'cpTw'; 'cpTt', 'pt', 'theta', 'phi').

Euler: ``bc_state(type, ...)`` is the macro ``bc_rsolve_state`` of
``pyfr/solvers/euler/kernels/bcs/<type>.mako`` and ``bccflux_euler`` the
kernel ``pyfr/solvers/euler/kernels/bccflux.mako:7-25``.
Navier-Stokes: ``bc_rsolve_state`` / ``bc_ldg_state`` /
``bc_ldg_grad_state`` of ``pyfr/solvers/navstokes/kernels/bcs/<type>.mako``,
``bcconu`` (``navstokes/kernels/bcconu.mako:6-15``) and ``bccflux_ns`` with
the common-flux states ``ghost`` / ``ghost-imperm`` / ``slp-adia-wall``
(``navstokes/kernels/bcs/ghost.mako``, ``ghost-imperm.mako``,
``slp-adia-wall.mako:21-37``; the state per type is ``cflux_state`` of
``pyfr/solvers/navstokes/inters.py:111-219``).
"""
import numpy as np

from oracle.physics import viscous_flux_add

EULER_TYPES = ('sup-in-fa', 'sup-out-fn', 'char-riem-inv', 'slp-adia-wall',
               'char-riem-inv-mass-flow', 'char-riem-inv-pressure')
NS_TYPES = ('sup-in-fa', 'sup-out-fn', 'char-riem-inv', 'slp-adia-wall',
            'no-slp-adia-wall', 'no-slp-isot-wall', 'sub-in-frv',
            'sub-in-ftpttang', 'sub-out-fp', 'char-riem-inv-mass-flow',
            'char-riem-inv-pressure')


def _ke(u):
    """|rho v|^2 of a state (the templates' pyfr.dot('u[{i}]', i=(1, 4)))."""
    return u[1]*u[1] + u[2]*u[2] + u[3]*u[3]


def _sup_in(ul, n, vals, c):
    """``euler/kernels/bcs/sup-in-fa.mako:3-11``."""
    g = c['gamma']
    rho = vals['rho']*np.ones_like(ul[0])
    ur = np.empty_like(ul)
    ur[0] = rho
    for i, v in enumerate('uvw'):
        ur[i + 1] = rho*vals[v]
    ur[4] = vals['p']/(g - 1) + 0.5*(1.0/ur[0])*_ke(ur)
    return ur


def _copy(ul, n, vals, c):
    """``euler/kernels/bcs/sup-out-fn.mako:3-7``."""
    return ul.copy()


def _char_riem_inv(ul, n, vals, c):
    """``euler/kernels/bcs/char-riem-inv.mako:6-43``."""
    g = c['gamma']
    gmo = g - 1.0
    rho_e, p_e = vals['rho'], vals['p']
    uvw = [vals['u'], vals['v'], vals['w']]
    cs = np.sqrt(g*p_e/rho_e)
    s = p_e*np.power(rho_e, -g)
    ratio = cs*(2.0/gmo)
    inv = 1.0/ul[0]
    V_e = uvw[0]*n[0] + uvw[1]*n[1] + uvw[2]*n[2]
    V_i = inv*(ul[1]*n[0] + ul[2]*n[1] + ul[3]*n[2])
    p_i = gmo*ul[4] - (0.5*gmo)*inv*_ke(ul)
    c_i = np.sqrt(g*p_i*inv)
    R_e = np.where((np.abs(V_e) >= cs) & (V_i >= 0), V_i - c_i*(2.0/gmo),
                   V_e - ratio)
    R_i = np.where((np.abs(V_e) >= cs) & (V_i < 0), V_e + ratio,
                   V_i + c_i*(2.0/gmo))
    V_b = 0.5*(R_e + R_i)
    c_b = (0.25*gmo)*(R_i - R_e)
    rho_b = np.where(V_i < 0,
                     np.power((1.0/(g*s))*c_b*c_b, 1.0/gmo),
                     ul[0]*np.power(ul[0]*c_b*c_b/(g*p_i), 1.0/gmo))
    p_b = (1.0/g)*rho_b*c_b*c_b
    ur = np.empty_like(ul)
    ur[0] = rho_b
    for i in range(3):
        ur[i + 1] = np.where(V_i >= 0,
                             rho_b*(ul[i + 1]*inv + (V_b - V_i)*n[i]),
                             rho_b*(uvw[i] + (V_b - V_e)*n[i]))
    ur[4] = p_b*(1.0/gmo) + 0.5*(1.0/ur[0])*_ke(ur)
    return ur


def _reflect(ul, n, vals, c):
    """``euler/kernels/bcs/slp-adia-wall.mako:3-10``."""
    nor = ul[1]*n[0] + ul[2]*n[1] + ul[3]*n[2]
    ur = np.empty_like(ul)
    ur[0] = ul[0]
    for i in range(3):
        ur[i + 1] = ul[i + 1] - 2*nor*n[i]
    ur[4] = ul[4]
    return ur


def _noslp_adia_rs(ul, n, vals, c):
    """``navstokes/kernels/bcs/no-slp-adia-wall.mako:4-11``."""
    ur = np.empty_like(ul)
    ur[0] = ul[0]
    ur[1:4] = -ul[1:4]
    ur[4] = ul[4]
    return ur


def _noslp_adia_ldg(ul, n, vals, c):
    """``navstokes/kernels/bcs/no-slp-adia-wall.mako:13-20``."""
    ur = np.empty_like(ul)
    ur[0] = ul[0]
    ur[1:4] = 0.0
    ur[4] = ul[4] - (0.5/ul[0])*_ke(ul)
    return ur


def _isot_rs(ul, n, vals, c):
    """``navstokes/kernels/bcs/no-slp-isot-wall.mako:4-11``."""
    ur = np.empty_like(ul)
    ur[0] = ul[0]
    for i, v in enumerate('uvw'):
        ur[i + 1] = -ul[i + 1] + 2*vals[v]*ul[0]
    ur[4] = (c['cpTw']/c['gamma'])*ur[0] + 0.5*(1.0/ur[0])*_ke(ur)
    return ur


def _isot_ldg(ul, n, vals, c):
    """``navstokes/kernels/bcs/no-slp-isot-wall.mako:13-20``."""
    ur = np.empty_like(ul)
    ur[0] = ul[0]
    for i, v in enumerate('uvw'):
        ur[i + 1] = vals[v]*ul[0]
    ur[4] = (c['cpTw']/c['gamma'])*ur[0] + 0.5*(1.0/ur[0])*_ke(ur)
    return ur


def _sub_in_frv(ul, n, vals, c):
    """``navstokes/kernels/bcs/sub-in-frv.mako:4-13``."""
    rho = vals['rho']*np.ones_like(ul[0])
    ur = np.empty_like(ul)
    ur[0] = rho
    for i, v in enumerate('uvw'):
        ur[i + 1] = rho*vals[v]
    ur[4] = (ul[4] - 0.5*(1.0/ul[0])*_ke(ul) + 0.5*(1.0/ur[0])*_ke(ur))
    return ur


def ftpttang_consts(c):
    """Host constants of ``NavierStokesSubInflowFtpttangBCInters``
    (``pyfr/solvers/navstokes/inters.py:177-203``): Rdcp and the velocity
    direction from the inflow angles theta, phi (degrees)."""
    g = c['gamma']
    theta = c['theta']*np.pi/180.0
    phi = c['phi']*np.pi/180.0
    vc = np.array([np.cos(theta), np.sin(theta), 1.0])
    vc[:2] *= np.sin(phi)
    vc[2] *= np.cos(phi)
    return (g - 1.0)/g, vc


def _sub_in_ftpttang(ul, n, vals, c):
    """``navstokes/kernels/bcs/sub-in-ftpttang.mako:4-18``."""
    g = c['gamma']
    Rdcp, vc = ftpttang_consts(c)
    cpTt, pt = c['cpTt'], c['pt']
    pl = (g - 1.0)*(ul[4] - (0.5/ul[0])*_ke(ul))
    udotu = (2.0*cpTt)*(1.0 - pt**(-Rdcp)*np.power(pl, Rdcp))
    udotu = np.maximum(0, udotu)
    ur = np.empty_like(ul)
    ur[0] = (1.0/Rdcp)*pl/(cpTt - 0.5*udotu)
    for i in range(3):
        ur[i + 1] = vc[i]*ur[0]*np.sqrt(udotu)
    ur[4] = (1.0/(g - 1.0))*pl + 0.5*ur[0]*udotu
    return ur


def _sub_out_fp(ul, n, vals, c):
    """``navstokes/kernels/bcs/sub-out-fp.mako:4-10``."""
    ur = ul.copy()
    ur[4] = vals['p']/(c['gamma'] - 1) + 0.5*(1.0/ul[0])*_ke(ul)
    return ur


# rsolve state and LDG state per type (the NS templates alias the LDG state
# to the rsolve state unless they define both)
_RS = {'sup-in-fa': _sup_in, 'sup-out-fn': _copy,
       'char-riem-inv': _char_riem_inv, 'slp-adia-wall': _reflect,
       'no-slp-adia-wall': _noslp_adia_rs, 'no-slp-isot-wall': _isot_rs,
       'sub-in-frv': _sub_in_frv, 'sub-in-ftpttang': _sub_in_ftpttang,
       'sub-out-fp': _sub_out_fp}
_LDG = dict(_RS, **{'no-slp-adia-wall': _noslp_adia_ldg,
                    'no-slp-isot-wall': _isot_ldg})
# gradient state: 'zero' / 'copy' (navstokes/kernels/bcs/common.mako) or
# the adiabatic-wall correction
_GRAD = {'sup-in-fa': 'zero', 'sup-out-fn': 'copy', 'char-riem-inv': 'zero',
         'slp-adia-wall': None, 'no-slp-adia-wall': 'adia',
         'no-slp-isot-wall': 'copy', 'sub-in-frv': 'zero',
         'sub-in-ftpttang': 'copy', 'sub-out-fp': 'zero'}
CFLUX_STATE = {'no-slp-adia-wall': 'ghost-imperm',
               'no-slp-isot-wall': 'ghost-imperm', 'slp-adia-wall': None}
# the controlled types are char-riem-inv with p = t*im + ic
# (euler/kernels/bcs/char-riem-inv-controlled.mako; navstokes/kernels/bcs/
# char-riem-inv-mass-flow.mako, char-riem-inv-pressure.mako): their
# caller passes that p in vals
for _t in ('char-riem-inv-mass-flow', 'char-riem-inv-pressure'):
    _RS[_t], _LDG[_t], _GRAD[_t] = _RS['char-riem-inv'], _LDG[
        'char-riem-inv'], _GRAD['char-riem-inv']


def bc_state(bctype, ul, n, vals, c):
    """bc_rsolve_state."""
    return _RS[bctype](ul, n, vals, c)


def bc_ldg_state(bctype, ul, n, vals, c):
    return _LDG[bctype](ul, n, vals, c)


def bc_ldg_grad_state(bctype, ur, n, gradul):
    """gradul (3, 5, npts) -> gradur."""
    kind = _GRAD[bctype]
    if kind == 'zero':
        return np.zeros_like(gradul)
    gradur = gradul.copy()
    if kind == 'adia':
        # no-slp-adia-wall.mako:22-80 (3-D branch)
        rcprho = 1.0/ur[0]
        u, v, w = rcprho*ur[1], rcprho*ur[2], rcprho*ur[3]
        g = gradul
        vel = (u, v, w)
        dv = [[g[d][1 + i] - vel[i]*g[d][0] for d in range(3)]
              for i in range(3)]            # dv[i][d] = rho d(v_i)/dx_d
        Tl = [g[d][4] - (rcprho*g[d][0]*ur[4] + u*dv[0][d] + v*dv[1][d] +
                         w*dv[2][d]) for d in range(3)]
        for i in range(3):
            gradur[i][4] = gradur[i][4] - (n[i]*n[0]*Tl[0] + n[i]*n[1]*Tl[1] +
                                           n[i]*n[2]*Tl[2])
    return gradur


def bccflux_euler(bctype, rsolver, ul, nl, vals, c):
    """``euler/kernels/bccflux.mako:7-25``: the scaled common normal flux
    written back to the boundary flux points."""
    mag = np.sqrt(nl[0]*nl[0] + nl[1]*nl[1] + nl[2]*nl[2])
    nn = (1/mag)*nl
    ur = bc_state(bctype, ul, nn, vals, c)
    return mag*rsolver(ul, ur, nn, c['gamma'])


def bcconu(bctype, ul, nl, vals, c):
    """``navstokes/kernels/bcconu.mako:6-15``."""
    mag = np.sqrt(nl[0]*nl[0] + nl[1]*nl[1] + nl[2]*nl[2])
    return bc_ldg_state(bctype, ul, (1/mag)*nl, vals, c)


def bccflux_ns(bctype, rsolver, ul, gradul, nl, vals, c):
    """``navstokes/kernels/bccflux.mako:10-20`` with the type's common-flux
    state."""
    mag = np.sqrt(nl[0]*nl[0] + nl[1]*nl[1] + nl[2]*nl[2])
    nn = (1/mag)*nl
    state = CFLUX_STATE.get(bctype, 'ghost')
    if state is None:
        ur = bc_state(bctype, ul, nn, vals, c)
        return mag*rsolver(ul, ur, nn, c['gamma'])
    url = bc_ldg_state(bctype, ul, nn, vals, c)
    gsrc = url if state == 'ghost-imperm' else ul
    gradur = bc_ldg_grad_state(bctype, gsrc, nn, gradul)
    fvr = viscous_flux_add(url, gradur, np.zeros((3, 5) + ul.shape[1:]), c)
    ur = bc_state(bctype, ul, nn, vals, c)
    fi = rsolver(ul, ur, nn, c['gamma'])
    out = np.empty_like(fi)
    tau = c.get('ldg-tau', 0.0)
    for i in range(5):
        fv = nn[0]*fvr[0, i] + nn[1]*fvr[1, i] + nn[2]*fvr[2, i]
        if state == 'ghost' and tau != 0.0:
            fv = fv + tau*(ul[i] - ur[i])
        out[i] = mag*(fi[i] + fv)
    return out
