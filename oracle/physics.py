"""Oracle: Euler / Navier-Stokes point physics (test infrastructure).

All functions act on arrays whose first axis is the variable/dimension and
whose trailing axes are points.  Expression order follows the reference
templates so that rounding differs only where NumPy's evaluation differs
from C's.
"""
import numpy as np


def inviscid_flux(s, gamma):
    """``pyfr/solvers/euler/kernels/flux.mako:3-26``.

    s: (5, ...) -> f (3, 5, ...), p, v (3, ...)."""
    invrho = 1.0/s[0]
    E = s[4]
    rhov = s[1:4]
    v = invrho*rhov
    p = (gamma - 1)*(E - 0.5*invrho*(rhov[0]*rhov[0] + rhov[1]*rhov[1] +
                                     rhov[2]*rhov[2]))
    f = np.empty((3, 5) + s.shape[1:])
    for i in range(3):
        f[i, 0] = rhov[i]
        f[i, 4] = (E + p)*v[i]
        for j in range(3):
            f[i, j + 1] = rhov[i]*v[j] + (p if i == j else 0)
    return f, p, v


def rusanov(ul, ur, n, gamma):
    """``pyfr/solvers/euler/kernels/rsolvers/rusanov.mako:4-25``."""
    fl, pl, vl = inviscid_flux(ul, gamma)
    fr, pr, vr = inviscid_flux(ur, gamma)
    nv = (n[0]*(vl[0] + vr[0]) + n[1]*(vl[1] + vr[1]) + n[2]*(vl[2] + vr[2]))
    a = np.sqrt((0.25*gamma)*(pl + pr)/(ul[0] + ur[0])) + 0.25*np.abs(nv)
    nf = np.empty_like(ul)
    for i in range(5):
        nf[i] = 0.5*(n[0]*(fl[0, i] + fr[0, i]) + n[1]*(fl[1, i] + fr[1, i]) +
                     n[2]*(fl[2, i] + fr[2, i])) + a*(ul[i] - ur[i])
    return nf


def hllc(ul, ur, n, gamma):
    """``pyfr/solvers/euler/kernels/rsolvers/hllc.mako:4-77``."""
    fl, pl, vl = inviscid_flux(ul, gamma)
    fr, pr, vr = inviscid_flux(ur, gamma)
    nvl = n[0]*vl[0] + n[1]*vl[1] + n[2]*vl[2]
    nvr = n[0]*vr[0] + n[1]*vr[1] + n[2]*vr[2]
    al = np.sqrt(gamma*pl/ul[0])
    ar = np.sqrt(gamma*pr/ur[0])
    srl, srr = np.sqrt(ul[0]), np.sqrt(ur[0])
    nv = (srl*nvl + srr*nvr)/(srl + srr)
    H = (srl*(pr + ur[4]) + srr*(pl + ul[4]))/(srl*ur[0] + srr*ul[0])
    inv_rar = 1/(srl + srr)
    va = [(vl[i]*srl + vr[i]*srr)*inv_rar for i in range(3)]
    qq = va[0]*va[0] + va[1]*va[1] + va[2]*va[2]
    a = np.sqrt((gamma - 1)*(H - 0.5*qq))
    sl = np.minimum(nv - a, nvl - al)
    sr = np.maximum(nv + a, nvr + ar)
    sstar = ((pr - pl + ul[0]*nvl*(sl - nvl) - ur[0]*nvr*(sr - nvr)) /
             (ul[0]*(sl - nvl) - ur[0]*(sr - nvr)))
    ul_com = (sl - nvl)/(sl - sstar)
    ur_com = (sr - nvr)/(sr - sstar)
    usl = np.empty_like(ul)
    usr = np.empty_like(ur)
    usl[0] = ul_com*ul[0]
    usr[0] = ur_com*ur[0]
    for i in range(3):
        usl[i + 1] = usl[0]*(vl[i] + (sstar - nvl)*n[i])
        usr[i + 1] = usr[0]*(vr[i] + (sstar - nvr)*n[i])
    usl[4] = ul_com*(ul[4] + (sstar - nvl)*(ul[0]*sstar + pl/(sl - nvl)))
    usr[4] = ur_com*(ur[4] + (sstar - nvr)*(ur[0]*sstar + pr/(sr - nvr)))
    nf = np.empty_like(ul)
    for i in range(5):
        nf_fl = n[0]*fl[0, i] + n[1]*fl[1, i] + n[2]*fl[2, i]
        nf_fr = n[0]*fr[0, i] + n[1]*fr[1, i] + n[2]*fr[2, i]
        nf_fsl = nf_fl + sl*(usl[i] - ul[i])
        nf_fsr = nf_fr + sr*(usr[i] - ur[i])
        nf[i] = np.where(0 <= sl, nf_fl,
                         np.where((sl <= 0) & (0 <= sstar), nf_fsl,
                                  np.where((sstar <= 0) & (0 <= sr),
                                           nf_fsr, nf_fr)))
    return nf


RSOLVERS = {'rusanov': rusanov, 'hllc': hllc}


def viscous_flux_add(u, g, f, c):
    """Add the NS viscous flux (constant mu or Sutherland) to f in place.

    u: (5, ...), g: (3, 5, ...) physical gradient, f: (3, 5, ...);
    ``pyfr/solvers/navstokes/kernels/flux.mako:48-105``."""
    rho, rhou, rhov, rhow, E = u
    rcprho = 1.0/rho
    uu, vv, ww = rcprho*rhou, rcprho*rhov, rcprho*rhow
    rho_x, rho_y, rho_z = g[0, 0], g[1, 0], g[2, 0]
    u_x = g[0, 1] - uu*rho_x
    u_y = g[1, 1] - uu*rho_y
    u_z = g[2, 1] - uu*rho_z
    v_x = g[0, 2] - vv*rho_x
    v_y = g[1, 2] - vv*rho_y
    v_z = g[2, 2] - vv*rho_z
    w_x = g[0, 3] - ww*rho_x
    w_y = g[1, 3] - ww*rho_y
    w_z = g[2, 3] - ww*rho_z
    E_x, E_y, E_z = g[0, 4], g[1, 4], g[2, 4]

    if c.get('visc_corr', 'none') == 'sutherland':
        cpT = c['gamma']*(rcprho*E - 0.5*(uu*uu + vv*vv + ww*ww))
        Trat = (1/c['cpTref'])*cpT
        mu_c = (c['mu']*(c['cpTref'] + c['cpTs']))*Trat*np.sqrt(Trat) / \
            (cpT + c['cpTs'])
    else:
        mu_c = c['mu']

    T_x = rcprho*(E_x - (rcprho*rho_x*E + uu*u_x + vv*v_x + ww*w_x))
    T_y = rcprho*(E_y - (rcprho*rho_y*E + uu*u_y + vv*v_y + ww*w_y))
    T_z = rcprho*(E_z - (rcprho*rho_z*E + uu*u_z + vv*v_z + ww*w_z))

    third = 1.0/3.0
    t_xx = -2*mu_c*rcprho*(u_x - third*(u_x + v_y + w_z))
    t_yy = -2*mu_c*rcprho*(v_y - third*(u_x + v_y + w_z))
    t_zz = -2*mu_c*rcprho*(w_z - third*(u_x + v_y + w_z))
    t_xy = -mu_c*rcprho*(v_x + u_y)
    t_xz = -mu_c*rcprho*(u_z + w_x)
    t_yz = -mu_c*rcprho*(w_y + v_z)

    gpr = c['gamma']/c['Pr']
    f[0, 1] += t_xx
    f[1, 1] += t_xy
    f[2, 1] += t_xz
    f[0, 2] += t_xy
    f[1, 2] += t_yy
    f[2, 2] += t_yz
    f[0, 3] += t_xz
    f[1, 3] += t_yz
    f[2, 3] += t_zz
    f[0, 4] += uu*t_xx + vv*t_xy + ww*t_xz + -mu_c*gpr*T_x
    f[1, 4] += uu*t_xy + vv*t_yy + ww*t_yz + -mu_c*gpr*T_y
    f[2, 4] += uu*t_xz + vv*t_yz + ww*t_zz + -mu_c*gpr*T_z
    return f


def pri_to_con(pris, gamma):
    """``pyfr/solvers/euler/elements.py:103-114``."""
    rho, p = pris[0], pris[-1]
    rhovs = [rho*c for c in pris[1:-1]]
    E = p/(gamma - 1) + 0.5*rho*sum(c*c for c in pris[1:-1])
    return [rho, *rhovs, E]
