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


def hll(ul, ur, n, gamma):
    """``pyfr/solvers/euler/kernels/rsolvers/hll.mako:4-53`` (including its
    left wave-speed estimate min(nv - a, nvl - ar))."""
    fl, pl, vl = inviscid_flux(ul, gamma)
    fr, pr, vr = inviscid_flux(ur, gamma)
    nvl = n[0]*vl[0] + n[1]*vl[1] + n[2]*vl[2]
    nvr = n[0]*vr[0] + n[1]*vr[1] + n[2]*vr[2]
    ar = np.sqrt(gamma*pr/ur[0])
    srl, srr = np.sqrt(ul[0]), np.sqrt(ur[0])
    nv = (srl*nvl + srr*nvr)/(srl + srr)
    H = (srl*(pr + ur[4]) + srr*(pl + ul[4]))/(srl*ur[0] + srr*ul[0])
    inv_rar = 1/(srl + srr)
    va = [(vl[i]*srl + vr[i]*srr)*inv_rar for i in range(3)]
    qq = va[0]*va[0] + va[1]*va[1] + va[2]*va[2]
    a = np.sqrt((gamma - 1)*(H - 0.5*qq))
    sl = np.minimum(nv - a, nvl - ar)
    sr = np.maximum(nv + a, nvr + ar)
    rcpsrsl = 1/(sr - sl)
    nf = np.empty_like(ul)
    for i in range(5):
        nf_fl = n[0]*fl[0, i] + n[1]*fl[1, i] + n[2]*fl[2, i]
        nf_fr = n[0]*fr[0, i] + n[1]*fr[1, i] + n[2]*fr[2, i]
        nf_sub = (sr*nf_fl - sl*nf_fr + sl*sr*(ur[i] - ul[i]))*rcpsrsl
        nf[i] = np.where(0 <= sl, nf_fl, np.where(0 >= sr, nf_fr, nf_sub))
    return nf


def _transform_to(n, u):
    """Rotate u[1:4] into the frame whose first axis is n (Moler & Hughes;
    ``pyfr/solvers/baseadvec/kernels/transform.mako``, 3-D branch)."""
    t = np.empty((3,) + np.shape(u[1]))
    b1 = np.abs(n[0]) < 0.99
    b2 = ~b1 & (np.abs(n[1]) < np.abs(n[2]))
    b3 = ~b1 & ~b2
    with np.errstate(divide='ignore', invalid='ignore'):
        h1, h2, h3 = 1/(1 + n[0]), 1/(1 - n[1]), 1/(1 - n[2])
    u0, u1, u2 = u[1], u[2], u[3]
    t[0] = n[0]*u0 + n[1]*u1 + n[2]*u2
    t[1] = np.where(
        b1, (n[0] + h1*n[2]*n[2])*u1 - n[1]*u0 - h1*n[1]*n[2]*u2,
        np.where(b2, (1 - h2*n[0]*n[0])*u0 + n[0]*u1 - h2*n[0]*n[2]*u2,
                 (1 - h3*n[1]*n[1])*u1 - h3*n[0]*n[1]*u0 + n[1]*u2))
    t[2] = np.where(
        b1, (n[0] + h1*n[1]*n[1])*u2 - n[2]*u0 - h1*n[1]*n[2]*u1,
        np.where(b2, n[2]*u1 - h2*n[0]*n[2]*u0 + (1 - h2*n[2]*n[2])*u2,
                 (1 - h3*n[0]*n[0])*u0 - h3*n[0]*n[1]*u1 + n[0]*u2))
    return t


def _transform_from(n, t):
    """Inverse of ``_transform_to`` (``transform.mako``, 3-D branch)."""
    u = np.empty_like(t)
    b1 = np.abs(n[0]) < 0.99
    b2 = ~b1 & (np.abs(n[1]) < np.abs(n[2]))
    with np.errstate(divide='ignore', invalid='ignore'):
        h1, h2, h3 = 1/(1 + n[0]), 1/(1 - n[1]), 1/(1 - n[2])
    t0, t1, t2 = t
    u[0] = np.where(b1, n[0]*t0 - n[1]*t1 - n[2]*t2,
                    np.where(b2, n[0]*t0 + (1 - h2*n[0]*n[0])*t1
                             - h2*n[0]*n[2]*t2,
                             n[0]*t0 - h3*n[0]*n[1]*t1
                             + (1 - h3*n[0]*n[0])*t2))
    u[1] = np.where(b1, n[1]*t0 + (n[0] + h1*n[2]*n[2])*t1
                    - h1*n[1]*n[2]*t2,
                    np.where(b2, n[1]*t0 + n[0]*t1 + n[2]*t2,
                             n[1]*t0 + (1 - h3*n[1]*n[1])*t1
                             - h3*n[0]*n[1]*t2))
    u[2] = np.where(b1, n[2]*t0 - h1*n[1]*n[2]*t1 + (n[0] + h1*n[1]*n[1])*t2,
                    np.where(b2, n[2]*t0 - h2*n[0]*n[2]*t1
                             + (1 - h2*n[2]*n[2])*t2,
                             n[2]*t0 + n[1]*t1 + n[0]*t2))
    return u


def _flux_1d(s, gamma):
    """``flux.mako: inviscid_flux_1d``."""
    invrho = 1.0/s[0]
    E = s[4]
    v = [invrho*s[1], invrho*s[2], invrho*s[3]]
    p = (gamma - 1)*(E - 0.5*invrho*(s[1]*s[1] + s[2]*s[2] + s[3]*s[3]))
    f = [s[1], s[1]*v[0] + p, s[1]*v[1], s[1]*v[2], (E + p)*v[0]]
    return f, p, v


def roe(ul, ur, n, gamma):
    """Roe with a Harten entropy fix (eps = 0.001) in the face frame;
    ``rsolvers/roe.mako:4-76`` through ``rsolve1d.mako:1-21``."""
    with np.errstate(divide='ignore', invalid='ignore'):
        utl = np.stack([ul[0], *_transform_to(n, ul), ul[4]])
        utr = np.stack([ur[0], *_transform_to(n, ur), ur[4]])
    fl, pl, vl = _flux_1d(utl, gamma)
    fr, pr, vr = _flux_1d(utr, gamma)
    srl, srr = np.sqrt(utl[0]), np.sqrt(utr[0])
    roa = srl*srr
    ha = (srl*(pr + utr[4]) + srr*(pl + utl[4]))/(srl*utr[0] + srr*utl[0])
    inv_rar = 1/(srl + srr)
    va = [(vl[i]*srl + vr[i]*srr)*inv_rar for i in range(3)]
    qq = va[0]*va[0] + va[1]*va[1] + va[2]*va[2]
    a = np.sqrt((gamma - 1)*(ha - 0.5*qq))
    eps = 0.001
    l1, l2, l3 = np.abs(va[0] - a), np.abs(va[0]), np.abs(va[0] + a)
    l1 = np.where(l1 < eps, (0.5/eps)*(l1*l1 + eps**2), l1)
    l3 = np.where(l3 < eps, (0.5/eps)*(l3*l3 + eps**2), l3)
    dv = [vr[i] - vl[i] for i in range(3)]
    dro = utr[0] - utl[0]
    dp = pr - pl
    inv_a2 = 1/(2*a*a)
    ntf = np.empty_like(utl)
    v1 = (dp - roa*a*dv[0])*inv_a2
    v2 = dro - dp*2*inv_a2
    v3 = (dp + roa*a*dv[0])*inv_a2
    ntf[0] = 0.5*(fl[0] + fr[0]) - (l1*v1 + l2*v2 + l3*v3)
    for i in range(3):
        if i == 0:
            v1 = (dp - roa*a*dv[0])*inv_a2*(va[i] - a)
            v2 = (dro - dp*2*inv_a2)*va[i]
            v3 = (dp + roa*a*dv[0])*inv_a2*(va[i] + a)
        else:
            v1 = (dp - roa*a*dv[0])*inv_a2*va[i]
            v2 = (dro - dp*2*inv_a2)*va[i] + roa*dv[i]
            v3 = (dp + roa*a*dv[0])*inv_a2*va[i]
        ntf[i + 1] = 0.5*(fl[i + 1] + fr[i + 1]) - (l1*v1 + l2*v2 + l3*v3)
    v1 = (dp - roa*a*dv[0])*inv_a2*(ha - a*va[0])
    v2 = (roa*((va[0]*dv[0] + va[1]*dv[1] + va[2]*dv[2]) - va[0]*dv[0]) +
          (dro - dp*2*inv_a2)*qq*0.5)
    v3 = (dp + roa*a*dv[0])*inv_a2*(ha + a*va[0])
    ntf[4] = 0.5*(fl[4] + fr[4]) - (l1*v1 + l2*v2 + l3*v3)
    nf = np.empty_like(ntf)
    nf[0], nf[4] = ntf[0], ntf[4]
    with np.errstate(divide='ignore', invalid='ignore'):
        nf[1:4] = _transform_from(n, ntf[1:4])
    return nf


def roem(ul, ur, n, gamma):
    """RoeM (Kim, Kim & Rho 2003); ``rsolvers/roem.mako:4-86``."""
    fl, pl, vl = inviscid_flux(ul, gamma)
    fr, pr, vr = inviscid_flux(ur, gamma)
    hl = (ul[4] + pl)/ul[0]
    hr = (ur[4] + pr)/ur[0]
    contral = n[0]*vl[0] + n[1]*vl[1] + n[2]*vl[2]
    contrar = n[0]*vr[0] + n[1]*vr[1] + n[2]*vr[2]
    drho, dp, dh = ur[0] - ul[0], pr - pl, hr - hl
    dcontra = contrar - contral
    dv = [vr[i] - vl[i] for i in range(3)]
    rrr = np.sqrt(ur[0]/ul[0])
    ratl = 1.0/(1.0 + rrr)
    ratr = rrr*ratl
    ra = rrr*ul[0]
    ha = hl*ratl + hr*ratr
    va = [vl[i]*ratl + vr[i]*ratr for i in range(3)]
    qq = va[0]*va[0] + va[1]*va[1] + va[2]*va[2]
    contraa = n[0]*va[0] + n[1]*va[1] + n[2]*va[2]
    aa = np.sqrt((gamma - 1)*(ha - 0.5*qq))
    rcp_aa = 1.0/aa
    abs_ma = np.abs(contraa*rcp_aa) + 1e-15
    b1 = np.maximum(0.0, np.maximum(contraa + aa, contrar + aa))
    b2 = np.minimum(0.0, np.minimum(contraa - aa, contral - aa))
    b1b2 = b1*b2
    rcp_b1_b2 = 1.0/(b1 - b2)
    b1, b2, b1b2 = b1*rcp_b1_b2, b2*rcp_b1_b2, b1b2*rcp_b1_b2
    sdst = np.where(pl < pr, pl/pr, pr/pl)
    h = 1.0 - sdst
    f = np.power(abs_ma, h)
    g = f/(1.0 + abs_ma)
    du = [ur[i] - ul[i] for i in range(4)] + [ur[0]*hr - ul[0]*hl]
    bdq = [None]*5
    bdq[0] = drho - f*dp*rcp_aa*rcp_aa
    bdq[4] = bdq[0]*ha + ra*dh
    for i in range(3):
        bdq[i + 1] = bdq[0]*va[i] + ra*(dv[i] - n[i]*dcontra)
    nf = np.empty_like(ul)
    for i in range(5):
        nf[i] = (n[0]*(b1*fl[0, i] - b2*fr[0, i]) +
                 n[1]*(b1*fl[1, i] - b2*fr[1, i]) +
                 n[2]*(b1*fl[2, i] - b2*fr[2, i]) +
                 b1b2*(du[i] - g*bdq[i]))
    return nf


def _star_flux(p, ps, rs, cs, g):
    """Pressure function of one side and its derivative (Toro 2009 eqs.
    4.86-4.87, zero covolume); ``rsolvers/exact.mako:14-30``."""
    gprtg, trgm = (g + 1)/(2*g), 2/(g - 1)
    trgp, gmrgp = 2/(g + 1), (g - 1)/(g + 1)
    pr = p/ps
    pr_g = np.power(pr, -gprtg)
    f_raref, fd_raref = trgm*cs*(pr_g*pr - 1), pr_g/(rs*cs)
    pbs_inv = 1/(p + gmrgp*ps)
    sapb = np.sqrt(trgp*pbs_inv/rs)
    f_shock = (p - ps)*sapb
    fd_shock = sapb - 0.5*f_shock*pbs_inv
    raref = p <= ps
    return (np.where(raref, f_raref, f_shock),
            np.where(raref, fd_raref, fd_shock))


def _exact_1d(ul, ur, g, p_min, kmax=3):
    """``rsolvers/exact.mako:49-206`` in the face frame: initial guess
    (primitive-variable / two-rarefaction / two-shock), kmax Newton steps
    on the star pressure (negative iterates replaced by p_min), then the
    sampled state at x/t = 0 and its 1-D flux."""
    hgm, grgm, trgm = 0.5*(g - 1), g/(g - 1), 2/(g - 1)
    trgp, gmrtg, gprtg = 2/(g + 1), (g - 1)/(2*g), (g + 1)/(2*g)
    tgrgm, gmrgp = (2*g)/(g - 1), (g - 1)/(g + 1)

    def prim(s):
        invrho = 1/s[0]
        v = invrho*s[1:4]
        p = (g - 1)*(s[4] - 0.5*invrho*(s[1]*s[1] + s[2]*s[2] + s[3]*s[3]))
        return p, v, np.sqrt(g*p*invrho)

    pl, vl, cl = prim(ul)
    pr, vr, cr = prim(ur)
    rl, rr = ul[0], ur[0]
    bpv = np.maximum(0, 0.125*(vl[0] - vr[0])*(rl + rr)*(cl + cr) +
                     0.5*(pl + pr))
    pmin, pmax = np.minimum(pl, pr), np.maximum(pl, pr)
    pre = np.power(pl/pr, gmrtg)
    um = ((pre*vl[0]*cr + vr[0]*cl + trgm*(pre - 1)*cl*cr) /
          (pre*cr + cl))
    ptl = 1 - hgm*(um - vl[0])/cl
    ptr = 1 + hgm*(um - vr[0])/cr
    p_rr = 0.5*(pl*np.power(ptl, tgrgm) + pr*np.power(ptr, tgrgm))
    gl = np.sqrt(trgp/(rl*(gmrgp*pl + bpv)))
    gr = np.sqrt(trgp/(rr*(gmrgp*pr + bpv)))
    p_ss = (gl*pl + gr*pr - vr[0] + vl[0])/(gl + gr)
    p0 = np.where((pmax <= 2*pmin) & (pmin <= bpv) & (bpv <= pmax), bpv,
                  np.where(bpv < pmin, p_rr, p_ss))
    for _ in range(kmax):
        fsl, fdl = _star_flux(p0, pl, rl, cl, g)
        fsr, fdr = _star_flux(p0, pr, rr, cr, g)
        p1 = p0 - (fsl + fsr + vr[0] - vl[0])/(fdl + fdr)
        p0 = np.where(p1 < 0, p_min, p1)
    # u* from the last iteration's side functions, evaluated at the
    # pressure before the final update (exact.mako:91-97)
    us = 0.5*(vl[0] + vr[0] + fsr - fsl)

    # sampled primitive state (rho, u_n, p) by the decision tree
    def st(r, u, p):
        return np.array([r, u, p])
    L, R = st(rl, vl[0], pl), st(rr, vr[0], pr)
    # left of the contact
    cl_fan = trgp + gmrgp*vl[0]/cl
    left_raref = np.where(
        vl[0] >= cl, L,
        np.where(us < cl*np.power(p0/pl, gmrtg),
                 st(rl*np.power(p0/pl, 1/g), us, p0),
                 st(rl*np.power(cl_fan, trgm), trgp*cl + gmrgp*vl[0],
                    pl*np.power(cl_fan, tgrgm))))
    p0pl = p0/pl
    sl = vl[0] - cl*np.sqrt(gprtg*p0pl + gmrtg)
    left_shock = np.where(0 <= sl, L,
                          st(rl*(p0pl + gmrgp)/(p0pl*gmrgp + 1), us, p0))
    wl = np.where(p0 <= pl, left_raref, left_shock)
    # right of the contact
    p0pr = p0/pr
    sr = vr[0] + cr*np.sqrt(gprtg*p0pr + gmrtg)
    right_shock = np.where(sr <= 0, R,
                           st(rr*(p0pr + gmrgp)/(p0pr*gmrgp + 1), us, p0))
    cr_fan = trgp - gmrgp*vr[0]/cr
    right_raref = np.where(
        vr[0] + cr <= 0, R,
        np.where(us + cr*np.power(p0pr, gmrtg) >= 0,
                 st(rr*np.power(p0pr, 1/g), us, p0),
                 st(rr*np.power(cr_fan, trgm), gmrgp*vr[0] - trgp*cr,
                    pr*np.power(cr_fan, tgrgm))))
    wr = np.where(p0 > pr, right_shock, right_raref)
    left = 0 <= us
    w = np.where(left, wl, wr)
    vt = [np.where(left, vl[i], vr[i]) for i in (1, 2)]
    nf = np.empty_like(ul)
    nf[0] = w[0]*w[1]
    nf[1] = nf[0]*w[1] + w[2]
    nf[2] = nf[0]*vt[0]
    nf[3] = nf[0]*vt[1]
    nf[4] = (grgm*w[2] + 0.5*w[0]*(w[1]*w[1] + vt[0]*vt[0] +
                                   vt[1]*vt[1]))*w[1]
    return nf


def exact(ul, ur, n, gamma, p_min=5*np.finfo(np.float64).eps):
    """Exact Riemann solver (Toro 2009 ch. 4) in the face frame;
    ``rsolvers/exact.mako`` through ``rsolve1d.mako:1-21``.  p_min: the
    reference's ``[solver-interfaces] p-min`` default, 5 x machine epsilon
    of the working precision (``pyfr/solvers/euler/inters.py:12-26``)."""
    with np.errstate(divide='ignore', invalid='ignore', over='ignore'):
        utl = np.stack([ul[0], *_transform_to(n, ul), ul[4]])
        utr = np.stack([ur[0], *_transform_to(n, ur), ur[4]])
        ntf = _exact_1d(utl, utr, gamma, p_min)
        nf = np.empty_like(ntf)
        nf[0], nf[4] = ntf[0], ntf[4]
        nf[1:4] = _transform_from(n, ntf[1:4])
    return nf


RSOLVERS = {'rusanov': rusanov, 'hllc': hllc, 'hll': hll, 'roe': roe,
            'roem': roem, 'exact': exact}


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
