"""Oracle: FR right-hand side and explicit RK steppers on periodic hex boxes
(test infrastructure).

The RHS follows the reference's own kernel decomposition:

Euler (``pyfr/solvers/baseadvec/system.py:57-136``):
  disu -> intcflux -> tdisf(tflux) -> tdivtpcorf -> tdivtconf -> negdivconf
Navier-Stokes with gradient fusion (``pyfr/solvers/baseadvecdiff/system.py:53-231``):
  disu -> intconu -> tgradpcoru -> tgradcoru -> tdisf_fused -> gradcoru_fpts
  -> intcflux -> tdivtpcorf -> tdivtconf -> negdivconf
"""
import inspect

import numpy as np

from oracle.basis import HexShapeOracle
from oracle.bcs import bccflux_euler, bccflux_ns, bcconu
from oracle.geom import (hex_corners, hex_ploc, linear_smats_djac, pnorm_at,
                         rcpdjac_at, smats_at)
from oracle.mesh import box_hex_verts, pair_faces_bruteforce
from oracle.physics import RSOLVERS, inviscid_flux, pri_to_con, viscous_flux_add


TGV_ICS = {
    'rho': lambda x, y, z: np.ones_like(x),
    'u': lambda x, y, z: np.sin(x)*np.cos(y)*np.cos(z),
    'v': lambda x, y, z: -np.cos(x)*np.sin(y)*np.cos(z),
    'w': lambda x, y, z: np.zeros_like(x),
    'p': lambda x, y, z: 100.0/1.4 + (np.cos(2.0*x) + np.cos(2.0*y)) *
    (np.cos(2.0*z) + 2.0)/16.0,
}


class OracleHexSystem:
    """Periodic box of n^3 linear hexes at FR order ``order``.

    constants: dict with 'gamma' (Euler) and additionally 'mu', 'Pr',
    'ldg-beta', 'ldg-tau' (NS).
    """

    def __init__(self, n, order, system='euler', rsolver='rusanov',
                 constants=None, length=2*np.pi, verts=None,
                 periodic=(True, True, True), bcs=None, curved=None,
                 antialias=(), nq1=None, nf1=None):
        """periodic: per-axis periodicity; the sides of the other axes
        ('x0', 'x1', 'y0', ...) are boundaries, each needing an entry
        ``bcs[side] = (type, vals)`` (``oracle.bcs``; vals may hold
        callables of the flux-point coordinates (x, y, z)).  verts:
        (nspts, 3, E) shape points of any order; curved: per-element mask
        of the elements that take the stored metric at p >= 2 (the rest use
        the linear-element metric of their 8 vertices).  antialias:
        ('flux',) evaluates the flux at the element quadrature points
        (``pyfr/solvers/baseadvec/elements.py:78-94``,
        ``baseadvecdiff/elements.py:103-116``); ('surf-flux',) puts the
        flux points at the nf1 x nf1 face quadrature rule, with M3 fed by
        its L2 projection (``pyfr/shapes.py:103-113``)."""
        self.n, self.order, self.system = n, order, system
        self.rsolver = RSOLVERS[rsolver]
        self.c = dict(constants or {'gamma': 1.4})
        self.length = length

        self.shape = sh = HexShapeOracle(order)
        if 'surf-flux' in antialias:
            sh.set_surface_quadrature(nf1)
        self.nupts, self.nfpts = sh.nupts, sh.nfpts
        self.verts = box_hex_verts(n, length) if verts is None else verts
        self.neles = self.verts.shape[-1]

        # Operators (pyfr/solvers/baseadvec/elements.py:73-100,
        # pyfr/solvers/baseadvecdiff/elements.py:47-55)
        self.M0 = sh.opmat('M0')
        self.M132 = sh.opmat('M1 - M3*M2')
        self.M3 = sh.opmat('M3')
        self.M460 = sh.opmat('M4 - M6*M0')
        self.M6 = sh.opmat('M6')
        self.fluxaa = 'flux' in antialias
        if self.fluxaa:
            sh.set_quadrature(nq1)
            self.nqpts = len(sh.qpts)
            self.M7 = sh.opmat('M7')
            self.M1329 = sh.opmat('(M1 - M3*M2)*M9')

        # Geometry; p < 2 => every element uses the stored ("curved")
        # metric (pyfr/solvers/base/elements.py:224-227)
        self.linear = order >= 2
        cmask = (np.zeros(self.neles, dtype=bool) if curved is None
                 else np.asarray(curved, dtype=bool))
        if not self.linear:
            cmask[:] = True
        sm = smats_at(self.verts, order, sh.upts).transpose(1, 0, 2, 3)
        rd = rcpdjac_at(self.verts, order, sh.upts)
        if cmask.all():
            self.smats_u, self.rcpdjac_l = sm, rd
        else:
            s, d = linear_smats_djac(hex_corners(self.verts), sh.upts)
            self.smats_u = np.where(cmask, sm, s)   # (Nu, 3, 3, E)
            self.rcpdjac_l = np.where(cmask, rd, 1/d)
        self.rcpdjac = rcpdjac_at(self.verts, order, sh.upts)
        if self.fluxaa:
            # S at the quadrature points: stored metric for curved
            # elements, the vertex formula for linear ones
            sq = smats_at(self.verts, order, sh.qpts).transpose(1, 0, 2, 3)
            if not cmask.all():
                sl, _ = linear_smats_djac(hex_corners(self.verts), sh.qpts)
                sq = np.where(cmask, sq, sl)
            self.smats_q = sq
        self.pnorm = pnorm_at(self.verts, order, sh.fpts, sh.norm_fpts)

        # Physical flux-point locations and the interface pairing
        self.plocf = hex_ploc(self.verts, sh.fpts)
        bfaces = {}
        pairs = pair_faces_bruteforce(self.plocf, 6, sh.nfacefpts, length,
                                      periodic, bfaces)
        self._bcdefs = []
        self._ctrls = {}
        for side, faces in sorted(bfaces.items()):
            bctype, vals = (bcs or {})[side]
            bf = np.concatenate([f[2] for f in faces])
            be = np.concatenate([np.full(len(f[2]), f[0]) for f in faces])
            if bctype in ('char-riem-inv-mass-flow',
                          'char-riem-inv-pressure'):
                vals = dict(vals)
                self._ctrls[side] = _Controller(
                    self, bctype, vals.pop('ctrl'), faces)
            self._bcdefs.append((side, bctype, bf, be, vals))
        self._tdep = bool(self._ctrls) or any(
            callable(fn) and _takes_time(fn) for *_, vals in self._bcdefs
            for fn in vals.values())
        self._set_bcs(0.0)
        self.lf = np.concatenate([p[2] for p in pairs])
        self.le = np.concatenate([np.full(len(p[2]), p[0]) for p in pairs])
        self.rf = np.concatenate([p[5] for p in pairs])
        self.re = np.concatenate([np.full(len(p[5]), p[3]) for p in pairs])
        self.nl = self.pnorm[self.lf, self.le].T    # (3, npairs)

    def _set_bcs(self, t):
        """Boundary values at the boundary flux points; callables of
        (x, y, z) or of (x, y, z, t); a controlled boundary's pressure
        t*im + ic."""
        self.bcs = []
        for side, bctype, bf, be, vals in self._bcdefs:
            x, y, z = self.plocf[bf, :, be].T
            v = {}
            if side in self._ctrls:
                c = self._ctrls[side]
                v['p'] = t*c.interp_m + c.interp_c
            for k, fn in vals.items():
                if callable(fn):
                    v[k] = fn(x, y, z, t) if _takes_time(fn) else fn(x, y, z)
                else:
                    v[k] = fn
            self.bcs.append((side, bctype, bf, be, self.pnorm[bf, be].T, v))

    # -- initial conditions (pyfr/solvers/base/elements.py:86-129)
    def ics(self, ics=TGV_ICS, vscale=1.0):
        x, y, z = hex_ploc(self.verts, self.shape.upts).transpose(1, 0, 2)
        pri = [ics['rho'](x, y, z), vscale*ics['u'](x, y, z),
               vscale*ics['v'](x, y, z), vscale*ics['w'](x, y, z),
               ics['p'](x, y, z)]
        con = pri_to_con(pri, self.c['gamma'])
        return np.stack(con, axis=1)               # (Nu, 5, E)

    # -- helpers
    def _mul(self, A, B):
        return (A @ B.reshape(A.shape[1], -1)).reshape(A.shape[0], *B.shape[1:])

    def _tflux_inv(self, u):
        f, _, _ = inviscid_flux(u.transpose(1, 0, 2), self.c['gamma'])
        return f                                    # (3, 5, Nu, E)

    def _transform(self, f, s=None):
        # F[i] = sum_k s[i][k] f[k]; tflux.mako:24-28
        s = self.smats_u if s is None else s        # (Npts, 3, 3, E)
        F = np.zeros((3, s.shape[0], 5, self.neles))
        for i in range(3):
            acc = 0
            for k in range(3):
                acc = acc + s[:, i, k][:, None, :]*f[k].transpose(1, 0, 2)
            F[i] = acc
        return F

    def _intcflux_euler(self, uf):
        ul = uf[self.lf, :, self.le].T
        ur = uf[self.rf, :, self.re].T
        nl = self.nl
        mag = np.sqrt(nl[0]*nl[0] + nl[1]*nl[1] + nl[2]*nl[2])
        nn = (1/mag)*nl
        fn = self.rsolver(ul, ur, nn, self.c['gamma'])
        uf[self.lf, :, self.le] = (mag*fn).T
        uf[self.rf, :, self.re] = (-mag*fn).T

    def _bccflux_euler(self, uf):
        for _, bctype, bf, be, nl, vals in self.bcs:
            ul = uf[bf, :, be].T
            uf[bf, :, be] = bccflux_euler(bctype, self.rsolver, ul, nl, vals,
                                          self.c).T

    # -- Euler RHS
    def rhs_euler(self, u):
        uf = self._mul(self.M0, u)                          # disu
        self._intcflux_euler(uf)                            # intcflux
        self._bccflux_euler(uf)                             # bccflux
        if self.fluxaa:
            uq = self._mul(self.M7, u)                      # qptsu
            F = self._transform(self._tflux_inv(uq), self.smats_q)
            div = self._mul(self.M1329, F.reshape(3*self.nqpts, 5, -1))
        else:
            F = self._transform(self._tflux_inv(u))         # tdisf
            div = self._mul(self.M132, F.reshape(3*self.nupts, 5, -1))
        div = self._mul(self.M3, uf) + div                  # tdivtconf
        return -self.rcpdjac[:, None, :]*div                # negdivconf

    # -- NS RHS
    def rhs_ns(self, u):
        c = self.c
        beta, tau = c['ldg-beta'], c['ldg-tau']
        Nu, E = self.nupts, self.neles

        uf = self._mul(self.M0, u)                          # disu
        ul = uf[self.lf, :, self.le].T
        ur = uf[self.rf, :, self.re].T
        comm = np.zeros_like(uf)                            # intconu
        cu = ur*(0.5 + beta) + ul*(0.5 - beta)
        comm[self.lf, :, self.le] = cu.T
        comm[self.rf, :, self.re] = cu.T
        for _, bctype, bf, be, nl, vals in self.bcs:          # bcconu
            comm[bf, :, be] = bcconu(bctype, uf[bf, :, be].T, nl, vals, c).T

        grad = self._mul(self.M460, u)                      # tgradpcoru
        grad = self._mul(self.M6, comm) + grad              # tgradcoru
        grad = grad.reshape(3, Nu, 5, E)

        # tdisf_fused: transform_grad.mako:3-10 then F = Fi + Fv
        s = self.smats_u
        gp = np.zeros_like(grad)
        for i in range(3):
            acc = 0
            for k in range(3):
                acc = acc + s[:, k, i][:, None, :]*grad[k]
            gp[i] = self.rcpdjac_l[:, None, :]*acc
        if self.fluxaa:
            # gradcoru_qpts, qptsu and the flux at the quadrature points
            uq = self._mul(self.M7, u)
            gq = np.stack([self._mul(self.M7, gp[d]) for d in range(3)])
            f = self._tflux_inv(uq)
            viscous_flux_add(uq.transpose(1, 0, 2), gq.transpose(0, 2, 1, 3),
                             f, c)
            F = self._transform(f, self.smats_q)
        else:
            f = self._tflux_inv(u)                          # (3, 5, Nu, E)
            viscous_flux_add(u.transpose(1, 0, 2), gp.transpose(0, 2, 1, 3),
                             f, c)
            F = self._transform(f)

        gf = np.stack([self._mul(self.M0, gp[d]) for d in range(3)])

        # intcflux (navstokes/kernels/intcflux.mako:10-54)
        gl = gf[:, self.lf, :, self.le].transpose(1, 2, 0)  # (3, 5, npairs)
        gr = gf[:, self.rf, :, self.re].transpose(1, 2, 0)
        nl = self.nl
        mag = np.sqrt(nl[0]*nl[0] + nl[1]*nl[1] + nl[2]*nl[2])
        nn = (1/mag)*nl
        fi = self.rsolver(ul, ur, nn, c['gamma'])
        fvl = viscous_flux_add(ul, gl, np.zeros((3, 5, ul.shape[1])), c)
        fvr = viscous_flux_add(ur, gr, np.zeros((3, 5, ur.shape[1])), c)
        out = np.empty_like(fi)
        for i in range(5):
            nfl = nn[0]*fvl[0, i] + nn[1]*fvl[1, i] + nn[2]*fvl[2, i]
            nfr = nn[0]*fvr[0, i] + nn[1]*fvr[1, i] + nn[2]*fvr[2, i]
            if beta == -0.5:
                fv = nfr
            elif beta == 0.5:
                fv = nfl
            else:
                fv = (0.5 + beta)*nfl + (0.5 - beta)*nfr
            if tau != 0.0:
                fv = fv + tau*(ul[i] - ur[i])
            out[i] = fi[i] + fv
        bcout = [bccflux_ns(bctype, self.rsolver, uf[bf, :, be].T,
                            gf[:, bf, :, be].transpose(1, 2, 0), nl, vals, c)
                 for _, bctype, bf, be, nl, vals in self.bcs]
        uf[self.lf, :, self.le] = (mag*out).T
        uf[self.rf, :, self.re] = (-mag*out).T
        for (_, _, bf, be, _, _), o in zip(self.bcs, bcout):  # bccflux
            uf[bf, :, be] = o.T

        if self.fluxaa:
            div = self._mul(self.M1329, F.reshape(3*self.nqpts, 5, -1))
        else:
            div = self._mul(self.M132, F.reshape(3*Nu, 5, -1))  # tdivtpcorf
        div = self._mul(self.M3, uf) + div                  # tdivtconf
        return -self.rcpdjac[:, None, :]*div                # negdivconf

    def rhs(self, u, t=0.0):
        """The RHS at time t (boundary values that depend on t; the
        controlled boundaries update from u first, as the reference's
        preparefns do before each RHS)."""
        for c in self._ctrls.values():
            c.prepare(u, t)
        if self._tdep:
            self._set_bcs(t)
        return self.rhs_euler(u) if self.system == 'euler' else self.rhs_ns(u)

    @property
    def ndofs(self):
        return self.nupts*5*self.neles


# -- time integration (pyfr/integrators/...)

class _Controller:
    """char-riem-inv-mass-flow / -pressure control
    (``pyfr/solvers/euler/inters.py:108-294``) with the surface integral of
    ``pyfr/quadrules/surface.py:8-55`` on the boundary's faces: a
    Gauss-Legendre rule of degree >= quad-deg, the solution interpolated by
    the nodal basis, the physical (unnormalised) normals of the stored
    metric.  ctrl: target, alpha, eta, nsteps, tstart, qdeg and, for the
    mass flow, the initial pressure p."""

    def __init__(self, orc, bctype, ctrl, faces):
        from oracle.basis import HEX_FACES, _proj, gauss_legendre_quad
        self.mass = bctype.endswith('mass-flow')
        self.target, self.alpha, self.eta = (ctrl['target'], ctrl['alpha'],
                                             ctrl['eta'])
        self.nsteps, self.tstart = ctrl.get('nsteps', 100), ctrl.get(
            'tstart', 0.0)
        self.gamma = orc.c['gamma']
        self.interp_c = ctrl['p'] if self.mass else self.target
        self.interp_m = 0.0
        self.meas_avg, self.tprev, self.counter = 0.0, None, 0
        n1 = -(-(ctrl['qdeg'] + 1)//2)
        qp2, qw = gauss_legendre_quad(n1)
        sh = orc.shape
        nfp = sh.nfacefpts[0]
        self.terms = []
        self.area = 0.0
        for f in faces:
            e, fidx = f[0], int(f[2][0]) // nfp
            proj, norm = HEX_FACES[fidx]
            ppts = _proj(proj, qp2)
            m0 = sh.ubasis.nodal_basis_at(ppts)               # (nq, Nu)
            pn = pnorm_at(orc.verts[..., e:e + 1], orc.order, ppts,
                          np.array([norm]*len(ppts), dtype=float))[:, 0]
            self.terms.append((e, m0, qw, pn))                # pn (nq, 3)
            self.area += (np.sqrt((pn**2).sum(axis=1))*qw).sum()

    def _measure(self, u):
        tot = 0.0
        for e, m0, qw, pn in self.terms:
            uf = m0 @ u[:, :, e]                              # (nq, 5)
            if self.mass:
                tot += np.einsum('qd,qd,q', uf[:, 1:4], pn, qw)
            else:
                rho, E = uf[:, 0], uf[:, 4]
                p = (self.gamma - 1)*(E - 0.5*(uf[:, 1:4]**2).sum(axis=1)/rho)
                tot += (p*np.sqrt((pn**2).sum(axis=1))*qw).sum()
        return tot if self.mass else tot/self.area

    def prepare(self, u, t):
        update = self.counter % self.nsteps == 0
        if (update or not self.tprev) and t >= self.tstart:
            meas = self._measure(u)
            if not self.tprev:
                self.meas_avg, self.tprev = meas, t
            else:
                a = self.alpha
                self.meas_avg = a*meas + (1 - a)*self.meas_avg
                dt, self.tprev = t - self.tprev, t
                p0 = self.interp_m*t + self.interp_c
                if self.mass:
                    p1 = p0 + dt*self.eta*(1 - self.target/self.meas_avg)
                else:
                    p1 = p0 + dt*self.eta*(self.target - self.meas_avg)
                self.interp_m = (p1 - p0)/dt
                self.interp_c = p0 - self.interp_m*t
        self.counter += 1


def _takes_time(fn):
    """A boundary-value callable of (x, y, z, t) rather than (x, y, z)."""
    return len(inspect.signature(fn).parameters) == 4


def clamp_dt(dt_want, t, tcurr, lookahead=10, dtmin=1e-12):
    """``pyfr/integrators/base.py:289-296``."""
    remaining = t - tcurr
    nsteps = -(-remaining // dt_want)
    if nsteps > lookahead:
        return dt_want
    return max(remaining/nsteps, dtmin)


def rk4_step(rhs, u, dt, t=None):
    """One RK4 step with the reference's register choreography and axnpby
    special cases (``steppers.py:67-108``, ``hip/kernels/axnpby.mako:21-43``).
    With t, rhs is called as rhs(u, stage time) (stage times t, t + dt/2,
    t + dt/2, t + dt)."""
    f = rhs if t is None else (lambda v, c: rhs(v, t + c*dt))
    g = (lambda v, c: f(v)) if t is None else f
    r0 = u
    r1 = g(r0, 0.0)
    r2 = (1.0*r0 + (dt/2.0)*r1)
    r2 = g(r2, 0.5)
    r1 = ((dt/6.0)*r1 + 1.0*r0 + (dt/3.0)*r2)
    r2 = ((dt/2.0)*r2 + 1.0*r0)
    r2 = g(r2, 0.5)
    r1 = r1 + ((dt/3.0)*r2)
    r2 = (dt*r2 + 1.0*r0)
    r2 = g(r2, 1.0)
    r1 = r1 + ((dt/6.0)*r2)
    return r1


RK45_A = [970286171893 / 4311952581923, 6584761158862 / 12103376702013,
          2251764453980 / 15575788980749, 26877169314380 / 34165994151039]
RK45_B = [1153189308089 / 22510343858157, 1772645290293 / 4653164025191,
          -1672844663538 / 4480602732383, 2114624349019 / 3568978502595,
          5198255086312 / 14908931495163]


def rk45_step(rhs, u, dt, t=None):
    """Low-storage 2-register RK45 (``steppers.py:111-190``, ``:218-243``;
    ``kernels/rkvdh2.mako:4-31``).  With t, rhs is called as rhs(u, stage
    time), stage i at t + c_i dt with c = [0] + [sum(b[:i]) + a[i]]
    (``steppers.py:130``)."""
    a, b = RK45_A, RK45_B
    c = [0.0] + [sum(b[:i]) + ai for i, ai in enumerate(a)]
    r1, r2 = u.copy(), None
    for i in range(5):
        v = r1 if i == 0 else r2
        r2 = rhs(v) if t is None else rhs(v, t + c[i]*dt)
        if i < 4:
            n1 = r1 + (dt*a[i])*r2
            n2 = r1 + (dt*b[i])*r2
            r1, r2 = n1, n2
        else:
            r1 = r1 + (dt*b[i])*r2
        r1, r2 = r2, r1
    return r2


RK34_A = [11847461282814 / 36547543011857, 3943225443063 / 7078155732230,
          -346793006927 / 4029903576067]
RK34_B = [1017324711453 / 9774461848756, 8237718856693 / 13685301971492,
          57731312506979 / 19404895981398,
          -101169746363290 / 37734290219643]


def rkvdh2_step(rhs, u, dt, a, b, t=None):
    """One step of a 2-register low-storage scheme (``steppers.py:111-190``,
    ``kernels/rkvdh2.mako:4-31``) with coefficients a, b."""
    n = len(b)
    c = [0.0] + [sum(b[:i]) + ai for i, ai in enumerate(a)]
    r1, r2 = u.copy(), None
    for i in range(n):
        v = r1 if i == 0 else r2
        r2 = rhs(v) if t is None else rhs(v, t + c[i]*dt)
        if i < n - 1:
            n1 = r1 + (dt*a[i])*r2
            n2 = r1 + (dt*b[i])*r2
            r1, r2 = n1, n2
        else:
            r1 = r1 + (dt*b[i])*r2
        r1, r2 = r2, r1
    return r2


def rk34_step(rhs, u, dt, t=None):
    """Low-storage RK34 (``steppers.py:193-215``)."""
    return rkvdh2_step(rhs, u, dt, RK34_A, RK34_B, t)


def euler_step(rhs, u, dt, t=None):
    """Forward Euler (``steppers.py:11-23``): u + dt f(u)."""
    f = rhs(u) if t is None else rhs(u, t)
    return 1.0*u + dt*f


def tvdrk3_step(rhs, u, dt, t=None):
    """TVD-RK3 (``steppers.py:26-57``) with the reference's axnpby operand
    order; stage times t, t + dt, t + dt/2."""
    g = (lambda v, c: rhs(v)) if t is None else (lambda v, c: rhs(v, t + c*dt))
    r0 = u
    r2 = g(r0, 0.0)
    r1 = 1.0*r0 + dt*r2
    r2 = g(r1, 1.0)
    r1 = 0.25*r1 + 0.75*r0 + (0.25*dt)*r2
    r2 = g(r1, 0.5)
    r1 = (2.0/3.0)*r1 + (1.0/3.0)*r0 + (2.0/3.0*dt)*r2
    return r1


STEPS = {'euler': euler_step, 'tvd-rk3': tvdrk3_step, 'rk34': rk34_step}


def advance(rhs, u, dt, tend, scheme='rk4', tstart=0.0, timed=False):
    """Fixed-dt ``none`` controller with Kahan-summed time
    (``explicit/controllers.py:48-60``, ``integrators/base.py:282-296``).
    timed: call rhs(u, stage time).  Returns (u, nsteps)."""
    step = {'rk4': rk4_step, 'rk45': rk45_step, **STEPS}[scheme]
    tcurr, comp, nsteps = tstart, 0.0, 0
    while tcurr < tend:
        h = clamp_dt(dt, tend, tcurr)
        u = step(rhs, u, h, tcurr if timed else None)
        y = h - comp
        t = tcurr + y
        comp = (t - tcurr) - y
        tcurr = t
        nsteps += 1
    return u, nsteps
