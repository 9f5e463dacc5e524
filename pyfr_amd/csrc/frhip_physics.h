// Point physics shared by the pointwise, boundary and fused dense kernels
// (internal header): fluxes, Riemann solvers, NS constants.
#pragma once

#include "frhip_common.h"

#include <limits>

namespace frhip {

// Inviscid Euler flux; pyfr/solvers/euler/kernels/flux.mako:3-26.
template <typename T>
__device__ __forceinline__ void inviscid_flux(const T s[5], T f[3][5], T &p,
                                              T v[3], T gm1)
{
    const T invrho = T(1)/s[0], E = s[4];
    const T rhov[3] = {s[1], s[2], s[3]};
    for (int i = 0; i < 3; ++i) v[i] = invrho*rhov[i];
    p = gm1*(E - T(0.5)*invrho*((rhov[0])*(rhov[0]) + (rhov[1])*(rhov[1]) +
                               (rhov[2])*(rhov[2])));
    for (int i = 0; i < 3; ++i) {
        f[i][0] = rhov[i];
        f[i][4] = (E + p)*v[i];
    }
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            f[i][j + 1] = (i == j) ? rhov[i]*v[j] + p : rhov[i]*v[j];
}

// The base-state form: kernels take u - u0 for a constant state
// u0 = (rho0, 0, 0, 0, E0) at rest, and return f(u) - f(u0), where
// f(u0) = (0, p0 I, 0) with p0 = (gamma - 1) E0.  The flux of a constant
// state has zero discrete divergence on free-stream-preserving elements,
// so the right-hand side is unchanged, while the arrays no longer carry
// E0 and p0: at low Mach number the pressure variations then keep their
// digits in FP32.  No counterpart in the reference.
template <typename T>
struct BaseState {
    T rho0, E0, p0, h0;    // h0 = E0 + p0
};

template <typename T>
inline BaseState<T> base_state(const double *b, double gamma)
{
    BaseState<T> s{};
    if (b) {
        const double p0 = (gamma - 1)*b[1];
        s.rho0 = (T)b[0];
        s.E0 = (T)b[1];
        s.p0 = (T)p0;
        s.h0 = (T)(b[1] + p0);
    }
    return s;
}

// u = u0 + s
template <typename T>
__device__ __forceinline__ void base_full(const T s[5], T u[5],
                                          const BaseState<T> &b)
{
    u[0] = b.rho0 + s[0];
    u[1] = s[1];
    u[2] = s[2];
    u[3] = s[3];
    u[4] = b.E0 + s[4];
}

// f(u0 + s) - f(u0) of the inviscid flux; p and v of the full state.  With
// p' = (gamma - 1)(E' - |rho v|^2/(2 rho)) = p - p0 the momentum flux is
// rho v v + p' I and the energy flux (E0 + p0 + E' + p') v.
template <typename T>
__device__ __forceinline__ void inviscid_flux_base(const T s[5], T f[3][5],
                                                   T &p, T v[3], T gm1,
                                                   const BaseState<T> &b)
{
    const T invrho = T(1)/(b.rho0 + s[0]);
    const T rhov[3] = {s[1], s[2], s[3]};
    for (int i = 0; i < 3; ++i) v[i] = invrho*rhov[i];
    const T pp = gm1*(s[4] - T(0.5)*invrho*((rhov[0])*(rhov[0]) +
                                            (rhov[1])*(rhov[1]) +
                                            (rhov[2])*(rhov[2])));
    p = b.p0 + pp;
    const T hh = b.h0 + (s[4] + pp);
    for (int i = 0; i < 3; ++i) {
        f[i][0] = rhov[i];
        f[i][4] = hh*v[i];
    }
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            f[i][j + 1] = (i == j) ? rhov[i]*v[j] + pp : rhov[i]*v[j];
}

// ---------------------------------------------------------------------------
// Point physics (device)
// ---------------------------------------------------------------------------

// Rusanov; pyfr/solvers/euler/kernels/rsolvers/rusanov.mako:4-25.
template <typename T>
__device__ __forceinline__ void rsolve_rusanov(const T ul[5], const T ur[5],
                                               const T n[3], T nf[5], T gm1,
                                               T qgamma)
{
    T fl[3][5], fr[3][5], vl[3], vr[3], pl, pr;
    inviscid_flux(ul, fl, pl, vl, gm1);
    inviscid_flux(ur, fr, pr, vr, gm1);
    const T nv = ((n[0])*(vl[0] + vr[0]) + (n[1])*(vl[1] + vr[1]) +
                  (n[2])*(vl[2] + vr[2]));
    const T a = sqrt(qgamma*(pl + pr)/(ul[0] + ur[0])) + T(0.25)*fabs(nv);
    for (int i = 0; i < 5; ++i)
        nf[i] = T(0.5)*(n[0]*(fl[0][i] + fr[0][i]) + n[1]*(fl[1][i] + fr[1][i])
                        + n[2]*(fl[2][i] + fr[2][i])) + a*(ul[i] - ur[i]);
}

// Rusanov in the base-state form: ul, ur hold u - u0; the result is
// F(u) - f(u0).n (the dissipation a (ul - ur) is unchanged, and the wave
// speed is that of the full states).
template <typename T>
__device__ __forceinline__ void rsolve_rusanov_base(const T ul[5],
                                                    const T ur[5],
                                                    const T n[3], T nf[5],
                                                    T gm1, T qgamma,
                                                    const BaseState<T> &b)
{
    T fl[3][5], fr[3][5], vl[3], vr[3], pl, pr;
    inviscid_flux_base(ul, fl, pl, vl, gm1, b);
    inviscid_flux_base(ur, fr, pr, vr, gm1, b);
    const T nv = ((n[0])*(vl[0] + vr[0]) + (n[1])*(vl[1] + vr[1]) +
                  (n[2])*(vl[2] + vr[2]));
    const T a = sqrt(qgamma*(pl + pr)/((b.rho0 + ul[0]) + (b.rho0 + ur[0])))
              + T(0.25)*fabs(nv);
    for (int i = 0; i < 5; ++i)
        nf[i] = T(0.5)*(n[0]*(fl[0][i] + fr[0][i]) + n[1]*(fl[1][i] + fr[1][i])
                        + n[2]*(fl[2][i] + fr[2][i])) + a*(ul[i] - ur[i]);
}

// HLLC; pyfr/solvers/euler/kernels/rsolvers/hllc.mako:4-77.
template <typename T>
__device__ __forceinline__ void rsolve_hllc(const T ul[5], const T ur[5],
                                            const T n[3], T nf[5], T gm1,
                                            T gamma)
{
    T fl[3][5], fr[3][5], vl[3], vr[3], pl, pr;
    inviscid_flux(ul, fl, pl, vl, gm1);
    inviscid_flux(ur, fr, pr, vr, gm1);
    const T nvl = n[0]*vl[0] + n[1]*vl[1] + n[2]*vl[2];
    const T nvr = n[0]*vr[0] + n[1]*vr[1] + n[2]*vr[2];
    const T al = sqrt(gamma*pl/ul[0]);
    const T ar = sqrt(gamma*pr/ur[0]);
    const T srl = sqrt(ul[0]), srr = sqrt(ur[0]);
    const T nv = (srl*nvl + srr*nvr)/(srl + srr);
    const T H = (srl*(pr + ur[4]) + srr*(pl + ul[4]))/(srl*ur[0] + srr*ul[0]);
    const T inv_rar = T(1)/(srl + srr);
    T va[3];
    for (int i = 0; i < 3; ++i) va[i] = (vl[i]*srl + vr[i]*srr)*inv_rar;
    const T qq = (va[0])*(va[0]) + (va[1])*(va[1]) + (va[2])*(va[2]);
    const T a = sqrt(gm1*(H - T(0.5)*qq));
    const T sl = fmin(nv - a, nvl - al);
    const T sr = fmax(nv + a, nvr + ar);
    const T sstar = (pr - pl + ul[0]*nvl*(sl - nvl) - ur[0]*nvr*(sr - nvr)) /
                    (ul[0]*(sl - nvl) - ur[0]*(sr - nvr));
    const T ul_com = (sl - nvl)/(sl - sstar);
    const T ur_com = (sr - nvr)/(sr - sstar);
    T usl[5], usr[5];
    usl[0] = ul_com*ul[0];
    usr[0] = ur_com*ur[0];
    for (int i = 0; i < 3; ++i) {
        usl[i + 1] = usl[0]*(vl[i] + (sstar - nvl)*n[i]);
        usr[i + 1] = usr[0]*(vr[i] + (sstar - nvr)*n[i]);
    }
    usl[4] = ul_com*(ul[4] + (sstar - nvl)*(ul[0]*sstar + pl/(sl - nvl)));
    usr[4] = ur_com*(ur[4] + (sstar - nvr)*(ur[0]*sstar + pr/(sr - nvr)));
    for (int i = 0; i < 5; ++i) {
        const T nf_fl = n[0]*fl[0][i] + n[1]*fl[1][i] + n[2]*fl[2][i];
        const T nf_fr = n[0]*fr[0][i] + n[1]*fr[1][i] + n[2]*fr[2][i];
        const T nf_fsl = nf_fl + sl*(usl[i] - ul[i]);
        const T nf_fsr = nf_fr + sr*(usr[i] - ur[i]);
        nf[i] = (T(0) <= sl) ? nf_fl
              : (sl <= T(0) && T(0) <= sstar) ? nf_fsl
              : (sstar <= T(0) && T(0) <= sr) ? nf_fsr : nf_fr;
    }
}

// HLL; pyfr/solvers/euler/kernels/rsolvers/hll.mako:4-53, including its
// left wave-speed estimate min(nv - a, nvl - ar).
template <typename T>
__device__ __forceinline__ void rsolve_hll(const T ul[5], const T ur[5],
                                           const T n[3], T nf[5], T gm1,
                                           T gamma)
{
    T fl[3][5], fr[3][5], vl[3], vr[3], pl, pr;
    inviscid_flux(ul, fl, pl, vl, gm1);
    inviscid_flux(ur, fr, pr, vr, gm1);
    const T nvl = n[0]*vl[0] + n[1]*vl[1] + n[2]*vl[2];
    const T nvr = n[0]*vr[0] + n[1]*vr[1] + n[2]*vr[2];
    const T ar = sqrt(gamma*pr/ur[0]);
    const T srl = sqrt(ul[0]), srr = sqrt(ur[0]);
    const T nv = (srl*nvl + srr*nvr)/(srl + srr);
    const T H = (srl*(pr + ur[4]) + srr*(pl + ul[4]))/(srl*ur[0] + srr*ul[0]);
    const T inv_rar = T(1)/(srl + srr);
    T va[3];
    for (int i = 0; i < 3; ++i) va[i] = (vl[i]*srl + vr[i]*srr)*inv_rar;
    const T qq = va[0]*va[0] + va[1]*va[1] + va[2]*va[2];
    const T a = sqrt(gm1*(H - T(0.5)*qq));
    const T sl = fmin(nv - a, nvl - ar);
    const T sr = fmax(nv + a, nvr + ar);
    const T rcpsrsl = T(1)/(sr - sl);
    for (int i = 0; i < 5; ++i) {
        const T nf_fl = n[0]*fl[0][i] + n[1]*fl[1][i] + n[2]*fl[2][i];
        const T nf_fr = n[0]*fr[0][i] + n[1]*fr[1][i] + n[2]*fr[2][i];
        const T nf_sub = (sr*nf_fl - sl*nf_fr + sl*sr*(ur[i] - ul[i]))*rcpsrsl;
        nf[i] = (T(0) <= sl) ? nf_fl : (T(0) >= sr) ? nf_fr : nf_sub;
    }
}

// Rotation of the momentum into the frame whose first axis is n, and back
// (Moler & Hughes); pyfr/solvers/baseadvec/kernels/transform.mako, 3-D.
template <typename T>
__device__ __forceinline__ void transform_to(const T n[3], const T u[3],
                                             T t[3])
{
    if (fabs(n[0]) < T(0.99)) {
        const T h = T(1)/(T(1) + n[0]);
        t[0] = n[0]*u[0] + n[1]*u[1] + n[2]*u[2];
        t[1] = (n[0] + h*n[2]*n[2])*u[1] - n[1]*u[0] - h*n[1]*n[2]*u[2];
        t[2] = (n[0] + h*n[1]*n[1])*u[2] - n[2]*u[0] - h*n[1]*n[2]*u[1];
    }
    else if (fabs(n[1]) < fabs(n[2])) {
        const T h = T(1)/(T(1) - n[1]);
        t[0] = n[0]*u[0] + n[1]*u[1] + n[2]*u[2];
        t[1] = (T(1) - h*n[0]*n[0])*u[0] + n[0]*u[1] - h*n[0]*n[2]*u[2];
        t[2] = n[2]*u[1] - h*n[0]*n[2]*u[0] + (T(1) - h*n[2]*n[2])*u[2];
    }
    else {
        const T h = T(1)/(T(1) - n[2]);
        t[0] = n[0]*u[0] + n[1]*u[1] + n[2]*u[2];
        t[1] = (T(1) - h*n[1]*n[1])*u[1] - h*n[0]*n[1]*u[0] + n[1]*u[2];
        t[2] = (T(1) - h*n[0]*n[0])*u[0] - h*n[0]*n[1]*u[1] + n[0]*u[2];
    }
}

template <typename T>
__device__ __forceinline__ void transform_from(const T n[3], const T t[3],
                                               T u[3])
{
    if (fabs(n[0]) < T(0.99)) {
        const T h = T(1)/(T(1) + n[0]);
        u[0] = n[0]*t[0] - n[1]*t[1] - n[2]*t[2];
        u[1] = n[1]*t[0] + (n[0] + h*n[2]*n[2])*t[1] - h*n[1]*n[2]*t[2];
        u[2] = n[2]*t[0] - h*n[1]*n[2]*t[1] + (n[0] + h*n[1]*n[1])*t[2];
    }
    else if (fabs(n[1]) < fabs(n[2])) {
        const T h = T(1)/(T(1) - n[1]);
        u[0] = n[0]*t[0] + (T(1) - h*n[0]*n[0])*t[1] - h*n[0]*n[2]*t[2];
        u[1] = n[1]*t[0] + n[0]*t[1] + n[2]*t[2];
        u[2] = n[2]*t[0] - h*n[0]*n[2]*t[1] + (T(1) - h*n[2]*n[2])*t[2];
    }
    else {
        const T h = T(1)/(T(1) - n[2]);
        u[0] = n[0]*t[0] - h*n[0]*n[1]*t[1] + (T(1) - h*n[0]*n[0])*t[2];
        u[1] = n[1]*t[0] + (T(1) - h*n[1]*n[1])*t[1] - h*n[0]*n[1]*t[2];
        u[2] = n[2]*t[0] + n[1]*t[1] + n[0]*t[2];
    }
}

// 1-D inviscid flux along the first axis; flux.mako: inviscid_flux_1d.
template <typename T>
__device__ __forceinline__ void inviscid_flux_1d(const T s[5], T f[5], T &p,
                                                 T v[3], T gm1)
{
    const T invrho = T(1)/s[0], E = s[4];
    for (int i = 0; i < 3; ++i) v[i] = invrho*s[i + 1];
    p = gm1*(E - T(0.5)*invrho*(s[1]*s[1] + s[2]*s[2] + s[3]*s[3]));
    f[0] = s[1];
    f[4] = (E + p)*v[0];
    f[1] = s[1]*v[0] + p;
    f[2] = s[1]*v[1];
    f[3] = s[1]*v[2];
}

// Roe with Harten's entropy fix (eps = 0.001), solved in the face frame;
// rsolvers/roe.mako:4-76 through rsolvers/rsolve1d.mako:1-21.
template <typename T>
__device__ __forceinline__ void rsolve_roe(const T ul[5], const T ur[5],
                                           const T n[3], T nf[5], T gm1)
{
    T tl[5], tr[5];
    tl[0] = ul[0]; tr[0] = ur[0]; tl[4] = ul[4]; tr[4] = ur[4];
    transform_to(n, ul + 1, tl + 1);
    transform_to(n, ur + 1, tr + 1);

    T fl[5], fr[5], vl[3], vr[3], pl, pr;
    inviscid_flux_1d(tl, fl, pl, vl, gm1);
    inviscid_flux_1d(tr, fr, pr, vr, gm1);
    const T srl = sqrt(tl[0]), srr = sqrt(tr[0]);
    const T roa = srl*srr;
    const T ha = (srl*(pr + tr[4]) + srr*(pl + tl[4]))/(srl*tr[0] + srr*tl[0]);
    const T inv_rar = T(1)/(srl + srr);
    T va[3], dv[3];
    for (int i = 0; i < 3; ++i) va[i] = (vl[i]*srl + vr[i]*srr)*inv_rar;
    const T qq = va[0]*va[0] + va[1]*va[1] + va[2]*va[2];
    const T a = sqrt(gm1*(ha - T(0.5)*qq));
    T l1 = fabs(va[0] - a), l2 = fabs(va[0]), l3 = fabs(va[0] + a);
    // the reference's literals: 0.5/eps and eps**2 for eps = 0.001
    const T eps = T(0.001), hre = T(500.0), eps2 = T(1.0000000000000002e-06);
    l1 = (l1 < eps) ? hre*(l1*l1 + eps2) : l1;
    l3 = (l3 < eps) ? hre*(l3*l3 + eps2) : l3;
    for (int i = 0; i < 3; ++i) dv[i] = vr[i] - vl[i];
    const T dro = tr[0] - tl[0], dp = pr - pl;
    const T inv_a2 = T(1)/(T(2)*a*a);

    T nt[5];
    T v1 = (dp - roa*a*dv[0])*inv_a2;
    T v2 = dro - dp*T(2)*inv_a2;
    T v3 = (dp + roa*a*dv[0])*inv_a2;
    nt[0] = T(0.5)*(fl[0] + fr[0]) - (l1*v1 + l2*v2 + l3*v3);
    for (int i = 0; i < 3; ++i) {
        if (i == 0) {
            v1 = (dp - roa*a*dv[0])*inv_a2*(va[i] - a);
            v2 = (dro - dp*T(2)*inv_a2)*va[i];
            v3 = (dp + roa*a*dv[0])*inv_a2*(va[i] + a);
        }
        else {
            v1 = (dp - roa*a*dv[0])*inv_a2*va[i];
            v2 = (dro - dp*T(2)*inv_a2)*va[i] + roa*dv[i];
            v3 = (dp + roa*a*dv[0])*inv_a2*va[i];
        }
        nt[i + 1] = T(0.5)*(fl[i + 1] + fr[i + 1]) - (l1*v1 + l2*v2 + l3*v3);
    }
    v1 = (dp - roa*a*dv[0])*inv_a2*(ha - a*va[0]);
    v2 = roa*((va[0]*dv[0] + va[1]*dv[1] + va[2]*dv[2]) - va[0]*dv[0]) +
         (dro - dp*T(2)*inv_a2)*qq*T(0.5);
    v3 = (dp + roa*a*dv[0])*inv_a2*(ha + a*va[0]);
    nt[4] = T(0.5)*(fl[4] + fr[4]) - (l1*v1 + l2*v2 + l3*v3);

    nf[0] = nt[0];
    nf[4] = nt[4];
    transform_from(n, nt + 1, nf + 1);
}

// RoeM (Kim, Kim and Rho, JCP 185(2)); rsolvers/roem.mako:4-86.
template <typename T>
__device__ __forceinline__ void rsolve_roem(const T ul[5], const T ur[5],
                                            const T n[3], T nf[5], T gm1)
{
    T fl[3][5], fr[3][5], vl[3], vr[3], pl, pr;
    inviscid_flux(ul, fl, pl, vl, gm1);
    inviscid_flux(ur, fr, pr, vr, gm1);
    const T hl = (ul[4] + pl)/ul[0], hr = (ur[4] + pr)/ur[0];
    const T contral = n[0]*vl[0] + n[1]*vl[1] + n[2]*vl[2];
    const T contrar = n[0]*vr[0] + n[1]*vr[1] + n[2]*vr[2];
    const T drho = ur[0] - ul[0], dp = pr - pl, dh = hr - hl;
    const T dcontra = contrar - contral;
    T dv[3], va[3];
    for (int i = 0; i < 3; ++i) dv[i] = vr[i] - vl[i];
    const T rrr = sqrt(ur[0]/ul[0]);
    const T ratl = T(1.0)/(T(1.0) + rrr), ratr = rrr*ratl;
    const T ra = rrr*ul[0], ha = hl*ratl + hr*ratr;
    for (int i = 0; i < 3; ++i) va[i] = vl[i]*ratl + vr[i]*ratr;
    const T qq = va[0]*va[0] + va[1]*va[1] + va[2]*va[2];
    const T contraa = n[0]*va[0] + n[1]*va[1] + n[2]*va[2];
    const T aa = sqrt(gm1*(ha - T(0.5)*qq));
    const T rcp_aa = T(1.0)/aa;
    const T abs_ma = fabs(contraa*rcp_aa) + T(1e-15);
    T b1 = fmax(T(0.0), fmax(contraa + aa, contrar + aa));
    T b2 = fmin(T(0.0), fmin(contraa - aa, contral - aa));
    T b1b2 = b1*b2;
    const T rcp_b1_b2 = T(1.0)/(b1 - b2);
    b1 = b1*rcp_b1_b2;
    b2 = b2*rcp_b1_b2;
    b1b2 = b1b2*rcp_b1_b2;
    const T sdst = (pl < pr) ? pl/pr : pr/pl;
    const T h = T(1.0) - sdst;
    const T f = pow(abs_ma, h);
    const T g = f/(T(1.0) + abs_ma);
    T du[5], bdq[5];
    for (int i = 0; i < 4; ++i) du[i] = ur[i] - ul[i];
    du[4] = ur[0]*hr - ul[0]*hl;
    bdq[0] = drho - f*dp*rcp_aa*rcp_aa;
    bdq[4] = bdq[0]*ha + ra*dh;
    for (int i = 0; i < 3; ++i)
        bdq[i + 1] = bdq[0]*va[i] + ra*(dv[i] - n[i]*dcontra);
    for (int i = 0; i < 5; ++i)
        nf[i] = n[0]*(b1*fl[0][i] - b2*fr[0][i]) +
                n[1]*(b1*fl[1][i] - b2*fr[1][i]) +
                n[2]*(b1*fl[2][i] - b2*fr[2][i]) + b1b2*(du[i] - g*bdq[i]);
}

// Exact Riemann solver (Toro 2009 ch. 4, zero covolume) in the face frame;
// rsolvers/exact.mako:14-206 through rsolve1d.mako:1-21.  The gamma ratios
// are formed in double, as the reference forms its template literals; pfloor
// replaces negative Newton iterates of the star pressure: [solver-interfaces]
// p-min, by default 5 x machine epsilon of T (pyfr/solvers/euler/
// inters.py:12-26).
template <typename T>
__device__ __forceinline__ void exact_star_flux(T p, T ps, T rs, T cs, T &f,
                                                T &fd, T gprtg, T trgm,
                                                T trgp, T gmrgp)
{
    if (p <= ps) {
        const T pr = p/ps;
        const T pr_g = pow(pr, -gprtg);
        fd = pr_g/(rs*cs);
        f = trgm*cs*(pr_g*pr - 1);
    }
    else {
        const T pbs_inv = T(1)/(p + gmrgp*ps);
        const T sapb = sqrt(trgp*pbs_inv/rs);
        f = (p - ps)*sapb;
        fd = sapb - T(0.5)*f*pbs_inv;
    }
}

template <typename T>
__device__ __forceinline__ void rsolve_exact(const T ul[5], const T ur[5],
                                             const T n[3], T nf[5], T gm1,
                                             T gamma, T pfloor)
{
    const double g = (double)gamma, gm = g - 1.0;
    const T hgm = T(0.5*gm), grgm = T(g/gm), trgm = T(2/gm);
    const T trgp = T(2/(g + 1)), gmrtg = T(gm/(2*g));
    const T gprtg = T((g + 1)/(2*g)), tgrgm = T((2*g)/gm);
    const T gmrgp = T(gm/(g + 1)), rgamma = T(1/g);
    const T p_min = pfloor;

    T tl[5], tr[5];
    tl[0] = ul[0]; tr[0] = ur[0]; tl[4] = ul[4]; tr[4] = ur[4];
    transform_to(n, ul + 1, tl + 1);
    transform_to(n, ur + 1, tr + 1);

    // primitives (exact.mako:32-45)
    T vl[3], vr[3];
    const T il = T(1)/tl[0], ir = T(1)/tr[0];
    for (int i = 0; i < 3; ++i) { vl[i] = il*tl[i + 1]; vr[i] = ir*tr[i + 1]; }
    const T pl = gm1*(tl[4] - T(0.5)*il*(tl[1]*tl[1] + tl[2]*tl[2] +
                                         tl[3]*tl[3]));
    const T pr = gm1*(tr[4] - T(0.5)*ir*(tr[1]*tr[1] + tr[2]*tr[2] +
                                         tr[3]*tr[3]));
    const T cl = sqrt(gamma*pl*il), cr = sqrt(gamma*pr*ir);
    const T rl = tl[0], rr = tr[0];

    // initial guess
    const T bpv = fmax(T(0), T(0.125)*(vl[0] - vr[0])*(rl + rr)*(cl + cr) +
                             T(0.5)*(pl + pr));
    const T pmin = fmin(pl, pr), pmax = fmax(pl, pr);
    T p0;
    if (pmax <= 2*pmin && pmin <= bpv && bpv <= pmax)
        p0 = bpv;
    else if (bpv < pmin) {
        const T pre = pow(pl/pr, gmrtg);
        const T um = (pre*vl[0]*cr + vr[0]*cl + trgm*(pre - 1)*cl*cr) /
                     (pre*cr + cl);
        const T ptl = 1 - hgm*(um - vl[0])/cl;
        const T ptr = 1 + hgm*(um - vr[0])/cr;
        p0 = T(0.5)*(pl*pow(ptl, tgrgm) + pr*pow(ptr, tgrgm));
    }
    else {
        const T gl = sqrt(trgp/(rl*(gmrgp*pl + bpv)));
        const T gr = sqrt(trgp/(rr*(gmrgp*pr + bpv)));
        p0 = (gl*pl + gr*pr - vr[0] + vl[0])/(gl + gr);
    }

    // three Newton steps on the star pressure
    T fsl, fdl, fsr, fdr;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        exact_star_flux(p0, pl, rl, cl, fsl, fdl, gprtg, trgm, trgp, gmrgp);
        exact_star_flux(p0, pr, rr, cr, fsr, fdr, gprtg, trgm, trgp, gmrgp);
        const T p1 = p0 - (fsl + fsr + vr[0] - vl[0])/(fdl + fdr);
        p0 = (p1 < 0) ? p_min : p1;
    }
    const T us = T(0.5)*(vl[0] + vr[0] + fsr - fsl);

    // sample x/t = 0: w = (rho, u_n, p) and the tangential velocity
    T w0, w1, w4, wt1, wt2;
    if (0 <= us) {
        wt1 = vl[1]; wt2 = vl[2];
        if (p0 <= pl) {
            if (vl[0] >= cl) { w0 = rl; w1 = vl[0]; w4 = pl; }
            else if (us < cl*pow(p0/pl, gmrtg)) {
                w0 = rl*pow(p0/pl, rgamma); w1 = us; w4 = p0;
            }
            else {
                const T c = trgp + gmrgp*vl[0]/cl;
                w0 = rl*pow(c, trgm); w1 = trgp*cl + gmrgp*vl[0];
                w4 = pl*pow(c, tgrgm);
            }
        }
        else {
            const T p0p = p0/pl;
            const T sl = vl[0] - cl*sqrt(gprtg*p0p + gmrtg);
            if (0 <= sl) { w0 = rl; w1 = vl[0]; w4 = pl; }
            else {
                w0 = rl*(p0p + gmrgp)/(p0p*gmrgp + 1); w1 = us; w4 = p0;
            }
        }
    }
    else {
        wt1 = vr[1]; wt2 = vr[2];
        if (p0 > pr) {
            const T p0p = p0/pr;
            const T sr = vr[0] + cr*sqrt(gprtg*p0p + gmrtg);
            if (sr <= 0) { w0 = rr; w1 = vr[0]; w4 = pr; }
            else {
                w0 = rr*(p0p + gmrgp)/(p0p*gmrgp + 1); w1 = us; w4 = p0;
            }
        }
        else if (vr[0] + cr <= 0) { w0 = rr; w1 = vr[0]; w4 = pr; }
        else {
            const T p0p = p0/pr;
            if (us + cr*pow(p0p, gmrtg) >= 0) {
                w0 = rr*pow(p0p, rgamma); w1 = us; w4 = p0;
            }
            else {
                const T c = trgp - gmrgp*vr[0]/cr;
                w0 = rr*pow(c, trgm); w1 = gmrgp*vr[0] - trgp*cr;
                w4 = pr*pow(c, tgrgm);
            }
        }
    }

    T nt[5];
    nt[0] = w0*w1;
    nt[1] = nt[0]*w1 + w4;
    nt[2] = nt[0]*wt1;
    nt[3] = nt[0]*wt2;
    nt[4] = (grgm*w4 + T(0.5)*w0*(w1*w1 + wt1*wt1 + wt2*wt2))*w1;
    nf[0] = nt[0];
    nf[4] = nt[4];
    transform_from(n, nt + 1, nf + 1);
}

template <typename T, int RS>
__device__ __forceinline__ void rsolve(const T ul[5], const T ur[5],
                                       const T n[3], T nf[5], T gm1, T gamma,
                                       T qgamma, T pmin)
{
    if constexpr (RS == FRHIP_RS_HLLC) rsolve_hllc(ul, ur, n, nf, gm1, gamma);
    else if constexpr (RS == FRHIP_RS_HLL) rsolve_hll(ul, ur, n, nf, gm1, gamma);
    else if constexpr (RS == FRHIP_RS_ROE) rsolve_roe(ul, ur, n, nf, gm1);
    else if constexpr (RS == FRHIP_RS_ROEM) rsolve_roem(ul, ur, n, nf, gm1);
    else if constexpr (RS == FRHIP_RS_EXACT)
        rsolve_exact(ul, ur, n, nf, gm1, gamma, pmin);
    else rsolve_rusanov(ul, ur, n, nf, gm1, qgamma);
}

// NS constants prepared on the host in double, then cast.
template <typename T>
struct NSConsts {
    T gamma, gm1, qgamma, mu, gpr, beta, tau, rcpTref, musuth, cpTs, suthfac;
    T pmin;     // the exact solver's pressure floor
    int suth;

    // Copy with the viscosity replaced (per-element / per-pair viscosity)
    __device__ NSConsts with_mu(T m) const
    {
        NSConsts c = *this;
        c.mu = m;
        c.musuth = m*suthfac;
        return c;
    }
};

// Viscous flux (added to f); pyfr/solvers/navstokes/kernels/flux.mako:48-105.
template <typename T>
__device__ __forceinline__ void viscous_flux_add(const T u[5], const T g[3][5],
                                                 T f[3][5],
                                                 const NSConsts<T> &c)
{
    const T rho = u[0], rhou = u[1], rhov = u[2], rhow = u[3], E = u[4];
    const T rcprho = T(1)/rho;
    const T uu = rcprho*rhou, vv = rcprho*rhov, ww = rcprho*rhow;
    const T rho_x = g[0][0], rho_y = g[1][0], rho_z = g[2][0];
    const T u_x = g[0][1] - uu*rho_x, u_y = g[1][1] - uu*rho_y,
            u_z = g[2][1] - uu*rho_z;
    const T v_x = g[0][2] - vv*rho_x, v_y = g[1][2] - vv*rho_y,
            v_z = g[2][2] - vv*rho_z;
    const T w_x = g[0][3] - ww*rho_x, w_y = g[1][3] - ww*rho_y,
            w_z = g[2][3] - ww*rho_z;
    const T E_x = g[0][4], E_y = g[1][4], E_z = g[2][4];

    T mu_c;
    if (c.suth) {
        const T cpT = c.gamma*(rcprho*E - T(0.5)*(uu*uu + vv*vv + ww*ww));
        const T Trat = c.rcpTref*cpT;
        mu_c = c.musuth*Trat*sqrt(Trat)/(cpT + c.cpTs);
    }
    else
        mu_c = c.mu;

    const T T_x = rcprho*(E_x - (rcprho*rho_x*E + uu*u_x + vv*v_x + ww*w_x));
    const T T_y = rcprho*(E_y - (rcprho*rho_y*E + uu*u_y + vv*v_y + ww*w_y));
    const T T_z = rcprho*(E_z - (rcprho*rho_z*E + uu*u_z + vv*v_z + ww*w_z));

    const T third = T(1.0/3.0);
    const T t_xx = -2*mu_c*rcprho*(u_x - third*(u_x + v_y + w_z));
    const T t_yy = -2*mu_c*rcprho*(v_y - third*(u_x + v_y + w_z));
    const T t_zz = -2*mu_c*rcprho*(w_z - third*(u_x + v_y + w_z));
    const T t_xy = -mu_c*rcprho*(v_x + u_y);
    const T t_xz = -mu_c*rcprho*(u_z + w_x);
    const T t_yz = -mu_c*rcprho*(w_y + v_z);

    f[0][1] += t_xx; f[1][1] += t_xy; f[2][1] += t_xz;
    f[0][2] += t_xy; f[1][2] += t_yy; f[2][2] += t_yz;
    f[0][3] += t_xz; f[1][3] += t_yz; f[2][3] += t_zz;
    f[0][4] += uu*t_xx + vv*t_xy + ww*t_xz + -mu_c*c.gpr*T_x;
    f[1][4] += uu*t_xy + vv*t_yy + ww*t_yz + -mu_c*c.gpr*T_y;
    f[2][4] += uu*t_xz + vv*t_yz + ww*t_zz + -mu_c*c.gpr*T_z;
}

// The exact solver's pressure floor: pmin, or where negative the
// default 5 x machine epsilon of T.
template <typename T>
inline T exact_pmin(double pmin)
{
    return pmin >= 0 ? (T)pmin : T(5)*std::numeric_limits<T>::epsilon();
}

template <typename T>
inline NSConsts<T> ns_consts(const double *c, double pmin)
{
    NSConsts<T> k;
    k.pmin = exact_pmin<T>(pmin);
    const double gamma = c[0], mu = c[1], Pr = c[2];
    k.gamma = (T)gamma;
    k.gm1 = (T)(gamma - 1);
    k.qgamma = (T)(0.25*gamma);
    k.mu = (T)mu;
    k.gpr = (T)(gamma/Pr);
    k.beta = (T)c[3];
    k.tau = (T)c[4];
    k.suth = c[5] != 0.0;
    k.rcpTref = (T)(k.suth ? 1/c[6] : 0.0);
    k.musuth = (T)(k.suth ? mu*(c[6] + c[7]) : 0.0);
    k.cpTs = (T)c[7];
    k.suthfac = (T)(k.suth ? c[6] + c[7] : 0.0);
    return k;
}

}  // namespace frhip

#define FR_DTYPE_OK(dt) \
    FR_CHECK((dt) == FRHIP_F64 || (dt) == FRHIP_F32, "invalid dtype")

#define FR_RS_OK(rs)                                                        \
    FR_CHECK((rs) >= FRHIP_RS_RUSANOV && (rs) <= FRHIP_RS_EXACT,            \
             "invalid Riemann solver")

// Expand M(RS) for the run-time solver id (one instantiation per solver)
#define FR_RS_DISPATCH(rs, M)                                               \
    switch (rs) {                                                           \
        case FRHIP_RS_RUSANOV: M(FRHIP_RS_RUSANOV); break;                  \
        case FRHIP_RS_HLLC: M(FRHIP_RS_HLLC); break;                        \
        case FRHIP_RS_HLL: M(FRHIP_RS_HLL); break;                          \
        case FRHIP_RS_ROE: M(FRHIP_RS_ROE); break;                          \
        case FRHIP_RS_ROEM: M(FRHIP_RS_ROEM); break;                        \
        default: M(FRHIP_RS_EXACT); break;                                  \
    }
