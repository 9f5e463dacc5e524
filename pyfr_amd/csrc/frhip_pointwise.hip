// Pointwise, interface, RK and reduction kernels of the FR right-hand side.
//
// Element kernels run on a 2-D grid: x covers 64 consecutive elements (one
// wavefront, so each (point, variable) access is one coalesced 512 B FP64
// transaction in the AoSoA layout) and y covers solution points.  Interface
// kernels are 1-D over flux-point pairs and gather through int32 view maps.
#include "frhip_common.h"
#include "frhip_physics.h"

#include <algorithm>
#include <cmath>

namespace frhip {

// S = |J| J^-T rows and |J| of a linear hex at reference point x from its 8
// vertices; pyfr/solvers/baseadvec/kernels/smats.mako:13-25 with the
// derivative of HexShape.interp_expr (pyfr/shapes.py:388-432).
template <typename T>
__device__ __forceinline__ void hex_smats(const T V[8][3], const T x[3],
                                          T s[3][3], T &d)
{
    T j[3][3];
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) j[a][b] = T(0);
#pragma unroll
    for (int v = 0; v < 8; ++v) {
        const T s0 = (v & 1) ? T(1) : T(-1);
        const T s1 = (v & 2) ? T(1) : T(-1);
        const T s2 = (v & 4) ? T(1) : T(-1);
        const T f0 = T(1) + s0*x[0], f1 = T(1) + s1*x[1], f2 = T(1) + s2*x[2];
        const T d0 = s0*f1*f2*T(0.125), d1 = s1*f0*f2*T(0.125),
                d2 = s2*f0*f1*T(0.125);
        for (int b = 0; b < 3; ++b) {
            j[0][b] += d0*V[v][b];
            j[1][b] += d1*V[v][b];
            j[2][b] += d2*V[v][b];
        }
    }
    const int ja[3] = {1, 2, 0}, jb[3] = {2, 0, 1};
    for (int i = 0; i < 3; ++i) {
        const int a = ja[i], b = jb[i];
        s[i][0] = j[a][1]*j[b][2] - j[a][2]*j[b][1];
        s[i][1] = j[a][2]*j[b][0] - j[a][0]*j[b][2];
        s[i][2] = j[a][0]*j[b][1] - j[a][1]*j[b][0];
    }
    d = j[0][0]*s[0][0] + j[0][1]*s[0][1] + j[0][2]*s[0][2];
}

// ---------------------------------------------------------------------------
// Element kernels
// ---------------------------------------------------------------------------

static dim3 elem_grid(int npts, int neles)
{
    const unsigned gx = cdiv(neles, 64);
    unsigned gy = std::max(1u, std::min(cdiv(npts, 4), 2048u/std::max(gx, 1u)));
    return dim3(gx, gy);
}

// Metric terms S at point pt of element e in geometry mode GEO (the
// FRHIP_GEO_* modes): V holds a linear hex's vertices (GEO 1) or the
// element's constant S (GEO 2).
template <typename T, int GEO>
__device__ __forceinline__ void load_geo(const T *__restrict__ geo, long ldg,
                                         int e, T V[8][3])
{
    if constexpr (GEO == 1) {
        for (int v = 0; v < 8; ++v)
            for (int d = 0; d < 3; ++d)
                V[v][d] = geo[v*ldg + soa_ix(e, d, 3)];
    }
    else if constexpr (GEO == 2) {
        for (int i = 0; i < 3; ++i)
            for (int k = 0; k < 3; ++k)
                V[i][k] = geo[(3*i + k)*ldg + e];
    }
}

template <typename T, int GEO>
__device__ __forceinline__ void smats_at(const T *__restrict__ geo, long ldg,
                                         int e, int pt, int npts,
                                         const T *__restrict__ upts,
                                         const T V[8][3], T s[3][3], T &djac)
{
    if constexpr (GEO == 1) {
        const T x[3] = {upts[3*pt], upts[3*pt + 1], upts[3*pt + 2]};
        hex_smats(V, x, s, djac);
    }
    else if constexpr (GEO == 2) {
        for (int i = 0; i < 3; ++i)
            for (int k = 0; k < 3; ++k) s[i][k] = V[i][k];
    }
    else
        for (int i = 0; i < 3; ++i)
            for (int k = 0; k < 3; ++k)
                s[i][k] = geo[((long)i*npts + pt)*ldg + soa_ix(e, k, 3)];
}

// tflux (Euler); pyfr/solvers/euler/kernels/tflux.mako:8-30.  BASE: u
// holds u - u0 and f receives the flux less f(u0) (BaseState).
template <typename T, int GEO, bool BASE>
__global__ __launch_bounds__(256) void k_tflux_euler(
    int npts, int neles, const T *__restrict__ u, long ldu, T *__restrict__ f,
    long ldf, const T *__restrict__ geo, long ldg, const T *__restrict__ upts,
    T gm1, BaseState<T> bs)
{
    const int e = blockIdx.x*64 + threadIdx.x;
    if (e >= neles) return;

    T V[8][3];
    load_geo<T, GEO>(geo, ldg, e, V);

    for (int pt = blockIdx.y*4 + threadIdx.y; pt < npts; pt += 4*gridDim.y) {
        T s[3][3], djac;
        smats_at<T, GEO>(geo, ldg, e, pt, npts, upts, V, s, djac);

        T uu[5];
        for (int v = 0; v < 5; ++v) uu[v] = u[pt*ldu + soa_ix(e, v, 5)];
        T ft[3][5], p, vel[3];
        if constexpr (BASE)
            inviscid_flux_base(uu, ft, p, vel, gm1, bs);
        else
            inviscid_flux(uu, ft, p, vel, gm1);
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 5; ++j)
                f[((long)i*npts + pt)*ldf + soa_ix(e, j, 5)] =
                    s[i][0]*ft[0][j] + s[i][1]*ft[1][j] + s[i][2]*ft[2][j];
    }
}

// tflux (NS, gradient-fused); pyfr/solvers/navstokes/kernels/tflux.mako:13-50.
template <typename T, int GEO, bool BASE>
__global__ __launch_bounds__(256) void k_tflux_ns(
    int npts, int neles, const T *__restrict__ u, long ldu, T *__restrict__ f,
    long ldf, T *__restrict__ gradu, long ldgr, const T *__restrict__ geo,
    long ldg, const T *__restrict__ upts, const T *__restrict__ rcpdjac,
    long ldr, NSConsts<T> c0, const T *__restrict__ mu_e, BaseState<T> bs)
{
    const int e = blockIdx.x*64 + threadIdx.x;
    if (e >= neles) return;
    const NSConsts<T> c = mu_e ? c0.with_mu(mu_e[e]) : c0;

    T V[8][3];
    load_geo<T, GEO>(geo, ldg, e, V);

    for (int pt = blockIdx.y*4 + threadIdx.y; pt < npts; pt += 4*gridDim.y) {
        T s[3][3], rcpd, djac;
        smats_at<T, GEO>(geo, ldg, e, pt, npts, upts, V, s, djac);
        if constexpr (GEO == 1)
            rcpd = T(1)/djac;
        else
            rcpd = rcpdjac[pt*ldr + e];

        T uu[5], g[3][5], tg[3][5];
        for (int v = 0; v < 5; ++v) uu[v] = u[pt*ldu + soa_ix(e, v, 5)];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 5; ++j)
                tg[i][j] = gradu[((long)i*npts + pt)*ldgr + soa_ix(e, j, 5)];

        // transform_grad.mako:3-10
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 5; ++j) {
                g[i][j] = rcpd*(s[0][i]*tg[0][j] + s[1][i]*tg[1][j] +
                                s[2][i]*tg[2][j]);
                gradu[((long)i*npts + pt)*ldgr + soa_ix(e, j, 5)] = g[i][j];
            }

        T ft[3][5], p, vel[3];
        if constexpr (BASE) {
            // the viscous flux of the full state; it vanishes at u0
            T uf[5];
            base_full(uu, uf, bs);
            inviscid_flux_base(uu, ft, p, vel, c.gm1, bs);
            viscous_flux_add(uf, g, ft, c);
        }
        else {
            inviscid_flux(uu, ft, p, vel, c.gm1);
            viscous_flux_add(uu, g, ft, c);
        }
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 5; ++j)
                f[((long)i*npts + pt)*ldf + soa_ix(e, j, 5)] =
                    s[i][0]*ft[0][j] + s[i][1]*ft[1][j] + s[i][2]*ft[2][j];
    }
}

// gradcoru: the corrected reference gradient made physical in place,
// g <- |J|^-1 S^T g; pyfr/solvers/baseadvecdiff/kernels/gradcoru.mako:10-23
// with transform_grad.mako:3-10 (the separate step of the unfused NS path).
template <typename T, int GEO>
__global__ __launch_bounds__(256) void k_gradcoru(
    int npts, int neles, T *__restrict__ gradu, long ldgr,
    const T *__restrict__ geo, long ldg, const T *__restrict__ upts,
    const T *__restrict__ rcpdjac, long ldr)
{
    const int e = blockIdx.x*64 + threadIdx.x;
    if (e >= neles) return;
    T V[8][3];
    load_geo<T, GEO>(geo, ldg, e, V);
    for (int pt = blockIdx.y*4 + threadIdx.y; pt < npts; pt += 4*gridDim.y) {
        T s[3][3], djac;
        smats_at<T, GEO>(geo, ldg, e, pt, npts, upts, V, s, djac);
        const T rcpd = GEO == 1 ? T(1)/djac : rcpdjac[pt*ldr + e];
        T tg[3][5];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 5; ++j)
                tg[i][j] = gradu[((long)i*npts + pt)*ldgr + soa_ix(e, j, 5)];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 5; ++j)
                gradu[((long)i*npts + pt)*ldgr + soa_ix(e, j, 5)] =
                    rcpd*(s[0][i]*tg[0][j] + s[1][i]*tg[1][j] +
                          s[2][i]*tg[2][j]);
    }
}

// tflux (NS) from a physical gradient: f holds the gradient on entry and
// the transformed flux on exit; pyfr/solvers/navstokes/kernels/tflux.mako:
// 13-50 without 'fused' (the anti-aliased path, at quadrature points).
template <typename T, int GEO>
__global__ __launch_bounds__(256) void k_tflux_ns_grad(
    int npts, int neles, const T *__restrict__ u, long ldu,
    T *__restrict__ f, long ldf, const T *__restrict__ geo, long ldg,
    const T *__restrict__ upts, NSConsts<T> c0, const T *__restrict__ mu_e)
{
    const int e = blockIdx.x*64 + threadIdx.x;
    if (e >= neles) return;
    const NSConsts<T> c = mu_e ? c0.with_mu(mu_e[e]) : c0;
    T V[8][3];
    load_geo<T, GEO>(geo, ldg, e, V);
    for (int pt = blockIdx.y*4 + threadIdx.y; pt < npts; pt += 4*gridDim.y) {
        T s[3][3], djac;
        smats_at<T, GEO>(geo, ldg, e, pt, npts, upts, V, s, djac);
        T uu[5], g[3][5];
        for (int v = 0; v < 5; ++v) uu[v] = u[pt*ldu + soa_ix(e, v, 5)];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 5; ++j)
                g[i][j] = f[((long)i*npts + pt)*ldf + soa_ix(e, j, 5)];
        T ft[3][5], p, vel[3];
        inviscid_flux(uu, ft, p, vel, c.gm1);
        viscous_flux_add(uu, g, ft, c);
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 5; ++j)
                f[((long)i*npts + pt)*ldf + soa_ix(e, j, 5)] =
                    s[i][0]*ft[0][j] + s[i][1]*ft[1][j] + s[i][2]*ft[2][j];
    }
}

// negdivconf; pyfr/solvers/baseadvec/kernels/negdivconf.mako:7-22.
template <typename T>
__global__ __launch_bounds__(256) void k_negdivconf(
    int npts, int neles, T *__restrict__ t, long ldt,
    const T *__restrict__ rcpdjac, long ldr)
{
    const int e = blockIdx.x*64 + threadIdx.x;
    if (e >= neles) return;
    for (int pt = blockIdx.y*4 + threadIdx.y; pt < npts; pt += 4*gridDim.y) {
        const T r = rcpdjac[pt*ldr + e];
        for (int v = 0; v < 5; ++v) {
            T *tp = t + pt*ldt + soa_ix(e, v, 5);
            *tp = -r*(*tp) + T(0);
        }
    }
}

// ---------------------------------------------------------------------------
// Interface kernels
// ---------------------------------------------------------------------------

// intcflux (Euler); pyfr/solvers/euler/kernels/intcflux.mako:6-22.
template <typename T, int RS, bool BASE>
__global__ __launch_bounds__(256) void k_intcflux_euler(
    int n, T *__restrict__ base, const int32_t *__restrict__ lmap,
    const int32_t *__restrict__ rmap, const T *__restrict__ nl, long ldn,
    T gamma, T gm1, T qgamma, T pmin, BaseState<T> bs)
{
    const int i = blockIdx.x*blockDim.x + threadIdx.x;
    if (i >= n) return;
    const long li = lmap[i], ri = rmap[i];
    T ul[5], ur[5];
    for (int v = 0; v < 5; ++v) {
        ul[v] = base[li + v*SOA];
        ur[v] = base[ri + v*SOA];
    }
    const T n0 = nl[i], n1 = nl[ldn + i], n2 = nl[2*ldn + i];
    const T mag = sqrt((n0)*(n0) + (n1)*(n1) + (n2)*(n2));
    const T nn[3] = {(T(1)/mag)*n0, (T(1)/mag)*n1, (T(1)/mag)*n2};
    T fn[5];
    if constexpr (BASE) {
        static_assert(RS == FRHIP_RS_RUSANOV, "base-state form: Rusanov");
        rsolve_rusanov_base(ul, ur, nn, fn, gm1, qgamma, bs);
    }
    else
        rsolve<T, RS>(ul, ur, nn, fn, gm1, gamma, qgamma, pmin);
    for (int v = 0; v < 5; ++v) {
        base[li + v*SOA] = mag*fn[v];
        base[ri + v*SOA] = -mag*fn[v];
    }
}

// intconu; pyfr/solvers/navstokes/kernels/intconu.mako:4-19.
template <typename T>
__global__ __launch_bounds__(256) void k_intconu(
    int n, const T *__restrict__ ub, const int32_t *__restrict__ ulin,
    const int32_t *__restrict__ urin, T *__restrict__ cb,
    const int32_t *__restrict__ ulout, const int32_t *__restrict__ urout,
    T beta, T wr, T wl)
{
    const int i = blockIdx.x*blockDim.x + threadIdx.x;
    if (i >= n) return;
    const long li = ulin[i], ri = urin[i], lo = ulout[i], ro = urout[i];
    for (int v = 0; v < 5; ++v) {
        if (beta == T(-0.5))
            cb[ro + v*SOA] = ub[li + v*SOA];
        else if (beta == T(0.5))
            cb[lo + v*SOA] = ub[ri + v*SOA];
        else {
            const T cu = ub[ri + v*SOA]*wr + ub[li + v*SOA]*wl;
            cb[lo + v*SOA] = cu;
            cb[ro + v*SOA] = cu;
        }
    }
}

// intcflux (NS); pyfr/solvers/navstokes/kernels/intcflux.mako:10-54.
template <typename T, int RS, bool BASE>
__global__ __launch_bounds__(256) void k_intcflux_ns(
    int n, T *__restrict__ base, const int32_t *__restrict__ lmap,
    const int32_t *__restrict__ rmap, const T *__restrict__ gb,
    const int32_t *__restrict__ glmap, const int32_t *__restrict__ grmap,
    long gstride, const int32_t *__restrict__ glstr,
    const int32_t *__restrict__ grstr, const T *__restrict__ nl, long ldn,
    NSConsts<T> c0, const T *__restrict__ mu_p, BaseState<T> bs)
{
    const int i = blockIdx.x*blockDim.x + threadIdx.x;
    if (i >= n) return;
    const NSConsts<T> c = mu_p ? c0.with_mu(mu_p[i]) : c0;
    const long li = lmap[i], ri = rmap[i], gli = glmap[i], gri = grmap[i];
    // per-point dimension strides of a view over several element types
    const long gsl = glstr ? glstr[i] : gstride;
    const long gsr = grstr ? grstr[i] : gstride;
    T ul[5], ur[5];
    for (int v = 0; v < 5; ++v) {
        ul[v] = base[li + v*SOA];
        ur[v] = base[ri + v*SOA];
    }
    const T n0 = nl[i], n1 = nl[ldn + i], n2 = nl[2*ldn + i];
    const T mag = sqrt((n0)*(n0) + (n1)*(n1) + (n2)*(n2));
    const T nn[3] = {(T(1)/mag)*n0, (T(1)/mag)*n1, (T(1)/mag)*n2};

    T fi[5];
    // BASE: ul, ur hold u - u0; the viscous fluxes take the full states
    T ulf[5], urf[5];
    if constexpr (BASE) {
        static_assert(RS == FRHIP_RS_RUSANOV, "base-state form: Rusanov");
        rsolve_rusanov_base(ul, ur, nn, fi, c.gm1, c.qgamma, bs);
        base_full(ul, ulf, bs);
        base_full(ur, urf, bs);
    }
    else {
        rsolve<T, RS>(ul, ur, nn, fi, c.gm1, c.gamma, c.qgamma, c.pmin);
        for (int v = 0; v < 5; ++v) {
            ulf[v] = ul[v];
            urf[v] = ur[v];
        }
    }

    T fvl[3][5] = {}, fvr[3][5] = {};
    if (c.beta != T(-0.5)) {
        T gl[3][5];
        for (int d = 0; d < 3; ++d)
            for (int v = 0; v < 5; ++v)
                gl[d][v] = gb[gli + d*gsl + v*SOA];
        viscous_flux_add(ulf, gl, fvl, c);
    }
    if (c.beta != T(0.5)) {
        T gr[3][5];
        for (int d = 0; d < 3; ++d)
            for (int v = 0; v < 5; ++v)
                gr[d][v] = gb[gri + d*gsr + v*SOA];
        viscous_flux_add(urf, gr, fvr, c);
    }

    for (int v = 0; v < 5; ++v) {
        const T nfl = nn[0]*fvl[0][v] + nn[1]*fvl[1][v] + nn[2]*fvl[2][v];
        const T nfr = nn[0]*fvr[0][v] + nn[1]*fvr[1][v] + nn[2]*fvr[2][v];
        T fvcomm;
        if (c.beta == T(-0.5)) fvcomm = nfr;
        else if (c.beta == T(0.5)) fvcomm = nfl;
        else fvcomm = (T(0.5) + c.beta)*nfl + (T(0.5) - c.beta)*nfr;
        if (c.tau != T(0)) fvcomm += c.tau*(ul[v] - ur[v]);
        base[li + v*SOA] = mag*(fi[v] + fvcomm);
        base[ri + v*SOA] = -mag*(fi[v] + fvcomm);
    }
}

// ---------------------------------------------------------------------------
// BLAS extensions
// ---------------------------------------------------------------------------

template <int NV>
struct Ptrs {
    void *p[NV];
};

template <typename T, int NV>
struct Coefs {
    T a[NV];
};

// axnpby; pyfr/backends/hip/kernels/axnpby.mako:4-44.
template <typename T, int NV>
__global__ __launch_bounds__(256) void k_axnpby(long ncol, long ld,
                                                Ptrs<NV> xs, Coefs<T, NV> ca)
{
    const long j = (long)blockIdx.x*blockDim.x + threadIdx.x;
    if (j >= ncol) return;
    const long idx = (long)blockIdx.y*ld + j;
    const T *a = ca.a;
    T *x0 = static_cast<T *>(xs.p[0]);
    T acc;
    if (a[0] == T(0)) {
        acc = (a[1])*(static_cast<const T *>(xs.p[1])[idx]);
        for (int i = 2; i < NV; ++i)
            acc = acc + (a[i])*(static_cast<const T *>(xs.p[i])[idx]);
        x0[idx] = acc;
    }
    else if (a[0] == T(1)) {
        acc = (a[1])*(static_cast<const T *>(xs.p[1])[idx]);
        for (int i = 2; i < NV; ++i)
            acc = acc + (a[i])*(static_cast<const T *>(xs.p[i])[idx]);
        x0[idx] += acc;
    }
    else {
        acc = (a[0])*(x0[idx]);
        for (int i = 1; i < NV; ++i)
            acc = acc + (a[i])*(static_cast<const T *>(xs.p[i])[idx]);
        x0[idx] = acc;
    }
}

// axnpby over a pool of independent solves: coefficient i of element e is
// dt[e]/a[i] when bit i of dtmask is set (dt/6.0 rounds like the host's
// dt/6.0), else a[i].  Same a0 special cases as the uniform kernel for an
// unscaled a0.
template <typename T, int NV>
__global__ __launch_bounds__(256) void k_axnpby_pool(long ncol, long ld,
                                                     Ptrs<NV> xs,
                                                     Coefs<T, NV> ca,
                                                     unsigned dtmask,
                                                     const T *__restrict__ dt,
                                                     int nvar)
{
    const long j = (long)blockIdx.x*blockDim.x + threadIdx.x;
    if (j >= ncol) return;
    const long idx = (long)blockIdx.y*ld + j;
    const long e = (j/(SOA*nvar))*SOA + (j & (SOA - 1));
    const T dte = dt[e];
    T a[NV];
    for (int i = 0; i < NV; ++i)
        a[i] = ((dtmask >> i) & 1) ? dte/ca.a[i] : ca.a[i];
    T *x0 = static_cast<T *>(xs.p[0]);
    const bool a0dt = dtmask & 1;
    T acc = (a[1])*(static_cast<const T *>(xs.p[1])[idx]);
    for (int i = 2; i < NV; ++i)
        acc = acc + (a[i])*(static_cast<const T *>(xs.p[i])[idx]);
    if (!a0dt && ca.a[0] == T(0))
        x0[idx] = acc;
    else if (!a0dt && ca.a[0] == T(1))
        x0[idx] += acc;
    else
        x0[idx] = (a[0])*(x0[idx]) + acc;
}

// rkvdh2 without error estimate; rkvdh2.mako:4-31.
template <typename T>
__global__ __launch_bounds__(256) void k_rkvdh2(long ncol, long ld,
                                                T *__restrict__ r1,
                                                T *__restrict__ r2, T dta,
                                                T dtb, int last)
{
    const long j = (long)blockIdx.x*blockDim.x + threadIdx.x;
    if (j >= ncol) return;
    const long idx = (long)blockIdx.y*ld + j;
    const T t1 = r1[idx], t2 = r2[idx];
    if (last)
        r1[idx] = t1 + dtb*t2;
    else {
        r1[idx] = t1 + dta*t2;
        r2[idx] = t1 + dtb*t2;
    }
}

// Mixed-precision rkvdh2 stage: the accumulated solution stays FP64, the
// RHS k and the next stage input live in one FP32 array (read k, then write
// the next input in place).  acc' = acc + dtb*k; in = (float)(acc + dta*k),
// or in = (float)acc' after the last stage.
__global__ __launch_bounds__(256) void k_rkvdh2_mixed(long ncol, long ld,
                                                      double *__restrict__ acc,
                                                      float *__restrict__ kin,
                                                      double dta, double dtb,
                                                      int last)
{
    const long j = (long)blockIdx.x*blockDim.x + threadIdx.x;
    if (j >= ncol) return;
    const long idx = (long)blockIdx.y*ld + j;
    const double t1 = acc[idx], t2 = (double)kin[idx];
    const double na = t1 + dtb*t2;
    acc[idx] = na;
    kin[idx] = last ? (float)na : (float)(t1 + dta*t2);
}

template <typename D, typename S>
__global__ __launch_bounds__(256) void k_convert(long n, D *__restrict__ d,
                                                 const S *__restrict__ s)
{
    const long i = (long)blockIdx.x*blockDim.x + threadIdx.x;
    if (i < n) d[i] = (D)s[i];
}

template <typename T, int NV>
static void launch_axnpby(int nrow, long ncol, long ld, void *const *xs,
                          const double *a, hipStream_t s)
{
    Ptrs<NV> px;
    Coefs<T, NV> pa;
    for (int i = 0; i < NV; ++i) {
        px.p[i] = xs[i];
        pa.a[i] = (T)a[i];
    }
    dim3 grid(cdiv(ncol, 256), nrow), block(256);
    hipLaunchKernelGGL((k_axnpby<T, NV>), grid, block, 0, s, ncol, ld, px, pa);
}

template <typename T, int NV>
static void launch_axnpby_pool(int nrow, long ncol, long ld, void *const *xs,
                               const double *a, unsigned dtmask,
                               const void *dt, int nvar, hipStream_t s)
{
    Ptrs<NV> px;
    Coefs<T, NV> pa;
    for (int i = 0; i < NV; ++i) {
        px.p[i] = xs[i];
        pa.a[i] = (T)a[i];
    }
    dim3 grid(cdiv(ncol, 256), nrow), block(256);
    hipLaunchKernelGGL((k_axnpby_pool<T, NV>), grid, block, 0, s, ncol, ld, px,
                       pa, dtmask, (const T *)dt, nvar);
}

}  // namespace frhip

using namespace frhip;

extern "C" {

int frhip_tflux_euler(int dtype, int linear, int npts, int neles,
                      const void *u, long ldu, void *f, long ldf,
                      const void *geo, long ldgeo, const void *upts,
                      double gamma, const double *base, void *stream)
{
    FR_API_BEGIN
    FR_DTYPE_OK(dtype);
    FR_CHECK(npts > 0 && neles >= 0, "invalid extents");
    FR_CHECK(linear >= 0 && linear <= 2, "invalid geometry mode");
    if (neles == 0) return 0;
    FR_CHECK(linear != FRHIP_GEO_HEX_LINEAR || upts,
             "linear tflux needs upts");
    dim3 grid = elem_grid(npts, neles), block(64, 4);
    hipStream_t s = as_stream(stream);
#define FR_LAUNCH_TF(T, G)                                                  \
    if (base)                                                               \
        hipLaunchKernelGGL((k_tflux_euler<T, G, true>), grid, block, 0, s,  \
                           npts, neles, (const T *)u, ldu, (T *)f, ldf,     \
                           (const T *)geo, ldgeo, (const T *)upts,          \
                           (T)(gamma - 1), base_state<T>(base, gamma));     \
    else                                                                    \
        hipLaunchKernelGGL((k_tflux_euler<T, G, false>), grid, block, 0, s, \
                           npts, neles, (const T *)u, ldu, (T *)f, ldf,     \
                           (const T *)geo, ldgeo, (const T *)upts,          \
                           (T)(gamma - 1), BaseState<T>{})
    if (dtype == FRHIP_F64) {
        if (linear == 0) FR_LAUNCH_TF(double, 0);
        else if (linear == 1) FR_LAUNCH_TF(double, 1);
        else FR_LAUNCH_TF(double, 2);
    }
    else {
        if (linear == 0) FR_LAUNCH_TF(float, 0);
        else if (linear == 1) FR_LAUNCH_TF(float, 1);
        else FR_LAUNCH_TF(float, 2);
    }
#undef FR_LAUNCH_TF
    FR_LAUNCHED();
    return 0;
    FR_API_END
}

int frhip_gradcoru(int dtype, int linear, int npts, int neles, void *gradu,
                   long ldg, const void *geo, long ldgeo, const void *upts,
                   const void *rcpdjac, long ldr, void *stream)
{
    FR_API_BEGIN
    FR_DTYPE_OK(dtype);
    FR_CHECK(npts > 0 && neles >= 0, "invalid extents");
    FR_CHECK(linear >= 0 && linear <= 2, "invalid geometry mode");
    if (neles == 0) return 0;
    FR_CHECK(linear == FRHIP_GEO_HEX_LINEAR ? upts != nullptr
                                            : rcpdjac != nullptr,
             "missing geometry argument");
    dim3 grid = elem_grid(npts, neles), block(64, 4);
    hipStream_t s = as_stream(stream);
#define FR_LAUNCH_GC(T, G)                                                  \
    hipLaunchKernelGGL((k_gradcoru<T, G>), grid, block, 0, s, npts, neles,  \
                       (T *)gradu, ldg, (const T *)geo, ldgeo,              \
                       (const T *)upts, (const T *)rcpdjac, ldr)
    if (dtype == FRHIP_F64) {
        if (linear == 0) FR_LAUNCH_GC(double, 0);
        else if (linear == 1) FR_LAUNCH_GC(double, 1);
        else FR_LAUNCH_GC(double, 2);
    }
    else {
        if (linear == 0) FR_LAUNCH_GC(float, 0);
        else if (linear == 1) FR_LAUNCH_GC(float, 1);
        else FR_LAUNCH_GC(float, 2);
    }
#undef FR_LAUNCH_GC
    FR_LAUNCHED();
    return 0;
    FR_API_END
}

int frhip_tflux_ns_grad(int dtype, int linear, int npts, int neles,
                        const void *u, long ldu, void *f, long ldf,
                        const void *geo, long ldgeo, const void *upts,
                        const double *consts, const void *mu_e, void *stream)
{
    FR_API_BEGIN
    FR_DTYPE_OK(dtype);
    FR_CHECK(npts > 0 && neles >= 0 && consts, "invalid arguments");
    FR_CHECK(linear >= 0 && linear <= 2, "invalid geometry mode");
    if (neles == 0) return 0;
    FR_CHECK(linear != FRHIP_GEO_HEX_LINEAR || upts,
             "linear tflux needs the point coordinates");
    dim3 grid = elem_grid(npts, neles), block(64, 4);
    hipStream_t s = as_stream(stream);
#define FR_LAUNCH_NG(T, G)                                                  \
    hipLaunchKernelGGL((k_tflux_ns_grad<T, G>), grid, block, 0, s, npts,    \
                       neles, (const T *)u, ldu, (T *)f, ldf,               \
                       (const T *)geo, ldgeo, (const T *)upts, c,           \
                       (const T *)mu_e)
    if (dtype == FRHIP_F64) {
        auto c = ns_consts<double>(consts, -1.0);
        if (linear == 0) FR_LAUNCH_NG(double, 0);
        else if (linear == 1) FR_LAUNCH_NG(double, 1);
        else FR_LAUNCH_NG(double, 2);
    }
    else {
        auto c = ns_consts<float>(consts, -1.0);
        if (linear == 0) FR_LAUNCH_NG(float, 0);
        else if (linear == 1) FR_LAUNCH_NG(float, 1);
        else FR_LAUNCH_NG(float, 2);
    }
#undef FR_LAUNCH_NG
    FR_LAUNCHED();
    return 0;
    FR_API_END
}

int frhip_tflux_ns(int dtype, int linear, int npts, int neles,
                   const void *u, long ldu, void *f, long ldf, void *gradu,
                   long ldg, const void *geo, long ldgeo, const void *upts,
                   const void *rcpdjac, long ldr, const double *consts,
                   const void *mu_e, const double *base, void *stream)
{
    FR_API_BEGIN
    FR_DTYPE_OK(dtype);
    FR_CHECK(npts > 0 && neles >= 0 && consts, "invalid arguments");
    FR_CHECK(linear >= 0 && linear <= 2, "invalid geometry mode");
    if (neles == 0) return 0;
    FR_CHECK(linear == FRHIP_GEO_HEX_LINEAR ? upts != nullptr
                                            : rcpdjac != nullptr,
             "missing geometry argument");
    dim3 grid = elem_grid(npts, neles), block(64, 4);
    hipStream_t s = as_stream(stream);
#define FR_LAUNCH_NS(T, G)                                                  \
    if (base)                                                               \
        hipLaunchKernelGGL((k_tflux_ns<T, G, true>), grid, block, 0, s,     \
                           npts, neles, (const T *)u, ldu, (T *)f, ldf,     \
                           (T *)gradu, ldg, (const T *)geo, ldgeo,          \
                           (const T *)upts, (const T *)rcpdjac, ldr, c,     \
                           (const T *)mu_e, base_state<T>(base, consts[0]));\
    else                                                                    \
        hipLaunchKernelGGL((k_tflux_ns<T, G, false>), grid, block, 0, s,    \
                           npts, neles, (const T *)u, ldu, (T *)f, ldf,     \
                           (T *)gradu, ldg, (const T *)geo, ldgeo,          \
                           (const T *)upts, (const T *)rcpdjac, ldr, c,     \
                           (const T *)mu_e, BaseState<T>{})
    if (dtype == FRHIP_F64) {
        auto c = ns_consts<double>(consts, -1.0);
        if (linear == 0) FR_LAUNCH_NS(double, 0);
        else if (linear == 1) FR_LAUNCH_NS(double, 1);
        else FR_LAUNCH_NS(double, 2);
    }
    else {
        auto c = ns_consts<float>(consts, -1.0);
        if (linear == 0) FR_LAUNCH_NS(float, 0);
        else if (linear == 1) FR_LAUNCH_NS(float, 1);
        else FR_LAUNCH_NS(float, 2);
    }
#undef FR_LAUNCH_NS
    FR_LAUNCHED();
    return 0;
    FR_API_END
}

int frhip_negdivconf(int dtype, int npts, int neles, void *t, long ldt,
                     const void *rcpdjac, long ldr, void *stream)
{
    FR_API_BEGIN
    FR_DTYPE_OK(dtype);
    if (neles == 0) return 0;
    dim3 grid = elem_grid(npts, neles), block(64, 4);
    if (dtype == FRHIP_F64)
        hipLaunchKernelGGL(k_negdivconf<double>, grid, block, 0,
                           as_stream(stream), npts, neles, (double *)t, ldt,
                           (const double *)rcpdjac, ldr);
    else
        hipLaunchKernelGGL(k_negdivconf<float>, grid, block, 0,
                           as_stream(stream), npts, neles, (float *)t, ldt,
                           (const float *)rcpdjac, ldr);
    FR_LAUNCHED();
    return 0;
    FR_API_END
}

int frhip_intcflux_euler(int dtype, int rsolver, int n, void *base,
                         const int32_t *lmap, const int32_t *rmap,
                         const void *nl, long ldn, double gamma,
                         double pmin, const double *bstate, void *stream)
{
    FR_API_BEGIN
    FR_DTYPE_OK(dtype);
    FR_RS_OK(rsolver);
    FR_CHECK(!bstate || rsolver == FRHIP_RS_RUSANOV,
             "the base-state form takes the Rusanov solver");
    if (n == 0) return 0;
    dim3 grid(cdiv(n, 256)), block(256);
    const double pm64 = exact_pmin<double>(pmin);
    const float pm32 = exact_pmin<float>(pmin);
#define FR_L64(RS)                                                          \
    hipLaunchKernelGGL((k_intcflux_euler<double, RS, false>), grid, block,  \
                       0, as_stream(stream), n, (double *)base, lmap, rmap, \
                       (const double *)nl, ldn, gamma, gamma - 1,           \
                       0.25*gamma, pm64, BaseState<double>{})
#define FR_L32(RS)                                                          \
    hipLaunchKernelGGL((k_intcflux_euler<float, RS, false>), grid, block,   \
                       0, as_stream(stream), n, (float *)base, lmap, rmap,  \
                       (const float *)nl, ldn, (float)gamma,                \
                       (float)(gamma - 1), (float)(0.25*gamma), pm32,       \
                       BaseState<float>{})
    if (bstate && dtype == FRHIP_F64)
        hipLaunchKernelGGL(
            (k_intcflux_euler<double, FRHIP_RS_RUSANOV, true>), grid, block,
            0, as_stream(stream), n, (double *)base, lmap, rmap,
            (const double *)nl, ldn, gamma, gamma - 1, 0.25*gamma, pm64,
            base_state<double>(bstate, gamma));
    else if (bstate)
        hipLaunchKernelGGL(
            (k_intcflux_euler<float, FRHIP_RS_RUSANOV, true>), grid, block,
            0, as_stream(stream), n, (float *)base, lmap, rmap,
            (const float *)nl, ldn, (float)gamma, (float)(gamma - 1),
            (float)(0.25*gamma), pm32, base_state<float>(bstate, gamma));
    else if (dtype == FRHIP_F64) {
        FR_RS_DISPATCH(rsolver, FR_L64)
    }
    else {
        FR_RS_DISPATCH(rsolver, FR_L32)
    }
#undef FR_L64
#undef FR_L32
    FR_LAUNCHED();
    return 0;
    FR_API_END
}

int frhip_intconu(int dtype, int n, const void *ubase, const int32_t *ulin,
                  const int32_t *urin, void *cbase, const int32_t *ulout,
                  const int32_t *urout, double beta, void *stream)
{
    FR_API_BEGIN
    FR_DTYPE_OK(dtype);
    if (n == 0) return 0;
    dim3 grid(cdiv(n, 256)), block(256);
    if (dtype == FRHIP_F64)
        hipLaunchKernelGGL(k_intconu<double>, grid, block, 0, as_stream(stream),
                           n, (const double *)ubase, ulin, urin,
                           (double *)cbase, ulout, urout, beta, 0.5 + beta,
                           0.5 - beta);
    else
        hipLaunchKernelGGL(k_intconu<float>, grid, block, 0, as_stream(stream),
                           n, (const float *)ubase, ulin, urin, (float *)cbase,
                           ulout, urout, (float)beta, (float)(0.5 + beta),
                           (float)(0.5 - beta));
    FR_LAUNCHED();
    return 0;
    FR_API_END
}

int frhip_intcflux_ns(int dtype, int rsolver, int n, void *ubase,
                      const int32_t *lmap, const int32_t *rmap,
                      const void *gbase, const int32_t *glmap,
                      const int32_t *grmap, long gstride,
                      const int32_t *glstr, const int32_t *grstr,
                      const void *nl, long ldn, const double *consts,
                      double pmin, const void *mu_p, const double *bstate,
                      void *stream)
{
    FR_API_BEGIN
    FR_DTYPE_OK(dtype);
    FR_RS_OK(rsolver);
    FR_CHECK(consts, "null constants");
    FR_CHECK(!bstate || rsolver == FRHIP_RS_RUSANOV,
             "the base-state form takes the Rusanov solver");
    if (n == 0) return 0;
    dim3 grid(cdiv(n, 256)), block(256);
#define FR_LNS(T, RS, B, BS)                                                \
    hipLaunchKernelGGL((k_intcflux_ns<T, RS, B>), grid, block, 0,           \
                       as_stream(stream), n, (T *)ubase, lmap, rmap,        \
                       (const T *)gbase, glmap, grmap, gstride, glstr,      \
                       grstr, (const T *)nl, ldn, ns_consts<T>(consts, pmin),     \
                       (const T *)mu_p, BS)
#define FR_L64(RS) FR_LNS(double, RS, false, BaseState<double>{})
#define FR_L32(RS) FR_LNS(float, RS, false, BaseState<float>{})
    if (bstate && dtype == FRHIP_F64)
        FR_LNS(double, FRHIP_RS_RUSANOV, true,
               base_state<double>(bstate, consts[0]));
    else if (bstate)
        FR_LNS(float, FRHIP_RS_RUSANOV, true,
               base_state<float>(bstate, consts[0]));
    else if (dtype == FRHIP_F64) {
        FR_RS_DISPATCH(rsolver, FR_L64)
    }
    else {
        FR_RS_DISPATCH(rsolver, FR_L32)
    }
#undef FR_L64
#undef FR_L32
#undef FR_LNS
    FR_LAUNCHED();
    return 0;
    FR_API_END
}

int frhip_axnpby(int dtype, int nrow, long ncol, long ld, int nv,
                 void *const *xs, const double *a, void *stream)
{
    FR_API_BEGIN
    FR_DTYPE_OK(dtype);
    FR_CHECK(nv >= 2 && nv <= 8, "axnpby supports 2..8 operands");
    FR_CHECK(nrow > 0 && nrow < 65536, "invalid row count");
    if (ncol == 0) return 0;
    hipStream_t s = as_stream(stream);
#define FR_AX(NVV)                                                        \
    case NVV:                                                             \
        if (dtype == FRHIP_F64)                                           \
            launch_axnpby<double, NVV>(nrow, ncol, ld, xs, a, s);         \
        else                                                              \
            launch_axnpby<float, NVV>(nrow, ncol, ld, xs, a, s);          \
        break;
    switch (nv) {
        FR_AX(2) FR_AX(3) FR_AX(4) FR_AX(5) FR_AX(6) FR_AX(7) FR_AX(8)
    }
#undef FR_AX
    FR_LAUNCHED();
    return 0;
    FR_API_END
}

int frhip_axnpby_pool(int dtype, int nrow, long ncol, long ld, int nv,
                      void *const *xs, const double *a, unsigned dtmask,
                      const void *dt, int nvar, void *stream)
{
    FR_API_BEGIN
    FR_DTYPE_OK(dtype);
    FR_CHECK(nv >= 2 && nv <= 4, "pool axnpby supports 2..4 operands");
    FR_CHECK(nrow > 0 && nrow < 65536 && nvar > 0 && dt, "invalid arguments");
    FR_CHECK(ncol % (SOA*nvar) == 0, "ncol must cover whole SoA blocks");
    if (ncol == 0) return 0;
    hipStream_t s = as_stream(stream);
#define FR_AXP(NVV)                                                        \
    case NVV:                                                              \
        if (dtype == FRHIP_F64)                                            \
            launch_axnpby_pool<double, NVV>(nrow, ncol, ld, xs, a, dtmask, \
                                            dt, nvar, s);                 \
        else                                                               \
            launch_axnpby_pool<float, NVV>(nrow, ncol, ld, xs, a, dtmask,  \
                                           dt, nvar, s);                  \
        break;
    switch (nv) {
        FR_AXP(2) FR_AXP(3) FR_AXP(4)
    }
#undef FR_AXP
    FR_LAUNCHED();
    return 0;
    FR_API_END
}

int frhip_rkvdh2(int dtype, int nrow, long ncol, long ld, void *r1, void *r2,
                 double dta, double dtb, int last, void *stream)
{
    FR_API_BEGIN
    FR_DTYPE_OK(dtype);
    FR_CHECK(nrow > 0 && nrow < 65536, "invalid row count");
    if (ncol == 0) return 0;
    dim3 grid(cdiv(ncol, 256), nrow), block(256);
    if (dtype == FRHIP_F64)
        hipLaunchKernelGGL(k_rkvdh2<double>, grid, block, 0, as_stream(stream),
                           ncol, ld, (double *)r1, (double *)r2, dta, dtb,
                           last);
    else
        hipLaunchKernelGGL(k_rkvdh2<float>, grid, block, 0, as_stream(stream),
                           ncol, ld, (float *)r1, (float *)r2, (float)dta,
                           (float)dtb, last);
    FR_LAUNCHED();
    return 0;
    FR_API_END
}

int frhip_rkvdh2_mixed(int nrow, long ncol, long ld, void *acc, void *kin,
                       double dta, double dtb, int last, void *stream)
{
    FR_API_BEGIN
    FR_CHECK(nrow > 0 && nrow < 65536, "invalid row count");
    if (ncol == 0) return 0;
    dim3 grid(cdiv(ncol, 256), nrow), block(256);
    hipLaunchKernelGGL(k_rkvdh2_mixed, grid, block, 0, as_stream(stream), ncol,
                       ld, (double *)acc, (float *)kin, dta, dtb, last);
    FR_LAUNCHED();
    return 0;
    FR_API_END
}

int frhip_convert(int dst_dtype, void *dst, int src_dtype, const void *src,
                  long n, void *stream)
{
    FR_API_BEGIN
    FR_DTYPE_OK(dst_dtype);
    FR_DTYPE_OK(src_dtype);
    if (n == 0) return 0;
    dim3 grid(cdiv(n, 256)), block(256);
    hipStream_t s = as_stream(stream);
    if (dst_dtype == FRHIP_F64 && src_dtype == FRHIP_F32)
        hipLaunchKernelGGL((k_convert<double, float>), grid, block, 0, s, n,
                           (double *)dst, (const float *)src);
    else if (dst_dtype == FRHIP_F32 && src_dtype == FRHIP_F64)
        hipLaunchKernelGGL((k_convert<float, double>), grid, block, 0, s, n,
                           (float *)dst, (const double *)src);
    else
        FR_HIP(hipMemcpyAsync(dst, src, n*(dst_dtype == FRHIP_F64 ? 8 : 4),
                              hipMemcpyDeviceToDevice, s));
    FR_LAUNCHED();
    return 0;
    FR_API_END
}

}  // extern "C"
