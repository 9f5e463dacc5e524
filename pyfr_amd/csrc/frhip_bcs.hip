// Boundary-condition interface kernels: bccflux (Euler, NS) and bcconu.
//
// One thread per boundary flux point, gathering the interior state through
// the int32 view map like the internal-interface kernels.  The ghost
// ("right") state of each boundary type is the reference's bc_rsolve_state
// / bc_ldg_state / bc_ldg_grad_state macro:
//   pyfr/solvers/euler/kernels/bcs/*.mako (sup-in-fa, sup-out-fn,
//   char-riem-inv, slp-adia-wall) and pyfr/solvers/navstokes/kernels/bcs/
//   *.mako (adds no-slp-adia-wall, no-slp-isot-wall, sub-in-frv,
//   sub-in-ftpttang, sub-out-fp).
// The boundary's prescribed values (which the reference splices into the
// kernel source as expressions of ploc and t) arrive as per-point rows bcv,
// evaluated on the host once, or, for values that depend on t, by the
// stack-program kernel below at each RHS; constants derived from gamma and the BC options
// are formed in double on the host, like the reference's template literals.
#include "frhip_common.h"
#include "frhip_physics.h"
#include "frhip_expr.h"

#include <cmath>
#include <type_traits>

namespace frhip {

template <typename T>
struct BCConsts {
    NSConsts<T> ns;
    T gmo, two_gmo, half_gmo, quarter_gmo, rcp_gmo, rcp_gamma;
    T cpTw_g;
    T two_cpTt, pt_mRdcp, Rdcp, rcp_Rdcp, cpTt;
    T vc[3];
};

template <typename T>
static BCConsts<T> bc_consts(const double *c, double pmin)
{
    BCConsts<T> k;
    k.ns = ns_consts<T>(c, pmin);
    const double g = c[0], gmo = g - 1.0;
    const double cpTw = c[8], cpTt = c[9], pt = c[10];
    const double Rdcp = gmo/g;
    k.gmo = (T)gmo;
    k.two_gmo = (T)(2.0/gmo);
    k.half_gmo = (T)(0.5*gmo);
    k.quarter_gmo = (T)(0.25*gmo);
    k.rcp_gmo = (T)(1.0/gmo);
    k.rcp_gamma = (T)(1.0/g);
    k.cpTw_g = (T)(cpTw/g);
    k.two_cpTt = (T)(2.0*cpTt);
    k.pt_mRdcp = (T)(pt > 0 ? std::pow(pt, -Rdcp) : 0.0);
    k.Rdcp = (T)Rdcp;
    k.rcp_Rdcp = (T)(1.0/Rdcp);
    k.cpTt = (T)cpTt;
    for (int i = 0; i < 3; ++i) k.vc[i] = (T)c[11 + i];
    return k;
}

// Rows of bcv per type
__host__ __device__ constexpr int bc_nvals(int bc)
{
    return (bc == FRHIP_BC_SUP_IN_FA || bc == FRHIP_BC_CHAR_RIEM_INV) ? 5
         : bc == FRHIP_BC_NO_SLP_ISOT_WALL ? 3
         : bc == FRHIP_BC_SUB_IN_FRV ? 4
         : bc == FRHIP_BC_SUB_OUT_FP ? 1 : 0;
}

template <typename T>
__device__ __forceinline__ T dot_m(const T u[5])
{
    return u[1]*u[1] + u[2]*u[2] + u[3]*u[3];
}

// bc_rsolve_state; LDG: the same unless the type defines its own
template <typename T, int BC, bool LDG>
__device__ __forceinline__ void bc_state(const T ul[5], const T n[3],
                                         const T *bv, const BCConsts<T> &k,
                                         T ur[5])
{
    if constexpr (BC == FRHIP_BC_SUP_IN_FA) {
        // euler/kernels/bcs/sup-in-fa.mako:3-11
        const T rho = bv[0], p = bv[1];
        ur[0] = rho;
        for (int i = 0; i < 3; ++i) ur[i + 1] = (rho)*(bv[2 + i]);
        ur[4] = p/k.gmo + T(0.5)*(T(1)/ur[0])*dot_m(ur);
    }
    else if constexpr (BC == FRHIP_BC_SUP_OUT_FN) {
        // euler/kernels/bcs/sup-out-fn.mako:3-7
        for (int i = 0; i < 5; ++i) ur[i] = ul[i];
    }
    else if constexpr (BC == FRHIP_BC_CHAR_RIEM_INV) {
        // euler/kernels/bcs/char-riem-inv.mako:6-43
        const T gamma = k.ns.gamma;
        const T rho_e = bv[0], p_e = bv[1];
        const T cs = sqrt(gamma*p_e/rho_e);
        const T s = p_e*pow(rho_e, -gamma);
        const T ratio = cs*k.two_gmo;
        const T inv = T(1)/ul[0];
        const T V_e = bv[2]*n[0] + bv[3]*n[1] + bv[4]*n[2];
        const T V_i = inv*(ul[1]*n[0] + ul[2]*n[1] + ul[3]*n[2]);
        const T p_i = k.gmo*ul[4] - k.half_gmo*inv*dot_m(ul);
        const T c_i = sqrt(gamma*p_i*inv);
        const T R_e = (fabs(V_e) >= cs && V_i >= 0) ? V_i - c_i*k.two_gmo
                                                     : V_e - ratio;
        const T R_i = (fabs(V_e) >= cs && V_i < 0) ? V_e + ratio
                                                    : V_i + c_i*k.two_gmo;
        const T V_b = T(0.5)*(R_e + R_i);
        const T c_b = k.quarter_gmo*(R_i - R_e);
        const T rho_b = (V_i < 0)
            ? pow((T(1)/(gamma*s))*c_b*c_b, k.rcp_gmo)
            : ul[0]*pow(ul[0]*c_b*c_b/(gamma*p_i), k.rcp_gmo);
        const T p_b = k.rcp_gamma*rho_b*c_b*c_b;
        ur[0] = rho_b;
        for (int i = 0; i < 3; ++i)
            ur[i + 1] = (V_i >= 0) ? rho_b*(ul[i + 1]*inv + (V_b - V_i)*n[i])
                                   : rho_b*(bv[2 + i] + (V_b - V_e)*n[i]);
        ur[4] = p_b*k.rcp_gmo + T(0.5)*(T(1)/ur[0])*dot_m(ur);
    }
    else if constexpr (BC == FRHIP_BC_SLP_ADIA_WALL) {
        // euler/kernels/bcs/slp-adia-wall.mako:3-10
        const T nor = ul[1]*n[0] + ul[2]*n[1] + ul[3]*n[2];
        ur[0] = ul[0];
        for (int i = 0; i < 3; ++i) ur[i + 1] = ul[i + 1] - 2*nor*n[i];
        ur[4] = ul[4];
    }
    else if constexpr (BC == FRHIP_BC_NO_SLP_ADIA_WALL) {
        // navstokes/kernels/bcs/no-slp-adia-wall.mako:4-20
        ur[0] = ul[0];
        for (int i = 0; i < 3; ++i) ur[i + 1] = LDG ? T(0) : -ul[i + 1];
        ur[4] = LDG ? ul[4] - (T(0.5)/ul[0])*dot_m(ul) : ul[4];
    }
    else if constexpr (BC == FRHIP_BC_NO_SLP_ISOT_WALL) {
        // navstokes/kernels/bcs/no-slp-isot-wall.mako:4-20
        ur[0] = ul[0];
        for (int i = 0; i < 3; ++i)
            ur[i + 1] = LDG ? bv[i]*ul[0] : -ul[i + 1] + 2*bv[i]*ul[0];
        ur[4] = k.cpTw_g*ur[0] + T(0.5)*(T(1)/ur[0])*dot_m(ur);
    }
    else if constexpr (BC == FRHIP_BC_SUB_IN_FRV) {
        // navstokes/kernels/bcs/sub-in-frv.mako:4-13
        const T rho = bv[0];
        ur[0] = rho;
        for (int i = 0; i < 3; ++i) ur[i + 1] = (rho)*(bv[1 + i]);
        ur[4] = ul[4] - T(0.5)*(T(1)/ul[0])*dot_m(ul)
              + T(0.5)*(T(1)/ur[0])*dot_m(ur);
    }
    else if constexpr (BC == FRHIP_BC_SUB_IN_FTPTTANG) {
        // navstokes/kernels/bcs/sub-in-ftpttang.mako:4-18
        const T pl = k.gmo*(ul[4] - (T(0.5)/ul[0])*dot_m(ul));
        T udotu = k.two_cpTt*(T(1) - k.pt_mRdcp*pow(pl, k.Rdcp));
        udotu = fmax(T(0), udotu);
        ur[0] = k.rcp_Rdcp*pl/(k.cpTt - T(0.5)*udotu);
        for (int i = 0; i < 3; ++i) ur[i + 1] = k.vc[i]*ur[0]*sqrt(udotu);
        ur[4] = k.rcp_gmo*pl + T(0.5)*ur[0]*udotu;
    }
    else {
        // navstokes/kernels/bcs/sub-out-fp.mako:4-10
        for (int i = 0; i < 4; ++i) ur[i] = ul[i];
        ur[4] = bv[0]/k.gmo + T(0.5)*(T(1)/ul[0])*dot_m(ul);
    }
}

// bc_ldg_grad_state: zero / copy (navstokes/kernels/bcs/common.mako:3-15)
// or the adiabatic-wall temperature-gradient correction
// (no-slp-adia-wall.mako:22-80, 3-D branch)
template <typename T, int BC>
__device__ __forceinline__ void bc_grad_state(const T ur[5], const T n[3],
                                              const T gl[3][5], T gr[3][5])
{
    constexpr bool zero = BC == FRHIP_BC_SUP_IN_FA ||
                          BC == FRHIP_BC_CHAR_RIEM_INV ||
                          BC == FRHIP_BC_SUB_IN_FRV ||
                          BC == FRHIP_BC_SUB_OUT_FP;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 5; ++j) gr[i][j] = zero ? T(0) : gl[i][j];
    if constexpr (BC == FRHIP_BC_NO_SLP_ADIA_WALL) {
        const T rcprho = T(1)/ur[0];
        const T u = rcprho*ur[1], v = rcprho*ur[2], w = rcprho*ur[3];
        T Tl[3];
        for (int d = 0; d < 3; ++d) {
            const T u_d = gl[d][1] - u*gl[d][0];
            const T v_d = gl[d][2] - v*gl[d][0];
            const T w_d = gl[d][3] - w*gl[d][0];
            Tl[d] = gl[d][4] - (rcprho*gl[d][0]*ur[4] + u*u_d + v*v_d +
                                w*w_d);
        }
        for (int i = 0; i < 3; ++i)
            gr[i][4] -= n[i]*n[0]*Tl[0] + n[i]*n[1]*Tl[1] + n[i]*n[2]*Tl[2];
    }
}

template <typename T>
__device__ __forceinline__ T unit_normal(const T *__restrict__ nl, long ldn,
                                         int i, T nn[3])
{
    const T n0 = nl[i], n1 = nl[ldn + i], n2 = nl[2*ldn + i];
    const T mag = sqrt((n0)*(n0) + (n1)*(n1) + (n2)*(n2));
    nn[0] = (T(1)/mag)*n0; nn[1] = (T(1)/mag)*n1; nn[2] = (T(1)/mag)*n2;
    return mag;
}

template <typename T, int BC>
__device__ __forceinline__ void load_bv(const T *__restrict__ bcv, long ldb,
                                        int i, T bv[5])
{
    for (int r = 0; r < bc_nvals(BC); ++r) bv[r] = bcv[r*ldb + i];
}

// bccflux (Euler); pyfr/solvers/euler/kernels/bccflux.mako:7-25
template <typename T, int RS, int BC>
__global__ __launch_bounds__(256) void k_bccflux_euler(
    int n, T *__restrict__ base, const int32_t *__restrict__ lmap,
    const T *__restrict__ nl, long ldn, const T *__restrict__ bcv, long ldb,
    BCConsts<T> k)
{
    const int i = blockIdx.x*blockDim.x + threadIdx.x;
    if (i >= n) return;
    const long li = lmap[i];
    T ul[5], ur[5], nn[3], bv[5], fn[5];
    for (int v = 0; v < 5; ++v) ul[v] = base[li + v*SOA];
    const T mag = unit_normal(nl, ldn, i, nn);
    load_bv<T, BC>(bcv, ldb, i, bv);
    bc_state<T, BC, false>(ul, nn, bv, k, ur);
    rsolve<T, RS>(ul, ur, nn, fn, k.ns.gm1, k.ns.gamma, k.ns.qgamma,
                  k.ns.pmin);
    for (int v = 0; v < 5; ++v) base[li + v*SOA] = mag*fn[v];
}

// bcconu; pyfr/solvers/navstokes/kernels/bcconu.mako:6-15
template <typename T, int BC>
__global__ __launch_bounds__(256) void k_bcconu(
    int n, const T *__restrict__ ub, const int32_t *__restrict__ ulin,
    T *__restrict__ cb, const int32_t *__restrict__ ulout,
    const T *__restrict__ nl, long ldn, const T *__restrict__ bcv, long ldb,
    BCConsts<T> k)
{
    const int i = blockIdx.x*blockDim.x + threadIdx.x;
    if (i >= n) return;
    const long li = ulin[i], lo = ulout[i];
    T ul[5], ur[5], nn[3], bv[5];
    for (int v = 0; v < 5; ++v) ul[v] = ub[li + v*SOA];
    unit_normal(nl, ldn, i, nn);
    load_bv<T, BC>(bcv, ldb, i, bv);
    bc_state<T, BC, true>(ul, nn, bv, k, ur);
    for (int v = 0; v < 5; ++v) cb[lo + v*SOA] = ur[v];
}

// bccflux (NS); pyfr/solvers/navstokes/kernels/bccflux.mako:10-20 with the
// type's common-flux state: slp-adia-wall.mako:21-37 (inviscid only),
// ghost-imperm.mako (no-slip walls) or ghost.mako (the rest)
template <typename T, int RS, int BC>
__global__ __launch_bounds__(256) void k_bccflux_ns(
    int n, T *__restrict__ base, const int32_t *__restrict__ lmap,
    const T *__restrict__ gb, const int32_t *__restrict__ glmap,
    long gstride, const int32_t *__restrict__ glstr,
    const T *__restrict__ nl, long ldn, const T *__restrict__ bcv, long ldb,
    BCConsts<T> k)
{
    const int i = blockIdx.x*blockDim.x + threadIdx.x;
    if (i >= n) return;
    const long li = lmap[i];
    T ul[5], ur[5], nn[3], bv[5], fi[5];
    for (int v = 0; v < 5; ++v) ul[v] = base[li + v*SOA];
    const T mag = unit_normal(nl, ldn, i, nn);
    load_bv<T, BC>(bcv, ldb, i, bv);

    if constexpr (BC == FRHIP_BC_SLP_ADIA_WALL) {
        bc_state<T, BC, false>(ul, nn, bv, k, ur);
        rsolve<T, RS>(ul, ur, nn, fi, k.ns.gm1, k.ns.gamma, k.ns.qgamma,
                  k.ns.pmin);
        for (int v = 0; v < 5; ++v) base[li + v*SOA] = mag*fi[v];
    }
    else {
        constexpr bool imperm = BC == FRHIP_BC_NO_SLP_ADIA_WALL ||
                                BC == FRHIP_BC_NO_SLP_ISOT_WALL;
        const long gli = glmap[i], gs = glstr ? glstr[i] : gstride;
        T gl[3][5], gr[3][5], url[5], fvr[3][5] = {};
        for (int d = 0; d < 3; ++d)
            for (int v = 0; v < 5; ++v)
                gl[d][v] = gb[gli + d*gs + v*SOA];
        bc_state<T, BC, true>(ul, nn, bv, k, url);
        bc_grad_state<T, BC>(imperm ? url : ul, nn, gl, gr);
        viscous_flux_add(url, gr, fvr, k.ns);
        bc_state<T, BC, false>(ul, nn, bv, k, ur);
        rsolve<T, RS>(ul, ur, nn, fi, k.ns.gm1, k.ns.gamma, k.ns.qgamma,
                  k.ns.pmin);
        for (int v = 0; v < 5; ++v) {
            T fvcomm = nn[0]*fvr[0][v] + nn[1]*fvr[1][v] + nn[2]*fvr[2][v];
            if (!imperm && k.ns.tau != T(0))
                fvcomm += k.ns.tau*(ul[v] - ur[v]);
            base[li + v*SOA] = mag*(fi[v] + fvcomm);
        }
    }
}

// Compile-time dispatch of a run-time id over [0, N)
template <int I, int N, typename F>
static void dispatch(int id, F &&f)
{
    if constexpr (I < N) {
        if (id == I) f(std::integral_constant<int, I>{});
        else dispatch<I + 1, N>(id, f);
    }
}


// ---- boundary values as expressions of (x, y, z, t) ------------------------
// The reference splices each value expression into the kernel source with
// t a run-time argument (pyfr/solvers/base/inters.py:54-71 _exp_opts,
// pyfr/solvers/euler/inters.py:57-102).  Here an expression is a stack
// program of (op, arg) pairs evaluated in double at each boundary point; t
// lives in a device scalar the host sets on the solver stream before each
// RHS, so captured graphs stay valid.
template <typename T>
__global__ __launch_bounds__(256) void k_bc_eval(
    int n, int nrows, const int32_t *__restrict__ code,
    const int32_t *__restrict__ offs, const double *__restrict__ pool,
    const T *__restrict__ ploc, long ldp, const double *__restrict__ tp,
    T *__restrict__ out, long ldo)
{
    const int i = blockIdx.x*blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double var[4] = {double(ploc[i]), double(ploc[ldp + i]),
                           double(ploc[2*ldp + i]), *tp};
    for (int r = 0; r < nrows; ++r)
        out[r*ldo + i] = T(expr_eval(code, offs[r], offs[r + 1], pool, var,
                                     4));
}

__global__ void k_set_scalar(double *p, double v)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) *p = v;
}

struct BCProgram {
    int nrows, ncode, npool;
    int32_t *code, *offs;   // device
    double *pool, *t;       // device
};

}  // namespace frhip

using namespace frhip;

extern "C" {

int frhip_bccflux_euler(int dtype, int bctype, int rsolver, int n,
                        void *base, const int32_t *lmap, const void *nl,
                        long ldn, const void *bcv, long ldb,
                        const double *consts, double pmin, void *stream)
{
    FR_API_BEGIN
    FR_DTYPE_OK(dtype);
    FR_RS_OK(rsolver);
    FR_CHECK(bctype >= 0 && bctype <= FRHIP_BC_SLP_ADIA_WALL,
             "boundary type not available for the Euler equations");
    FR_CHECK(consts, "missing constants");
    FR_CHECK(bc_nvals(bctype) == 0 || bcv, "missing boundary values");
    if (n == 0) return 0;
    const dim3 grid(cdiv(n, 256)), block(256);
    auto go = [&](auto tag) {
        using T = decltype(tag);
        const BCConsts<T> k = bc_consts<T>(consts, pmin);
        dispatch<0, FRHIP_RS_EXACT + 1>(rsolver, [&](auto rs) {
            dispatch<0, FRHIP_BC_SLP_ADIA_WALL + 1>(bctype, [&](auto bc) {
                hipLaunchKernelGGL(
                    (k_bccflux_euler<T, decltype(rs)::value,
                                     decltype(bc)::value>),
                    grid, block, 0, as_stream(stream), n, (T *)base, lmap,
                    (const T *)nl, ldn, (const T *)bcv, ldb, k);
            });
        });
    };
    if (dtype == FRHIP_F64) go(double());
    else go(float());
    FR_LAUNCHED();
    return 0;
    FR_API_END
}

int frhip_bcconu(int dtype, int bctype, int n, const void *ub,
                 const int32_t *ulin, void *cb, const int32_t *ulout,
                 const void *nl, long ldn, const void *bcv, long ldb,
                 const double *consts, void *stream)
{
    FR_API_BEGIN
    FR_DTYPE_OK(dtype);
    FR_CHECK(bctype >= 0 && bctype <= FRHIP_BC_SUB_OUT_FP,
             "invalid boundary type");
    FR_CHECK(consts, "missing constants");
    FR_CHECK(bc_nvals(bctype) == 0 || bcv, "missing boundary values");
    if (n == 0) return 0;
    const dim3 grid(cdiv(n, 256)), block(256);
    auto go = [&](auto tag) {
        using T = decltype(tag);
        const BCConsts<T> k = bc_consts<T>(consts, -1.0);
        dispatch<0, FRHIP_BC_SUB_OUT_FP + 1>(bctype, [&](auto bc) {
            hipLaunchKernelGGL((k_bcconu<T, decltype(bc)::value>), grid,
                               block, 0, as_stream(stream), n, (const T *)ub,
                               ulin, (T *)cb, ulout, (const T *)nl, ldn,
                               (const T *)bcv, ldb, k);
        });
    };
    if (dtype == FRHIP_F64) go(double());
    else go(float());
    FR_LAUNCHED();
    return 0;
    FR_API_END
}

int frhip_bccflux_ns(int dtype, int bctype, int rsolver, int n, void *base,
                     const int32_t *lmap, const void *gb,
                     const int32_t *glmap, long gstride,
                     const int32_t *glstr, const void *nl, long ldn,
                     const void *bcv, long ldb, const double *consts,
                     double pmin, void *stream)
{
    FR_API_BEGIN
    FR_DTYPE_OK(dtype);
    FR_RS_OK(rsolver);
    FR_CHECK(bctype >= 0 && bctype <= FRHIP_BC_SUB_OUT_FP,
             "invalid boundary type");
    FR_CHECK(consts, "missing constants");
    FR_CHECK(bc_nvals(bctype) == 0 || bcv, "missing boundary values");
    FR_CHECK(bctype == FRHIP_BC_SLP_ADIA_WALL || (gb && glmap),
             "missing gradient view");
    if (n == 0) return 0;
    const dim3 grid(cdiv(n, 256)), block(256);
    auto go = [&](auto tag) {
        using T = decltype(tag);
        const BCConsts<T> k = bc_consts<T>(consts, pmin);
        dispatch<0, FRHIP_RS_EXACT + 1>(rsolver, [&](auto rs) {
            dispatch<0, FRHIP_BC_SUB_OUT_FP + 1>(bctype, [&](auto bc) {
                hipLaunchKernelGGL(
                    (k_bccflux_ns<T, decltype(rs)::value,
                                  decltype(bc)::value>),
                    grid, block, 0, as_stream(stream), n, (T *)base, lmap,
                    (const T *)gb, glmap, gstride, glstr, (const T *)nl, ldn,
                    (const T *)bcv, ldb, k);
            });
        });
    };
    if (dtype == FRHIP_F64) go(double());
    else go(float());
    FR_LAUNCHED();
    return 0;
    FR_API_END
}

int frhip_bc_program_create(void **prog, int nrows, const int32_t *code,
                            const int32_t *offs, const double *pool,
                            int npool)
{
    FR_API_BEGIN
    FR_CHECK(prog && code && offs && nrows > 0 && nrows <= 8,
             "invalid program arguments");
    FR_CHECK(npool == 0 || pool, "missing constant pool");
    {
        const char *err = expr_validate(nrows, code, offs, npool, 4);
        FR_CHECK(!err, err ? err : "");
    }
    auto *p = new BCProgram{nrows, offs[nrows], npool, nullptr, nullptr,
                            nullptr, nullptr};
    auto fail = [&]() {
        (void)hipFree(p->code);
        (void)hipFree(p->offs);
        (void)hipFree(p->pool);
        (void)hipFree(p->t);
        delete p;
    };
    const size_t nc = (size_t)2*p->ncode*sizeof(int32_t);
    if (hipMalloc(&p->code, nc) != hipSuccess ||
        hipMalloc(&p->offs, (nrows + 1)*sizeof(int32_t)) != hipSuccess ||
        hipMalloc(&p->pool, (npool > 0 ? npool : 1)*sizeof(double)) !=
            hipSuccess ||
        hipMalloc(&p->t, sizeof(double)) != hipSuccess ||
        hipMemcpy(p->code, code, nc, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(p->offs, offs, (nrows + 1)*sizeof(int32_t),
                  hipMemcpyHostToDevice) != hipSuccess ||
        (npool > 0 && hipMemcpy(p->pool, pool, npool*sizeof(double),
                                hipMemcpyHostToDevice) != hipSuccess)) {
        fail();
        FR_CHECK(false, "device allocation for a boundary program failed");
    }
    const double zero = 0.0;
    if (hipMemcpy(p->t, &zero, sizeof(double), hipMemcpyHostToDevice) !=
            hipSuccess) {
        fail();
        FR_CHECK(false, "device copy for a boundary program failed");
    }
    *prog = p;
    return 0;
    FR_API_END
}

int frhip_bc_program_destroy(void *prog)
{
    FR_API_BEGIN
    auto *p = static_cast<BCProgram *>(prog);
    if (!p) return 0;
    (void)hipFree(p->code);
    (void)hipFree(p->offs);
    (void)hipFree(p->pool);
    (void)hipFree(p->t);
    delete p;
    return 0;
    FR_API_END
}

int frhip_bc_program_set_time(void *prog, double t, void *stream)
{
    FR_API_BEGIN
    auto *p = static_cast<BCProgram *>(prog);
    FR_CHECK(p, "null program");
    hipLaunchKernelGGL(k_set_scalar, dim3(1), dim3(64), 0, as_stream(stream),
                       p->t, t);
    FR_LAUNCHED();
    return 0;
    FR_API_END
}

int frhip_bc_program_set_const(void *prog, int idx, double v, void *stream)
{
    FR_API_BEGIN
    auto *p = static_cast<BCProgram *>(prog);
    FR_CHECK(p, "null program");
    FR_CHECK(idx >= 0 && idx < p->npool, "constant index out of range");
    hipLaunchKernelGGL(k_set_scalar, dim3(1), dim3(64), 0, as_stream(stream),
                       p->pool + idx, v);
    FR_LAUNCHED();
    return 0;
    FR_API_END
}

int frhip_bc_eval(void *prog, int dtype, int n, const void *ploc, long ldp,
                  void *out, long ldo, void *stream)
{
    FR_API_BEGIN
    auto *p = static_cast<BCProgram *>(prog);
    FR_CHECK(p, "null program");
    FR_DTYPE_OK(dtype);
    FR_CHECK(n >= 0 && ldp >= n && ldo >= n, "invalid extents");
    if (n == 0) return 0;
    FR_CHECK(ploc && out, "null arrays");
    const dim3 grid(cdiv(n, 256)), block(256);
    if (dtype == FRHIP_F64)
        hipLaunchKernelGGL(k_bc_eval<double>, grid, block, 0,
                           as_stream(stream), n, p->nrows, p->code, p->offs,
                           p->pool, (const double *)ploc, ldp, p->t,
                           (double *)out, ldo);
    else
        hipLaunchKernelGGL(k_bc_eval<float>, grid, block, 0,
                           as_stream(stream), n, p->nrows, p->code, p->offs,
                           p->pool, (const float *)ploc, ldp, p->t,
                           (float *)out, ldo);
    FR_LAUNCHED();
    return 0;
    FR_API_END
}

}  // extern "C"
