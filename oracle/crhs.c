/* Oracle: C/OpenMP restatement of the reference's OpenMP-backend FR
 * right-hand side for hexahedra (test infrastructure and CPU baseline).
 *
 * It runs the kernel sequence of pyfr/solvers/baseadvec/system.py:57-136
 * (Euler) and pyfr/solvers/baseadvecdiff/system.py:53-231 (Navier-Stokes,
 * gradient fusion) with the point physics of the kernel templates
 * (euler/kernels/flux.mako:3-26, rsolvers/rusanov.mako:4-25,
 * navstokes/kernels/flux.mako:48-105, intconu.mako:4-19,
 * navstokes/kernels/intcflux.mako:10-54, tflux.mako, transform_grad.mako,
 * negdivconf.mako:7-22).  Data use the reference OpenMP backend's blocked
 * layout: 8-element SoA blocks, [block][row][variable][8]
 * (pyfr/backends/openmp/base.py:20-26).  The operators are applied as
 * sparse row lists (the reference hands them to libxsmm's sparse kernels,
 * pyfr/backends/openmp/xsmm.py:54-135).
 *
 * The inputs (operators, metric terms, interface pairs) come from the NumPy
 * oracle (oracle/rhs.py), so the two restatements share set-up and differ
 * only in the arithmetic they are checked against each other for.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define SB 8   /* SoA block width (elements) */
#define NV 5

typedef struct {
    int m, k, nnz;
    int *rowptr, *col;
    double *val;
} csr_t;

typedef struct {
    int ns;                 /* 0 Euler, 1 Navier-Stokes */
    int E, nblk, Nu, Nf;
    csr_t m0, m132, m3, m460, m6;
    double *smats;          /* [blk][Nu][3][3][SB] */
    double *rcpdjac;        /* [blk][Nu][SB] */
    int npairs;
    long *lo, *ro;          /* offsets of (e, fpt) in a [blk][Nf][NV][SB]
                               array, without the variable term */
    double *nl;             /* [npairs][3] */
    double gamma, mu, pr, beta, tau;
    double *u, *uf, *F, *grad, *gf, *comm, *out;
} crhs_t;

static void csr_make(csr_t *c, int m, int k, const double *a)
{
    c->m = m;
    c->k = k;
    c->rowptr = malloc((m + 1)*sizeof(int));
    int nnz = 0;
    for (long i = 0; i < (long)m*k; ++i) nnz += a[i] != 0.0;
    c->nnz = nnz;
    c->col = malloc((nnz ? nnz : 1)*sizeof(int));
    c->val = malloc((nnz ? nnz : 1)*sizeof(double));
    int q = 0;
    for (int r = 0; r < m; ++r) {
        c->rowptr[r] = q;
        for (int j = 0; j < k; ++j)
            if (a[(long)r*k + j] != 0.0) {
                c->col[q] = j;
                c->val[q++] = a[(long)r*k + j];
            }
    }
    c->rowptr[m] = q;
}

static void csr_free(csr_t *c)
{
    free(c->rowptr);
    free(c->col);
    free(c->val);
}

/* out[blk][r][v][SB] (=|+=) sum_j A[r][j] in[blk][j][v][SB]; rows of in and
 * out are NV*SB wide, stride between blocks is rows*NV*SB */
static void spmm(const csr_t *A, const double *in, long istride, double *out,
                 long ostride, int nblk, int accumulate)
{
    #pragma omp parallel for schedule(static)
    for (int b = 0; b < nblk; ++b) {
        const double *ib = in + b*istride;
        double *ob = out + b*ostride;
        for (int r = 0; r < A->m; ++r) {
            double acc[NV*SB];
            for (int x = 0; x < NV*SB; ++x) acc[x] = 0.0;
            for (int q = A->rowptr[r]; q < A->rowptr[r + 1]; ++q) {
                const double a = A->val[q];
                const double *row = ib + (long)A->col[q]*NV*SB;
                for (int x = 0; x < NV*SB; ++x) acc[x] += a*row[x];
            }
            double *orow = ob + (long)r*NV*SB;
            if (accumulate)
                for (int x = 0; x < NV*SB; ++x) orow[x] += acc[x];
            else
                for (int x = 0; x < NV*SB; ++x) orow[x] = acc[x];
        }
    }
}

static inline void inviscid_flux(const double s[NV], double f[3][NV], double *p,
                          double v[3], double gamma)
{
    const double invrho = 1.0/s[0], E = s[4];
    const double rhov[3] = {s[1], s[2], s[3]};
    for (int i = 0; i < 3; ++i) v[i] = invrho*rhov[i];
    *p = (gamma - 1)*(E - 0.5*invrho*(rhov[0]*rhov[0] + rhov[1]*rhov[1] +
                                      rhov[2]*rhov[2]));
    for (int i = 0; i < 3; ++i) {
        f[i][0] = rhov[i];
        f[i][4] = (E + *p)*v[i];
        for (int j = 0; j < 3; ++j)
            f[i][j + 1] = rhov[i]*v[j] + (i == j ? *p : 0.0);
    }
}

static inline void rusanov(const double ul[NV], const double ur[NV],
                    const double n[3], double nf[NV], double gamma)
{
    double fl[3][NV], fr[3][NV], vl[3], vr[3], pl, pr;
    inviscid_flux(ul, fl, &pl, vl, gamma);
    inviscid_flux(ur, fr, &pr, vr, gamma);
    const double nv = n[0]*(vl[0] + vr[0]) + n[1]*(vl[1] + vr[1]) +
                      n[2]*(vl[2] + vr[2]);
    const double a = sqrt((0.25*gamma)*(pl + pr)/(ul[0] + ur[0])) +
                     0.25*fabs(nv);
    for (int i = 0; i < NV; ++i)
        nf[i] = 0.5*(n[0]*(fl[0][i] + fr[0][i]) + n[1]*(fl[1][i] + fr[1][i]) +
                     n[2]*(fl[2][i] + fr[2][i])) + a*(ul[i] - ur[i]);
}

static inline void viscous_flux_add(const double s[NV], const double g[3][NV],
                             double f[3][NV], double mu, double gamma,
                             double pr)
{
    const double rho = s[0], E = s[4], rcprho = 1.0/rho;
    const double uu = rcprho*s[1], vv = rcprho*s[2], ww = rcprho*s[3];
    const double rho_x = g[0][0], rho_y = g[1][0], rho_z = g[2][0];
    const double u_x = g[0][1] - uu*rho_x, u_y = g[1][1] - uu*rho_y,
                 u_z = g[2][1] - uu*rho_z;
    const double v_x = g[0][2] - vv*rho_x, v_y = g[1][2] - vv*rho_y,
                 v_z = g[2][2] - vv*rho_z;
    const double w_x = g[0][3] - ww*rho_x, w_y = g[1][3] - ww*rho_y,
                 w_z = g[2][3] - ww*rho_z;
    const double E_x = g[0][4], E_y = g[1][4], E_z = g[2][4];
    const double T_x = rcprho*(E_x - (rcprho*rho_x*E + uu*u_x + vv*v_x +
                                      ww*w_x));
    const double T_y = rcprho*(E_y - (rcprho*rho_y*E + uu*u_y + vv*v_y +
                                      ww*w_y));
    const double T_z = rcprho*(E_z - (rcprho*rho_z*E + uu*u_z + vv*v_z +
                                      ww*w_z));
    const double third = 1.0/3.0, div = u_x + v_y + w_z;
    const double t_xx = -2*mu*rcprho*(u_x - third*div);
    const double t_yy = -2*mu*rcprho*(v_y - third*div);
    const double t_zz = -2*mu*rcprho*(w_z - third*div);
    const double t_xy = -mu*rcprho*(v_x + u_y);
    const double t_xz = -mu*rcprho*(u_z + w_x);
    const double t_yz = -mu*rcprho*(w_y + v_z);
    const double gpr = gamma/pr;
    f[0][1] += t_xx; f[1][1] += t_xy; f[2][1] += t_xz;
    f[0][2] += t_xy; f[1][2] += t_yy; f[2][2] += t_yz;
    f[0][3] += t_xz; f[1][3] += t_yz; f[2][3] += t_zz;
    f[0][4] += uu*t_xx + vv*t_xy + ww*t_xz + -mu*gpr*T_x;
    f[1][4] += uu*t_xy + vv*t_yy + ww*t_yz + -mu*gpr*T_y;
    f[2][4] += uu*t_xz + vv*t_yz + ww*t_zz + -mu*gpr*T_z;
}

/* ---- set-up ------------------------------------------------------------ */

/* ops: M0 (Nf x Nu), M1-M3*M2 (Nu x 3Nu), M3 (Nu x Nf), M4-M6*M0 (3Nu x Nu),
 * M6 (3Nu x Nf), row-major.  smats (Nu, 3, 3, E), rcpdjac (Nu, E) in the
 * NumPy oracle's layout.  Pairs: left/right (fpt, element) per flux-point
 * pair and the left normals nl (3, npairs).  consts = {gamma, mu, Pr, beta,
 * tau}. */
void *crhs_create(int ns, int E, int Nu, int Nf, const double *m0,
                  const double *m132, const double *m3, const double *m460,
                  const double *m6, const double *smats,
                  const double *rcpdjac, int npairs, const long *lf,
                  const long *le, const long *rf, const long *re,
                  const double *nl, const double *consts)
{
    crhs_t *c = calloc(1, sizeof(crhs_t));
    c->ns = ns;
    c->E = E;
    c->nblk = (E + SB - 1)/SB;
    c->Nu = Nu;
    c->Nf = Nf;
    csr_make(&c->m0, Nf, Nu, m0);
    csr_make(&c->m132, Nu, 3*Nu, m132);
    csr_make(&c->m3, Nu, Nf, m3);
    if (ns) {
        csr_make(&c->m460, 3*Nu, Nu, m460);
        csr_make(&c->m6, 3*Nu, Nf, m6);
    }
    const long nb = c->nblk;
    c->smats = calloc(nb*Nu*9*SB, sizeof(double));
    c->rcpdjac = calloc(nb*Nu*SB, sizeof(double));
    for (int e = 0; e < E; ++e)
        for (int j = 0; j < Nu; ++j) {
            for (int q = 0; q < 9; ++q)
                c->smats[(((long)(e/SB)*Nu + j)*9 + q)*SB + e%SB] =
                    smats[((long)j*9 + q)*E + e];
            c->rcpdjac[((long)(e/SB)*Nu + j)*SB + e%SB] =
                rcpdjac[(long)j*E + e];
        }
    c->npairs = npairs;
    c->lo = malloc(npairs*sizeof(long));
    c->ro = malloc(npairs*sizeof(long));
    c->nl = malloc(3*(long)npairs*sizeof(double));
    for (int i = 0; i < npairs; ++i) {
        c->lo[i] = ((le[i]/SB)*Nf + lf[i])*NV*SB + le[i]%SB;
        c->ro[i] = ((re[i]/SB)*Nf + rf[i])*NV*SB + re[i]%SB;
        for (int d = 0; d < 3; ++d) c->nl[3*i + d] = nl[(long)d*npairs + i];
    }
    c->gamma = consts[0];
    c->mu = consts[1];
    c->pr = consts[2];
    c->beta = consts[3];
    c->tau = consts[4];
    const long su = nb*Nu*NV*SB, sf = nb*Nf*NV*SB;
    c->u = calloc(su, sizeof(double));
    c->out = calloc(su, sizeof(double));
    c->uf = calloc(sf, sizeof(double));
    c->F = calloc(3*su, sizeof(double));
    if (ns) {
        c->grad = calloc(3*su, sizeof(double));
        c->gf = calloc(3*sf, sizeof(double));
        c->comm = calloc(sf, sizeof(double));
    }
    return c;
}

void crhs_destroy(void *h)
{
    crhs_t *c = h;
    csr_free(&c->m0);
    csr_free(&c->m132);
    csr_free(&c->m3);
    if (c->ns) {
        csr_free(&c->m460);
        csr_free(&c->m6);
    }
    free(c->smats); free(c->rcpdjac); free(c->lo); free(c->ro); free(c->nl);
    free(c->u); free(c->out); free(c->uf); free(c->F);
    free(c->grad); free(c->gf); free(c->comm);
    free(c);
}

/* u / out in the NumPy oracle's (Nu, 5, E) layout */
void crhs_set_u(void *h, const double *u)
{
    crhs_t *c = h;
    for (int j = 0; j < c->Nu; ++j)
        for (int v = 0; v < NV; ++v)
            for (int e = 0; e < c->E; ++e)
                c->u[(((long)(e/SB)*c->Nu + j)*NV + v)*SB + e%SB] =
                    u[((long)j*NV + v)*c->E + e];
}

void crhs_get_out(void *h, double *out)
{
    crhs_t *c = h;
    for (int j = 0; j < c->Nu; ++j)
        for (int v = 0; v < NV; ++v)
            for (int e = 0; e < c->E; ++e)
                out[((long)j*NV + v)*c->E + e] =
                    c->out[(((long)(e/SB)*c->Nu + j)*NV + v)*SB + e%SB];
}

/* ---- the right-hand side ----------------------------------------------- */

static void tflux(crhs_t *c)
{
    const int Nu = c->Nu, nb = c->nblk;
    const long su = (long)nb*Nu*NV*SB;
    /* padding elements (rho = 0) produce non-finite values in their own
       columns only, which nothing reads back */
    #pragma omp parallel for schedule(static)
    for (int b = 0; b < nb; ++b)
        for (int j = 0; j < Nu; ++j)
            #pragma omp simd
            for (int x = 0; x < SB; ++x) {
                const long ub = ((long)b*Nu + j)*NV*SB + x;
                double s[NV], f[3][NV], p, vel[3];
                for (int v = 0; v < NV; ++v) s[v] = c->u[ub + v*SB];
                inviscid_flux(s, f, &p, vel, c->gamma);
                const double *S = c->smats + ((long)b*Nu + j)*9*SB + x;
                if (c->ns) {
                    /* transform_grad.mako:3-10, in place */
                    const double rcpd = c->rcpdjac[((long)b*Nu + j)*SB + x];
                    double tg[3][NV], g[3][NV];
                    for (int d = 0; d < 3; ++d)
                        for (int v = 0; v < NV; ++v)
                            tg[d][v] = c->grad[d*su + ub + v*SB];
                    for (int i = 0; i < 3; ++i)
                        for (int v = 0; v < NV; ++v) {
                            g[i][v] = rcpd*(S[(0*3 + i)*SB]*tg[0][v] +
                                            S[(1*3 + i)*SB]*tg[1][v] +
                                            S[(2*3 + i)*SB]*tg[2][v]);
                            c->grad[i*su + ub + v*SB] = g[i][v];
                        }
                    viscous_flux_add(s, g, f, c->mu, c->gamma, c->pr);
                }
                for (int i = 0; i < 3; ++i)
                    for (int v = 0; v < NV; ++v)
                        c->F[i*su + ub + v*SB] = S[(i*3 + 0)*SB]*f[0][v] +
                                                 S[(i*3 + 1)*SB]*f[1][v] +
                                                 S[(i*3 + 2)*SB]*f[2][v];
            }
}

/* the (D, Nu) rows of a vector array as one (D*Nu) row block per element
 * block: F and grad are stored [dim][blk][Nu][V][SB], so the operator
 * applications over D*Nu rows gather the dimension planes */
static void spmm_vec_in(const csr_t *A, const double *in, long dstride,
                        int Nu, double *out, long ostride, int nblk,
                        int accumulate)
{
    #pragma omp parallel for schedule(static)
    for (int b = 0; b < nblk; ++b) {
        double *ob = out + b*ostride;
        for (int r = 0; r < A->m; ++r) {
            double acc[NV*SB];
            for (int x = 0; x < NV*SB; ++x) acc[x] = 0.0;
            for (int q = A->rowptr[r]; q < A->rowptr[r + 1]; ++q) {
                const int cc = A->col[q], d = cc/Nu, j = cc % Nu;
                const double a = A->val[q];
                const double *row = in + d*dstride +
                                    ((long)b*Nu + j)*NV*SB;
                for (int x = 0; x < NV*SB; ++x) acc[x] += a*row[x];
            }
            double *orow = ob + (long)r*NV*SB;
            if (accumulate)
                for (int x = 0; x < NV*SB; ++x) orow[x] += acc[x];
            else
                for (int x = 0; x < NV*SB; ++x) orow[x] = acc[x];
        }
    }
}

static void spmm_vec_out(const csr_t *A, const double *in, long istride,
                         int Nu, double *out, long dstride, int nblk,
                         int accumulate)
{
    #pragma omp parallel for schedule(static)
    for (int b = 0; b < nblk; ++b) {
        const double *ib = in + b*istride;
        for (int r = 0; r < A->m; ++r) {
            double acc[NV*SB];
            for (int x = 0; x < NV*SB; ++x) acc[x] = 0.0;
            for (int q = A->rowptr[r]; q < A->rowptr[r + 1]; ++q) {
                const double a = A->val[q];
                const double *row = ib + (long)A->col[q]*NV*SB;
                for (int x = 0; x < NV*SB; ++x) acc[x] += a*row[x];
            }
            const int d = r/Nu, j = r % Nu;
            double *orow = out + d*dstride + ((long)b*Nu + j)*NV*SB;
            if (accumulate)
                for (int x = 0; x < NV*SB; ++x) orow[x] += acc[x];
            else
                for (int x = 0; x < NV*SB; ++x) orow[x] = acc[x];
        }
    }
}

void crhs_rhs(void *h)
{
    crhs_t *c = h;
    const int Nu = c->Nu, Nf = c->Nf, nb = c->nblk;
    const long su = (long)nb*Nu*NV*SB, sf = (long)nb*Nf*NV*SB;
    const long ustr = (long)Nu*NV*SB, fstr = (long)Nf*NV*SB;

    spmm(&c->m0, c->u, ustr, c->uf, fstr, nb, 0);                 /* disu */

    if (c->ns) {
        const double be = c->beta;
        #pragma omp parallel for schedule(static)                 /* intconu */
        for (int i = 0; i < c->npairs; ++i)
            for (int v = 0; v < NV; ++v) {
                const double ul = c->uf[c->lo[i] + v*SB];
                const double ur = c->uf[c->ro[i] + v*SB];
                const double cu = ur*(0.5 + be) + ul*(0.5 - be);
                c->comm[c->lo[i] + v*SB] = cu;
                c->comm[c->ro[i] + v*SB] = cu;
            }
        /* tgradpcoru_upts, tgradcoru_upts */
        spmm_vec_out(&c->m460, c->u, ustr, Nu, c->grad, su, nb, 0);
        spmm_vec_out(&c->m6, c->comm, fstr, Nu, c->grad, su, nb, 1);
    }

    tflux(c);                                              /* tdisf(_fused) */

    if (c->ns) {
        for (int d = 0; d < 3; ++d)                        /* gradcoru_fpts */
            spmm(&c->m0, c->grad + d*su, ustr, c->gf + d*sf, fstr, nb, 0);
    }

    const double gamma = c->gamma;
    #pragma omp parallel for schedule(static)                    /* intcflux */
    for (int i = 0; i < c->npairs; ++i) {
        double ul[NV], ur[NV], fn[NV];
        const double *n = c->nl + 3*i;
        const double mag = sqrt(n[0]*n[0] + n[1]*n[1] + n[2]*n[2]);
        const double nn[3] = {(1/mag)*n[0], (1/mag)*n[1], (1/mag)*n[2]};
        for (int v = 0; v < NV; ++v) {
            ul[v] = c->uf[c->lo[i] + v*SB];
            ur[v] = c->uf[c->ro[i] + v*SB];
        }
        rusanov(ul, ur, nn, fn, gamma);
        if (c->ns) {
            double gl[3][NV], gr[3][NV], fvl[3][NV] = {{0}}, fvr[3][NV] = {{0}};
            for (int d = 0; d < 3; ++d)
                for (int v = 0; v < NV; ++v) {
                    gl[d][v] = c->gf[d*sf + c->lo[i] + v*SB];
                    gr[d][v] = c->gf[d*sf + c->ro[i] + v*SB];
                }
            viscous_flux_add(ul, gl, fvl, c->mu, gamma, c->pr);
            viscous_flux_add(ur, gr, fvr, c->mu, gamma, c->pr);
            for (int v = 0; v < NV; ++v) {
                const double nfl = nn[0]*fvl[0][v] + nn[1]*fvl[1][v] +
                                   nn[2]*fvl[2][v];
                const double nfr = nn[0]*fvr[0][v] + nn[1]*fvr[1][v] +
                                   nn[2]*fvr[2][v];
                double fv;
                if (c->beta == -0.5) fv = nfr;
                else if (c->beta == 0.5) fv = nfl;
                else fv = (0.5 + c->beta)*nfl + (0.5 - c->beta)*nfr;
                if (c->tau != 0.0) fv = fv + c->tau*(ul[v] - ur[v]);
                fn[v] = fn[v] + fv;
            }
        }
        for (int v = 0; v < NV; ++v) {
            c->uf[c->lo[i] + v*SB] = mag*fn[v];
            c->uf[c->ro[i] + v*SB] = -mag*fn[v];
        }
    }

    spmm_vec_in(&c->m132, c->F, su, Nu, c->out, ustr, nb, 0); /* tdivtpcorf */
    spmm(&c->m3, c->uf, fstr, c->out, ustr, nb, 1);            /* tdivtconf */

    #pragma omp parallel for schedule(static)                  /* negdivconf */
    for (int b = 0; b < nb; ++b)
        for (int j = 0; j < Nu; ++j)
            for (int v = 0; v < NV; ++v)
                for (int x = 0; x < SB; ++x) {
                    double *o = c->out + (((long)b*Nu + j)*NV + v)*SB + x;
                    *o = -c->rcpdjac[((long)b*Nu + j)*SB + x]*(*o);
                }
}
