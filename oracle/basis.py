"""Oracle: polynomial bases and FR operator matrices (test infrastructure).

Restates ``pyfr/polys.py``, ``pyfr/nputil.py:10-57`` and the hex/quad parts of
``pyfr/shapes.py`` in NumPy.  Quadrature points come from
``numpy.polynomial.legendre.leggauss`` instead of the reference's tabulated
rule files; tests pin them against values read from those files.
"""
from functools import cached_property

import numpy as np


def block_diag(arrs):
    """``pyfr/nputil.py:10-20``."""
    shapes = [a.shape for a in arrs]
    out = np.zeros(np.sum(shapes, axis=0), dtype=arrs[0].dtype)
    r = c = 0
    for a, (rr, cc) in zip(arrs, shapes):
        out[r:r + rr, c:c + cc] = a
        r, c = r + rr, c + cc
    return out


def clean(arr, tol=1e-10):
    """Flush tiny entries and coalesce near-equal magnitudes to their median.

    Follows ``pyfr/nputil.py:23-57``: |x| < tol -> 0; runs of sorted
    magnitudes that are np.isclose(rtol=tol, atol=tol/10) to the run's first
    element are replaced by the run's median; signs are restored.
    """
    arr = np.array(arr, dtype=float, copy=True)
    arr[np.abs(arr) < tol] = 0
    if arr.size > 1:
        amfl = np.abs(arr.ravel()).copy()
        amix = np.argsort(amfl)
        i, ix = 0, amix[0]
        j = 0
        for j, jx in enumerate(amix[1:], start=1):
            if not np.isclose(amfl[jx], amfl[ix], rtol=tol, atol=0.1*tol):
                if j - i > 1:
                    amfl[amix[i:j]] = np.median(amfl[amix[i:j]])
                i, ix = j, jx
        if i != j:
            amfl[amix[i:]] = np.median(amfl[amix[i:]])
        arr = np.copysign(amfl, arr.ravel()).reshape(arr.shape)
    return arr


def jacobi(n, a, b, z):
    """Jacobi polynomials P_0..P_n^(a,b)(z); ``pyfr/polys.py:9-28``."""
    j = [np.ones_like(z)]
    if n >= 1:
        j.append(((a + b + 2)*z + a - b) / 2)
    if n >= 2:
        apb, bbmaa = a + b, b*b - a*a
        for q in range(2, n + 1):
            qapbpq, apbp2q = q*(apb + q), apb + 2*q
            apbp2qm1, apbp2qm2 = apbp2q - 1, apbp2q - 2
            aq = apbp2q*apbp2qm1/(2*qapbpq)
            bq = apbp2qm1*bbmaa/(2*qapbpq*apbp2qm2)
            cq = apbp2q*(a + q - 1)*(b + q - 1)/(qapbpq*apbp2qm2)
            j.append((aq*z - bq)*j[-1] - cq*j[-2])
    return j


def jacobi_diff(n, a, b, z):
    """Derivatives of the Jacobi polynomials; ``pyfr/polys.py:31-38``."""
    dj = [np.zeros_like(z)]
    if n >= 1:
        dj.extend(jp*(i + a + b + 2)/2
                  for i, jp in enumerate(jacobi(n - 1, a + 1, b + 1, z)))
    return dj


class _Basis:
    """Orthonormal + nodal basis; ``pyfr/polys.py:45-96``."""

    def __init__(self, order, pts):
        self.order = order
        self.pts = np.asarray(pts, dtype=float)

    def ortho_basis_at(self, pts, do_clean=True):
        pts = np.atleast_2d(np.atleast_1d(np.asarray(pts, dtype=float)).T)
        m = np.array(self._ortho(*pts))
        return clean(m) if do_clean else m

    def jac_ortho_basis_at(self, pts, do_clean=True):
        pts = np.atleast_2d(np.atleast_1d(np.asarray(pts, dtype=float)).T)
        m = np.array(self._jac_ortho(*pts)).swapaxes(0, 1)
        return clean(m) if do_clean else m

    @cached_property
    def vdm(self):
        return self.ortho_basis_at(self.pts)

    @cached_property
    def invvdm(self):
        return clean(np.linalg.inv(self.vdm))

    def nodal_basis_at(self, epts):
        return clean((self.invvdm @ self.ortho_basis_at(epts, False)).T)

    def jac_nodal_basis_at(self, epts):
        return clean(self.invvdm @ self.jac_ortho_basis_at(epts, False))


class LineBasis(_Basis):
    """``pyfr/polys.py:99-114``."""

    def _ortho(self, p):
        return [(i + 0.5)**0.5*jp for i, jp in enumerate(jacobi(self.order, 0, 0, p))]

    def _jac_ortho(self, p):
        return [((i + 0.5)**0.5*jp,)
                for i, jp in enumerate(jacobi_diff(self.order, 0, 0, p))]


class QuadBasis(_Basis):
    """``pyfr/polys.py:170-195``."""

    def _ortho(self, p, q):
        sk = [(k + 0.5)**0.5 for k in range(self.order + 1)]
        pa = [c*jp for c, jp in zip(sk, jacobi(self.order, 0, 0, p))]
        pb = [c*jp for c, jp in zip(sk, jacobi(self.order, 0, 0, q))]
        return [pi*pj for pi in pa for pj in pb]


class HexBasis(_Basis):
    """``pyfr/polys.py:397-433``."""

    def _ortho(self, p, q, r):
        sk = [(k + 0.5)**0.5 for k in range(self.order + 1)]
        pa = [c*jp for c, jp in zip(sk, jacobi(self.order, 0, 0, p))]
        pb = [c*jp for c, jp in zip(sk, jacobi(self.order, 0, 0, q))]
        pc = [c*jp for c, jp in zip(sk, jacobi(self.order, 0, 0, r))]
        return [pi*pj*pk for pi in pa for pj in pb for pk in pc]

    def _jac_ortho(self, p, q, r):
        sk = [(k + 0.5)**0.5 for k in range(self.order + 1)]
        pa = [c*jp for c, jp in zip(sk, jacobi(self.order, 0, 0, p))]
        pb = [c*jp for c, jp in zip(sk, jacobi(self.order, 0, 0, q))]
        pc = [c*jp for c, jp in zip(sk, jacobi(self.order, 0, 0, r))]
        dpa = [c*jp for c, jp in zip(sk, jacobi_diff(self.order, 0, 0, p))]
        dpb = [c*jp for c, jp in zip(sk, jacobi_diff(self.order, 0, 0, q))]
        dpc = [c*jp for c, jp in zip(sk, jacobi_diff(self.order, 0, 0, r))]
        return [[dpi*pj*pk, pi*dpj*pk, pi*pj*dpk]
                for pi, dpi in zip(pa, dpa)
                for pj, dpj in zip(pb, dpb)
                for pk, dpk in zip(pc, dpc)]


def gauss_legendre_line(n):
    """n-point Gauss-Legendre rule on [-1, 1] (points ascending)."""
    x, w = np.polynomial.legendre.leggauss(n)
    return x, w


def gauss_legendre_quad(n):
    """Tensor GL rule with the first coordinate fastest (the ordering of
    ``pyfr/quadrules/quad/gauss-legendre-n*.txt``)."""
    x, w = gauss_legendre_line(n)
    pts = np.array([(x[i], x[j]) for j in range(n) for i in range(n)])
    wts = np.array([w[i]*w[j] for j in range(n) for i in range(n)])
    return pts, wts


def gauss_legendre_hex(n):
    """Tensor GL rule, x fastest then y then z
    (``pyfr/quadrules/hex/gauss-legendre-n*.txt``)."""
    x, w = gauss_legendre_line(n)
    pts = np.array([(x[i], x[j], x[k])
                    for k in range(n) for j in range(n) for i in range(n)])
    wts = np.array([w[i]*w[j]*w[k]
                    for k in range(n) for j in range(n) for i in range(n)])
    return pts, wts


# Hex faces: (reference-to-face projection, outward normal); shapes.py:378-385
HEX_FACES = [
    (lambda s, t: (s, t, -1*np.ones_like(s)), (0, 0, -1)),
    (lambda s, t: (s, -1*np.ones_like(s), t), (0, -1, 0)),
    (lambda s, t: (np.ones_like(s), s, t), (1, 0, 0)),
    (lambda s, t: (s, np.ones_like(s), t), (0, 1, 0)),
    (lambda s, t: (-1*np.ones_like(s), s, t), (-1, 0, 0)),
    (lambda s, t: (s, t, np.ones_like(s)), (0, 0, 1)),
]


def _proj(fn, pts2):
    """``pyfr/shapes.py:12-14``."""
    return np.vstack(np.broadcast_arrays(*fn(*pts2.T))).T


class HexShapeOracle:
    """FR operators for a hex with GL solution and flux points.

    Follows ``pyfr/shapes.py:22-135`` (m0..m6), ``:231-259`` (gbasis) and the
    ``HexShape`` face table ``:375-385``.
    """
    ndims = 3
    nfaces = 6

    def __init__(self, order, upts='gauss-legendre'):
        """upts: 'gauss-legendre' or 'gauss-legendre-lobatto' solution
        points (``[solver-elements-hex] soln-pts``); the flux points are
        Gauss-Legendre either way."""
        self.order = order
        n = order + 1
        if upts == 'gauss-legendre':
            self.upts, self.upts_wts = gauss_legendre_hex(n)
        else:
            # Lobatto: the ends and the roots of P'_(n-1); weights
            # 2/(n(n - 1) P_(n-1)^2)
            c = np.zeros(n)
            c[-1] = 1
            x = np.concatenate([[-1.0], np.sort(np.polynomial.legendre.legroots(
                np.polynomial.legendre.legder(c))), [1.0]])
            w = 2/(n*(n - 1)*np.polynomial.legendre.legval(x, c)**2)
            self.upts = np.array([(x[i], x[j], x[k]) for k in range(n)
                                  for j in range(n) for i in range(n)])
            self.upts_wts = np.array([w[i]*w[j]*w[k] for k in range(n)
                                      for j in range(n) for i in range(n)])
        self.nupts = n**3
        self.ubasis = HexBasis(order, self.upts)
        fq, self.fq_wts = gauss_legendre_quad(n)
        self.face_pts2 = fq
        self.nfacefpts = [n*n]*6
        self.fpts = np.vstack([_proj(p, fq) for p, _ in HEX_FACES])
        self.nfpts = 6*n*n
        self.norm_fpts = np.vstack([[nrm]*(n*n) for _, nrm in HEX_FACES]).astype(float)
        nf = np.cumsum([0] + self.nfacefpts)
        self.facefpts = [list(range(nf[i], nf[i + 1])) for i in range(6)]

    @cached_property
    def m0(self):
        return self.ubasis.nodal_basis_at(self.fpts)

    @cached_property
    def m1(self):
        m = np.rollaxis(self.ubasis.jac_nodal_basis_at(self.upts), 2)
        return m.reshape(self.nupts, -1)

    @cached_property
    def m2(self):
        m = self.norm_fpts[..., None]*self.m0[:, None, :]
        return m.reshape(self.nfpts, -1)

    @cached_property
    def gbasis_coeffs(self):
        # shapes.py:231-255: quad faces use a (p+1)^2 GL rule
        n = self.order + 1
        qp, qw = gauss_legendre_quad(n)
        fbasis = QuadBasis(self.order, self.face_pts2)
        coeffs = []
        for proj, _ in HEX_FACES:
            L = fbasis.nodal_basis_at(qp)
            M = self.ubasis.ortho_basis_at(_proj(proj, qp))
            coeffs.append(np.einsum('i,ik,ji->kj', qw, L, M))
        return np.vstack(coeffs)

    def gbasis_at(self, pts):
        return clean((self.gbasis_coeffs @ self.ubasis.ortho_basis_at(pts)).T)

    @cached_property
    def m3(self):
        m = self.gbasis_at(self.upts)
        if getattr(self, 'iq_pts', None) is not None:
            # shapes.py:107-111 with proj_l2 (:18-19) of the interface rule
            # onto the face basis at the standard flux points
            fb = QuadBasis(self.order, self.face_pts2)
            fp = clean(fb.vdm.T @ (self.iq_wts*fb.ortho_basis_at(self.iq_pts)))
            nq = len(self.iq_pts)
            bd = np.zeros((len(fp)*6, nq*6))
            for f in range(6):
                bd[f*len(fp):(f + 1)*len(fp), f*nq:(f + 1)*nq] = fp
            m = m @ bd
        return m

    @cached_property
    def m4(self):
        m = self.m1.reshape(self.nupts, -1, self.nupts).swapaxes(0, 1)
        return m.reshape(-1, self.nupts)

    @cached_property
    def m6(self):
        m = self.norm_fpts.T[:, None, :]*self.m3
        return m.reshape(-1, self.nfpts)

    # Surface-flux anti-aliasing (pyfr/shapes.py:196-224, :297-303): the
    # flux points of each face are the interface rule (Gauss-Legendre,
    # nf1 x nf1 points; default (order + 2)^2, shapes.py:159-161); M3 as
    # above.  The face basis of gbasis stays on the standard flux points.
    def set_surface_quadrature(self, nf1=None):
        self.iq_pts, self.iq_wts = gauss_legendre_quad(nf1 or self.order + 2)
        nq = len(self.iq_pts)
        self.nfacefpts = [nq]*6
        self.fpts = np.vstack([_proj(p, self.iq_pts) for p, _ in HEX_FACES])
        self.nfpts = 6*nq
        self.norm_fpts = np.vstack([[nrm]*nq for _, nrm in HEX_FACES]
                                   ).astype(float)
        self.facefpts = [list(range(f*nq, (f + 1)*nq)) for f in range(6)]
        for k in ('m0', 'm2', 'm3', 'm6'):
            self.__dict__.pop(k, None)

    # Flux anti-aliasing (pyfr/shapes.py:125-135, :18-19): the element
    # quadrature rule (Gauss-Legendre, nq1 points per direction; default
    # (order + 2)^3 as pyfr/shapes.py:165-167), M7 interpolation to it, M8
    # the L2 projection basis.vdm.T @ (w * psi(qpts)), M9 = blockdiag(M8 x 3)
    def set_quadrature(self, nq1=None):
        self.qpts, self.qwts = gauss_legendre_hex(nq1 or self.order + 2)
        for k in ('m7', 'm8'):
            self.__dict__.pop(k, None)

    @cached_property
    def m7(self):
        return self.ubasis.nodal_basis_at(self.qpts)

    @cached_property
    def m8(self):
        ub = self.ubasis
        return ub.vdm.T @ (self.qwts*ub.ortho_basis_at(self.qpts))

    @property
    def m9(self):
        m8, (a, b) = self.m8, self.m8.shape
        out = np.zeros((3*a, 3*b))
        for d in range(3):
            out[d*a:(d + 1)*a, d*b:(d + 1)*b] = m8
        return out

    def opmat(self, expr):
        """Evaluate an operator expression such as 'M1 - M3*M2' and clean it
        (``pyfr/shapes.py:79-87``)."""
        env = {f'M{i}': getattr(self, f'm{i}') for i in (0, 1, 2, 3, 4, 6)}
        if hasattr(self, 'qpts'):
            env |= {'M7': self.m7, 'M8': self.m8, 'M9': self.m9}
        return clean(eval(expr.replace('*', '@'), {'__builtins__': {}}, env))

    # Shape points of a linear hex (std_ele(1)); shapes.py:278-286
    @staticmethod
    def linspts():
        pts1d = np.linspace(-1, 1, 2)
        import itertools as it
        return np.array([p[::-1] for p in it.product(pts1d, repeat=3)])


# ---------------------------------------------------------------------------
# Simplex, prism and pyramid bases (pyfr/polys.py:117-167, 198-276,
# 279-392) and the TetShape / PriShape / PyrShape operators
# (pyfr/shapes.py:476-640).  Point tables are
# passed in by the caller (the tests take them from the product's embedded
# tables, which are themselves checked against the reference's rule files).
# ---------------------------------------------------------------------------

class TriBasis(_Basis):
    """``pyfr/polys.py:117-167``."""

    def _ortho(self, p, q):
        with np.errstate(divide='ignore', invalid='ignore'):
            a = np.where(q != 1, 2*(1 + p)/(1 - q) - 1, -1)
        b = q
        ob = []
        for i, pi in enumerate(jacobi(self.order, 0, 0, a)):
            pa = pi*(1 - b)**i
            for j, pj in enumerate(jacobi(self.order - i, 2*i + 1, 0, b)):
                cij = ((2*i + 1)*(2*i + 2*j + 2))**0.5 / 2**(i + 1)
                ob.append(cij*pa*pj)
        return ob


class TetBasis(_Basis):
    """``pyfr/polys.py:198-276``."""

    def _ortho(self, p, q, r):
        with np.errstate(divide='ignore', invalid='ignore'):
            a = np.where(r != -q, -2*(1 + p)/(q + r) - 1, -1)
            b = np.where(r != 1, 2*(1 + q)/(1 - r) - 1, -1)
        c = r
        ob = []
        for i, pi in enumerate(jacobi(self.order, 0, 0, a)):
            ci = 2**(-2*i - 1)*(2*i + 1)**0.5*(1 - b)**i
            for j, pj in enumerate(jacobi(self.order - i, 2*i + 1, 0, b)):
                cj = (i + j + 1)**0.5*2**-j*(1 - c)**(i + j)
                cij, pij = ci*cj, pi*pj
                jp = jacobi(self.order - i - j, 2*(i + j + 1), 0, c)
                for k, pk in enumerate(jp):
                    ck = (2*(k + j + i) + 3)**0.5
                    ob.append(cij*ck*pij*pk)
        return ob

    def _jac_ortho(self, p, q, r):
        with np.errstate(divide='ignore', invalid='ignore'):
            a = np.where(r != -q, -2*(1 + p)/(q + r) - 1, -1)
            b = np.where(r != 1, 2*(1 + q)/(1 - r) - 1, -1)
        c = r
        f = jacobi(self.order, 0, 0, a)
        df = jacobi_diff(self.order, 0, 0, a)
        ob = []
        for i, (fi, dfi) in enumerate(zip(f, df)):
            ci = 2**(-2*i - 1)*(2*i + 1)**0.5
            g = jacobi(self.order - i, 2*i + 1, 0, b)
            dg = jacobi_diff(self.order - i, 2*i + 1, 0, b)
            for j, (gj, dgj) in enumerate(zip(g, dg)):
                cj = (i + j + 1)**0.5*2**-j
                cij = ci*cj
                h = jacobi(self.order - i - j, 2*(i + j + 1), 0, c)
                dh = jacobi_diff(self.order - i - j, 2*(i + j + 1), 0, c)
                for k, (hk, dhk) in enumerate(zip(h, dh)):
                    ck = (2*(k + j + i) + 3)**0.5
                    cijk = cij*ck
                    tmp1 = (1 - c)**(i + j - 1) if i + j > 0 else 1
                    tmp2 = tmp1*(1 - b)**(i - 1) if i > 0 else 1
                    pijk = 4*tmp2*dfi*gj*hk
                    qijk = 2*(tmp2*(-i*fi + (1 + a)*dfi)*gj
                              + tmp1*(1 - b)**i*fi*dgj)*hk
                    rijk = (2*(1 + a)*tmp2*dfi*gj*hk
                            + (1 + b)*tmp1*(1 - b)**i*fi*dgj*hk
                            + (1 - c)**(i + j)*(1 - b)**i*fi*gj*dhk
                            - (i*(1 + b)*tmp2 + (i + j)*tmp1*(1 - b)**i)
                            * fi*gj*hk)
                    ob.append([cijk*pijk, cijk*qijk, cijk*rijk])
        return ob


class PriBasis(_Basis):
    """``pyfr/polys.py:279-338``."""

    def _ortho(self, p, q, r):
        with np.errstate(divide='ignore', invalid='ignore'):
            a = np.where(q != 1, 2*(1 + p)/(1 - q) - 1, -1)
        b, c = q, r
        pab = []
        for i, pi in enumerate(jacobi(self.order, 0, 0, a)):
            ci = (1 - b)**i / 2**(i + 1)
            for j, pj in enumerate(jacobi(self.order - i, 2*i + 1, 0, b)):
                cij = ((2*i + 1)*(2*i + 2*j + 2))**0.5*ci
                pab.append(cij*pi*pj)
        sk = [(k + 0.5)**0.5 for k in range(self.order + 1)]
        pc = [s*jp for s, jp in zip(sk, jacobi(self.order, 0, 0, c))]
        return [pij*pk for pij in pab for pk in pc]

    def _jac_ortho(self, p, q, r):
        with np.errstate(divide='ignore', invalid='ignore'):
            a = np.where(q != 1, 2*(1 + p)/(1 - q) - 1, -1)
        b, c = q, r
        f = jacobi(self.order, 0, 0, a)
        df = jacobi_diff(self.order, 0, 0, a)
        pab = []
        for i, (fi, dfi) in enumerate(zip(f, df)):
            g = jacobi(self.order - i, 2*i + 1, 0, b)
            dg = jacobi_diff(self.order - i, 2*i + 1, 0, b)
            for j, (gj, dgj) in enumerate(zip(g, dg)):
                cij = ((2*i + 1)*(2*i + 2*j + 2))**0.5 / 2**(i + 1)
                tmp = (1 - b)**(i - 1) if i > 0 else 1
                pij = 2*tmp*dfi*gj
                qij = tmp*(-i*fi + (1 + a)*dfi)*gj + (1 - b)**i*fi*dgj
                rij = (1 - b)**i*fi*gj
                pab.append([cij*pij, cij*qij, cij*rij])
        sk = [(k + 0.5)**0.5 for k in range(self.order + 1)]
        hc = [s*jp for s, jp in zip(sk, jacobi(self.order, 0, 0, c))]
        dhc = [s*jp for s, jp in zip(sk, jacobi_diff(self.order, 0, 0, c))]
        return [[pij*hk, qij*hk, rij*dhk]
                for pij, qij, rij in pab for hk, dhk in zip(hc, dhc)]


class PyrBasis(_Basis):
    """``pyfr/polys.py:332-392``: (1 - c)^(i + j) P_i(a) P_j(b)
    P_k^(2(i+j+1), 0)(c), a = 2p/(1 - r), b = 2q/(1 - r), c = r, scaled to
    unit norm; i, j <= order, k <= order - max(i, j)."""

    def _abc(self, p, q, r):
        with np.errstate(divide='ignore', invalid='ignore'):
            a = np.where(r != 1, 2*p/(1 - r), 0)
            b = np.where(r != 1, 2*q/(1 - r), 0)
        return a, b, r

    def _scaled_leg(self, z, diff=False):
        f = jacobi_diff if diff else jacobi
        return [2**(-k - 0.25)*(k + 0.5)**0.5*jp
                for k, jp in enumerate(f(self.order, 0, 0, z))]

    def _ortho(self, p, q, r):
        a, b, c = self._abc(p, q, r)
        pa, pb = self._scaled_leg(a), self._scaled_leg(b)
        ob = []
        for i, pi in enumerate(pa):
            for j, pj in enumerate(pb):
                jc = jacobi(self.order - max(i, j), 2*(i + j + 1), 0, c)
                for k, pk in enumerate(jc):
                    ob.append((2*(i + j + k) + 3)**0.5*(1 - c)**(i + j)
                              * pi*pj*pk)
        return ob

    def _jac_ortho(self, p, q, r):
        a, b, c = self._abc(p, q, r)
        f, df = self._scaled_leg(a), self._scaled_leg(a, True)
        g, dg = self._scaled_leg(b), self._scaled_leg(b, True)
        ob = []
        for i, (fi, dfi) in enumerate(zip(f, df)):
            for j, (gj, dgj) in enumerate(zip(g, dg)):
                n = self.order - max(i, j)
                h = jacobi(n, 2*(i + j + 1), 0, c)
                dh = jacobi_diff(n, 2*(i + j + 1), 0, c)
                # d/dp and d/dq through a, b; d/dr through a, b and c
                w = (1 - c)**(i + j - 1) if i + j > 0 else 1
                for k, (hk, dhk) in enumerate(zip(h, dh)):
                    ck = (2*(i + j + k) + 3)**0.5
                    dp = 2*w*dfi*gj*hk
                    dq = 2*w*fi*dgj*hk
                    dr = (w*(a*dfi*gj + b*fi*dgj - (i + j)*fi*gj)*hk
                          + (1 - c)**(i + j)*fi*gj*dhk)
                    ob.append([ck*dp, ck*dq, ck*dr])
        return ob


TET_FACES = [
    ('tri', lambda s, t: (s, t, -1*np.ones_like(s)), (0, 0, -1)),
    ('tri', lambda s, t: (s, -1*np.ones_like(s), t), (0, -1, 0)),
    ('tri', lambda s, t: (-1*np.ones_like(s), t, s), (-1, 0, 0)),
    ('tri', lambda s, t: (s, t, -s - t - 1), (1, 1, 1)),
]
PRI_FACES = [
    ('tri', lambda s, t: (s, t, -1*np.ones_like(s)), (0, 0, -1)),
    ('tri', lambda s, t: (s, t, np.ones_like(s)), (0, 0, 1)),
    ('quad', lambda s, t: (s, -1*np.ones_like(s), t), (0, -1, 0)),
    ('quad', lambda s, t: (-s, s, t), (1, 1, 0)),
    ('quad', lambda s, t: (-1*np.ones_like(s), s, t), (-1, 0, 0)),
]
# pyfr/shapes.py:573-585
PYR_FACES = [
    ('quad', lambda s, t: (s, t, -1*np.ones_like(s)), (0, 0, -1)),
    ('tri', lambda s, t: (s + (t + 1)/2, (t - 1)/2, t), (0, -1, 0.5)),
    ('tri', lambda s, t: ((1 - t)/2, -s - (t + 1)/2, t), (1, 0, 0.5)),
    ('tri', lambda s, t: (-s - (t + 1)/2, (1 - t)/2, t), (0, 1, 0.5)),
    ('tri', lambda s, t: ((t - 1)/2, s + (t + 1)/2, t), (-1, 0, 0.5)),
]


class SimplexShapeOracle:
    """TetShape / PriShape / PyrShape operators by the reference's
    algorithm.

    tables: {'upts': (pts, wts), 'tri_fpts': pts, 'tri_q36': (pts, wts)}."""

    def __init__(self, name, order, tables):
        self.name, self.order = name, order
        self.faces = {'tet': TET_FACES, 'pri': PRI_FACES,
                      'pyr': PYR_FACES}[name]
        basis = {'tet': TetBasis, 'pri': PriBasis, 'pyr': PyrBasis}[name]
        self.upts = np.asarray(tables['upts'][0], dtype=float)
        self.nupts = len(self.upts)
        self.ubasis = basis(order, self.upts)
        n = order + 1
        gq, _ = gauss_legendre_quad(n)
        self._frule = {'tri': np.asarray(tables['tri_fpts'], dtype=float),
                       'quad': gq}
        self._q36 = tables['tri_q36']
        self.nfacefpts = [len(self._frule[k]) for k, _, _ in self.faces]
        self.nfpts = sum(self.nfacefpts)
        self.fpts = np.vstack([_proj(p, self._frule[k])
                               for k, p, _ in self.faces])
        self.norm_fpts = np.vstack([[nm]*c for (_, _, nm), c in
                                    zip(self.faces, self.nfacefpts)]) * 1.0

    def _facebasis(self, kind):
        cls = TriBasis if kind == 'tri' else QuadBasis
        return cls(self.order, self._frule[kind])

    # Surface-flux anti-aliasing (pyfr/shapes.py:103-113, :196-224,
    # :297-303): the flux points of each face are its interface rule
    # iqrules[kind] = (points, weights); M3 is fed by proj_l2 of that rule
    # onto the face basis, which stays on the standard flux points
    def set_surface_quadrature(self, iqrules):
        self._iqrules = {k: tuple(np.asarray(x, dtype=float) for x in v)
                         for k, v in iqrules.items()}
        self.nfacefpts = [len(self._iqrules[k][0]) for k, _, _ in self.faces]
        self.nfpts = sum(self.nfacefpts)
        self.fpts = np.vstack([_proj(p, self._iqrules[k][0])
                               for k, p, _ in self.faces])
        self.norm_fpts = np.vstack([[nm]*c for (_, _, nm), c in
                                    zip(self.faces, self.nfacefpts)]) * 1.0
        for k in ('m0', 'm2', 'm3', 'm6'):
            self.__dict__.pop(k, None)

    @cached_property
    def gbasis_coeffs(self):
        coeffs = []
        n = self.order + 1
        for kind, proj, _ in self.faces:
            if kind == 'tri':
                qp, qw = (np.asarray(x, dtype=float) for x in self._q36)
            else:
                qp, qw = gauss_legendre_quad(n)
            L = self._facebasis(kind).nodal_basis_at(qp)
            M = self.ubasis.ortho_basis_at(_proj(proj, qp))
            coeffs.append(np.einsum('i,ik,ji->kj', qw, L, M))
        return np.vstack(coeffs)

    @cached_property
    def m0(self):
        return self.ubasis.nodal_basis_at(self.fpts)

    @cached_property
    def m1(self):
        m = np.rollaxis(self.ubasis.jac_nodal_basis_at(self.upts), 2)
        return m.reshape(self.nupts, -1)

    @cached_property
    def m2(self):
        m = self.norm_fpts[..., None]*self.m0[:, None, :]
        return m.reshape(self.nfpts, -1)

    @cached_property
    def m3(self):
        m = clean((self.gbasis_coeffs @
                   self.ubasis.ortho_basis_at(self.upts)).T)
        if getattr(self, '_iqrules', None):
            fps = []
            for kind, _, _ in self.faces:
                fb = self._facebasis(kind)
                qp, qw = self._iqrules[kind]
                fps.append(clean(fb.vdm.T @ (qw*fb.ortho_basis_at(qp))))
            bd = np.zeros((sum(f.shape[0] for f in fps),
                           sum(f.shape[1] for f in fps)))
            r = c = 0
            for f in fps:
                bd[r:r + f.shape[0], c:c + f.shape[1]] = f
                r, c = r + f.shape[0], c + f.shape[1]
            m = m @ bd
        return m

    @cached_property
    def m4(self):
        m = self.m1.reshape(self.nupts, -1, self.nupts).swapaxes(0, 1)
        return m.reshape(-1, self.nupts)

    @cached_property
    def m6(self):
        m = self.norm_fpts.T[:, None, :]*self.m3
        return m.reshape(-1, self.nfpts)

    # Flux anti-aliasing (pyfr/shapes.py:125-135, :18-19): an element
    # quadrature rule (points, weights), M7 interpolation to it, M8 the L2
    # projection basis.vdm.T @ (w * psi(qpts)), M9 = blockdiag(M8 x 3)
    def set_quadrature(self, qpts, qwts):
        self.qpts = np.asarray(qpts, dtype=float)
        self.qwts = np.asarray(qwts, dtype=float)
        for k in ('m7', 'm8'):
            self.__dict__.pop(k, None)

    @cached_property
    def m7(self):
        return self.ubasis.nodal_basis_at(self.qpts)

    @cached_property
    def m8(self):
        ub = self.ubasis
        return ub.vdm.T @ (self.qwts*ub.ortho_basis_at(self.qpts))

    @property
    def m9(self):
        m8, (a, b) = self.m8, self.m8.shape
        out = np.zeros((3*a, 3*b))
        for d in range(3):
            out[d*a:(d + 1)*a, d*b:(d + 1)*b] = m8
        return out

    def opmat(self, expr):
        env = {f'M{i}': getattr(self, f'm{i}') for i in (0, 1, 2, 3, 4, 6)}
        if hasattr(self, 'qpts'):
            env |= {'M7': self.m7, 'M8': self.m8, 'M9': self.m9}
        return clean(eval(expr.replace('*', '@'), {'__builtins__': {}}, env))
