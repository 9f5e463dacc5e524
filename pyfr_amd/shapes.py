"""FR operator matrices for tensor-product hexahedra.

The reference builds its operators from a Vandermonde matrix of an
orthonormal basis (``pyfr/shapes.py:79-135``, ``pyfr/polys.py:45-96``).  Here
the hex operators are assembled in closed form from one-dimensional
Gauss-Legendre quantities, which exposes the sparsity the device kernels use:

* ``l_j(+-1)``  Lagrange basis of the n GL points evaluated at the ends,
* ``D[i, j] = l_j'(x_i)``  the 1-D differentiation matrix,
* ``h^{+-}(x_i) = sum_k (k + 1/2) (+-1)^k P_k(x_i)``  the 1-D reproducing
  kernel at the ends, which is the value of the DG/FR correction function of
  ``BaseShape.gbasis_at`` (``pyfr/shapes.py:231-259``) on the line normal to a
  face when the face flux points are the GL points.

Orderings follow the reference exactly: solution points are the GL tensor
rule with x fastest (``pyfr/quadrules/hex/gauss-legendre-n*.txt``), flux points
run face by face through ``HexShape.faces`` (``pyfr/shapes.py:378-385``) with
the first face coordinate fastest, and the operator layouts are those of
``m0``..``m6`` (``pyfr/shapes.py:89-123``).
"""
from functools import cached_property
import itertools as it
import math
import re

import numpy as np


# Face table of HexShape (pyfr/shapes.py:378-385): for each face the axis of
# the face normal, its side (-1/+1) and the axes that the face coordinates
# (s, t) map onto.
HEX_FACES = [
    # (normal axis, side, s axis, t axis)
    (2, -1, 0, 1),
    (1, -1, 0, 2),
    (0, +1, 1, 2),
    (1, +1, 0, 2),
    (0, -1, 1, 2),
    (2, +1, 0, 1),
]


def _legendre(n, x):
    """P_n(x) and P_n'(x) by the three-term recurrence (pure Python)."""
    p0, p1 = 1.0, x
    if n == 0:
        return 1.0, 0.0
    for k in range(2, n + 1):
        p0, p1 = p1, ((2*k - 1)*x*p1 - (k - 1)*p0)/k
    dp = n*(x*p1 - p0)/(x*x - 1.0)
    return p1, dp


def gauss_legendre(n):
    """n-point Gauss-Legendre nodes (ascending) and weights.

    Newton iteration on P_n in plain Python floats from cosine initial
    guesses, symmetrised, so that the result is bit-identical on every host
    (the operator kernels bake these values in at build time and match them
    bit for bit at run time)."""
    x = [0.0]*n
    w = [0.0]*n
    for i in range((n + 1)//2):
        z = math.cos(math.pi*(i + 0.75)/(n + 0.5))
        for _ in range(100):
            p, dp = _legendre(n, z)
            dz = p/dp
            z -= dz
            if abs(dz) < 1e-16:
                break
        p, dp = _legendre(n, z)
        wi = 2.0/((1.0 - z*z)*dp*dp)
        x[i], x[n - 1 - i] = -z, z
        w[i] = w[n - 1 - i] = wi
    if n % 2:
        x[n//2] = 0.0
    return np.array(x), np.array(w)


def gauss_legendre_lobatto(n):
    """n-point Gauss-Legendre-Lobatto nodes (ascending: -1, the roots of
    P'_(n-1), 1) and weights 2/(n(n - 1) P_(n-1)(x)^2), by the Newton
    iteration x <- x - (x P_N - P_(N-1))/((N + 1) P_N), N = n - 1, from the
    Chebyshev-Gauss-Lobatto points (Trefethen's formulation), in plain
    Python floats so every host gets the same bits."""
    if n < 2:
        raise ValueError('Gauss-Legendre-Lobatto rules need n >= 2')
    N = n - 1
    xs, ws = [], []
    for i in range(n):
        x = -math.cos(math.pi*i/N)
        for _ in range(100):
            p0, p1 = 1.0, x
            for k in range(2, N + 1):
                p0, p1 = p1, ((2*k - 1)*x*p1 - (k - 1)*p0)/k
            dx = (x*p1 - p0)/((N + 1)*p1)
            x -= dx
            if abs(dx) < 1e-17:
                break
        if i in (0, N):
            x = float(2*(i == N) - 1)
        p0, p1 = 1.0, x
        for k in range(2, N + 1):
            p0, p1 = p1, ((2*k - 1)*x*p1 - (k - 1)*p0)/k
        xs.append(x)
        ws.append(2/(N*(N + 1)*p1*p1))
    if n % 2:
        xs[N//2] = 0.0
    return np.array(xs), np.array(ws)


def lagrange_at(x, pts):
    """Values l_j(pts[q]) of the Lagrange basis through x: (len(pts), n)."""
    n = len(x)
    x = [float(v) for v in x]
    out = np.ones((len(pts), n))
    for q, t in enumerate(pts):
        for j in range(n):
            v = 1.0
            for m in range(n):
                if m != j:
                    v *= (float(t) - x[m])/(x[j] - x[m])
            out[q, j] = v
    return out


def diff_matrix(x):
    """D[i, j] = l_j'(x_i) via barycentric weights."""
    n = len(x)
    x = [float(v) for v in x]
    wb = []
    for j in range(n):
        p = 1.0
        for m in range(n):
            if m != j:
                p *= x[j] - x[m]
        wb.append(1.0/p)
    D = np.zeros((n, n))
    for i in range(n):
        for j in range(n):
            if i != j:
                D[i, j] = wb[j]/wb[i]/(x[i] - x[j])
        D[i, i] = -math.fsum(D[i, j] for j in range(n) if j != i)
    return D


def legendre_kernel_end(x, side):
    """h(x_i) = sum_{k<=p} (k + 1/2) side^k P_k(x_i) for n = len(x) points."""
    n = len(x)
    out = np.empty(n)
    for i, xi in enumerate(x):
        terms, p0, p1 = [], 1.0, float(xi)
        for k in range(n):
            pk = 1.0 if k == 0 else (float(xi) if k == 1 else None)
            if k >= 2:
                p0, p1 = p1, ((2*k - 1)*float(xi)*p1 - (k - 1)*p0)/k
                pk = p1
            terms.append((k + 0.5)*(side**k)*pk)
        out[i] = math.fsum(terms)
    return out


def sparse_matmul(a, b):
    """a @ b with each entry the correctly rounded (math.fsum) sum of its
    nonzero products: bit-reproducible on any host, unlike BLAS."""
    out = np.zeros((a.shape[0], b.shape[1]))
    bnz = [np.flatnonzero(b[k]) for k in range(b.shape[0])]
    for i in range(a.shape[0]):
        acc = {}
        for k in np.flatnonzero(a[i]):
            aik = float(a[i, k])
            for j in bnz[k]:
                acc.setdefault(j, []).append(aik*float(b[k, j]))
        for j, terms in acc.items():
            out[i, j] = math.fsum(terms)
    return out


def flush(a, tol=1e-10):
    """Zero entries below tol (the first step of ``pyfr/nputil.py:23-57``)."""
    a = np.array(a, dtype=float, copy=True)
    a[np.abs(a) < tol] = 0
    return a


def l2_projection_1d(x, nq):
    """(len(x), nq): values at the nodes x of the L2 projection, onto the
    polynomials of degree < len(x), of data at the nq-point Gauss-Legendre
    rule: sum_k (k + 1/2) P_k(x_i) sum_q w_q P_k(x_q) f_q."""
    xq, wq = gauss_legendre(nq)
    P = np.zeros((len(x), nq))
    for i, xi in enumerate(x):
        for q, (xqq, wqq) in enumerate(zip(xq, wq)):
            P[i, q] = wqq*sum((k + 0.5)*_legendre(k, float(xi))[0] *
                              _legendre(k, float(xqq))[0]
                              for k in range(len(x)))
    return P


class HexShape:
    """Hex FR element of order p with Gauss-Legendre (or
    Gauss-Legendre-Lobatto) solution points and Gauss-Legendre flux
    points."""
    name = 'hex'
    ndims = 3
    nfaces = 6

    def __init__(self, order, nq1=None, nf1=None, upts_name=None):
        """nq1: points per direction of the Gauss-Legendre quadrature rule
        of flux anti-aliasing (default order + 2, the reference's
        ``(order + 2)**3`` hex rule, ``pyfr/shapes.py:147-176``).  nf1:
        surface-flux anti-aliasing, the flux points of each face are the
        nf1 x nf1 Gauss-Legendre rule (``pyfr/shapes.py:196-224``,
        ``:297-303``; the reference's default is nf1 = order + 2)."""
        if order < 1:
            raise ValueError('Order must be >= 1')
        self.order = order
        self.n = n = order + 1
        # solution points ([solver-elements-hex] soln-pts): Gauss-Legendre,
        # whose lines through the solution points end at the standard face
        # points, or Gauss-Legendre-Lobatto, whose do not
        self.upts_name = upts_name or 'gauss-legendre'
        if self.upts_name == 'gauss-legendre':
            self.x1d, self.w1d = gauss_legendre(n)
        elif self.upts_name == 'gauss-legendre-lobatto':
            self.x1d, self.w1d = gauss_legendre_lobatto(n)
        else:
            raise NotImplementedError(f'hex soln-pts {self.upts_name} is '
                                      'not built')
        self.on_lines = self.upts_name == 'gauss-legendre'
        # the standard face points, which carry the face bases
        self.xs1d = gauss_legendre(n)[0]
        self.nupts = n**3
        self.nf1 = nf = nf1 or n
        self.surf_flux = nf1 is not None
        self.xf1d = gauss_legendre(nf)[0]
        self.nfacefpts = [nf*nf]*6
        self.nfpts = 6*nf*nf
        self.nq1 = nq1 or order + 2
        self.nqpts = self.nq1**3

    # -- point sets
    def uidx(self, i, j, k):
        return i + self.n*(j + self.n*k)

    @cached_property
    def upts(self):
        x, n = self.x1d, self.n
        return np.array([(x[i], x[j], x[k])
                         for k in range(n) for j in range(n) for i in range(n)])

    @cached_property
    def upts_wts(self):
        w, n = self.w1d, self.n
        return np.array([w[i]*w[j]*w[k]
                         for k in range(n) for j in range(n) for i in range(n)])

    @cached_property
    def fpts(self):
        x, n = self.xf1d, self.nf1
        pts = []
        for nax, side, sax, tax in HEX_FACES:
            for b in range(n):
                for a in range(n):
                    p = [0.0, 0.0, 0.0]
                    p[nax], p[sax], p[tax] = float(side), x[a], x[b]
                    pts.append(p)
        return np.array(pts)

    @cached_property
    def norm_fpts(self):
        out = np.zeros((self.nfpts, 3))
        m = self.nf1**2
        for f, (nax, side, _, _) in enumerate(HEX_FACES):
            out[f*m:(f + 1)*m, nax] = side
        return out

    @cached_property
    def facefpts(self):
        m = self.nf1**2
        return [list(range(f*m, (f + 1)*m)) for f in range(6)]

    def nodal_basis_at(self, pts):
        """(npts, nupts): the solution's nodal basis at reference points."""
        pts = np.atleast_2d(pts)
        L = [lagrange_at(self.x1d, pts[:, d]) for d in range(3)]
        return np.einsum('pk,pj,pi->pkji', L[2], L[1], L[0]).reshape(
            len(pts), -1)

    def fpt_line(self, f, a, b):
        """Solution points on the line normal to face f through its flux point
        (a, b), ordered along the normal axis, and that axis."""
        nax, side, sax, tax = HEX_FACES[f]
        line = []
        for l in range(self.n):
            c = [0, 0, 0]
            c[nax], c[sax], c[tax] = l, a, b
            line.append(self.uidx(*c))
        return line, nax, side

    # -- 1-D building blocks
    @cached_property
    def lend(self):
        """{-1: l_j(-1), +1: l_j(+1)}"""
        v = lagrange_at(self.x1d, np.array([-1.0, 1.0]))
        return {-1: v[0], +1: v[1]}

    @cached_property
    def dmat(self):
        return diff_matrix(self.x1d)

    @cached_property
    def hend(self):
        return {s: legendre_kernel_end(self.x1d, s) for s in (-1, +1)}

    # -- operators (layouts of pyfr/shapes.py:89-123)
    @cached_property
    def m0(self):
        n = self.n
        if self.surf_flux or not self.on_lines:
            return self._m0_surf
        m = np.zeros((self.nfpts, self.nupts))
        for f in range(6):
            for b in range(n):
                for a in range(n):
                    line, _, side = self.fpt_line(f, a, b)
                    m[f*n*n + a + n*b, line] = self.lend[side]
        return m

    @cached_property
    def _m0_surf(self):
        """Interpolation to face points that are not on the solution-point
        lines: the normal end value times the two in-plane 1-D Lagrange
        interpolants (the tensor form of ``ubasis.nodal_basis_at(fpts)``,
        ``pyfr/shapes.py:89-91``)."""
        n, nf = self.n, self.nf1
        L = lagrange_at(self.x1d, self.xf1d)             # (nf, n)
        m = np.zeros((self.nfpts, self.nupts))
        for f, (nax, side, sax, tax) in enumerate(HEX_FACES):
            for b, a in it.product(range(nf), repeat=2):
                row = f*nf*nf + a + nf*b
                for l, s, t in it.product(range(n), repeat=3):
                    c = [0, 0, 0]
                    c[nax], c[sax], c[tax] = l, s, t
                    m[row, self.uidx(*c)] = (self.lend[side][l]*L[a, s] *
                                             L[b, t])
        return m

    @cached_property
    def m1(self):
        n, Nu = self.n, self.nupts
        m = np.zeros((Nu, 3*Nu))
        D = self.dmat
        for k, j, i in it.product(range(n), repeat=3):
            u = self.uidx(i, j, k)
            c = (i, j, k)
            for d in range(3):
                for l in range(n):
                    cc = list(c)
                    cc[d] = l
                    m[u, d*Nu + self.uidx(*cc)] = D[c[d], l]
        return m

    @cached_property
    def m2(self):
        m = self.norm_fpts[..., None]*self.m0[:, None, :]
        return m.reshape(self.nfpts, -1)

    @cached_property
    def m3(self):
        n = self.n
        m = np.zeros((self.nupts, 6*n*n))
        if self.on_lines:
            for f in range(6):
                for b in range(n):
                    for a in range(n):
                        line, _, side = self.fpt_line(f, a, b)
                        m[line, f*n*n + a + n*b] = self.hend[side]
        else:
            # the lifting of face point (a, b): the 1-D kernel end function
            # along the normal times the face's Lagrange basis on the
            # standard face points, at the solution points' coordinates
            Ls = lagrange_at(self.xs1d, self.x1d)             # (n, n)
            for f, (nax, side, sax, tax) in enumerate(HEX_FACES):
                for b, a in it.product(range(n), repeat=2):
                    col = f*n*n + a + n*b
                    for l, s_, t in it.product(range(n), repeat=3):
                        c = [0, 0, 0]
                        c[nax], c[sax], c[tax] = l, s_, t
                        m[self.uidx(*c), col] = (self.hend[side][l] *
                                                 Ls[s_, a]*Ls[t, b])
        if self.surf_flux:
            # correction from the standard face points, fed by the L2
            # projection of each face's quadrature-point values onto the
            # order-p face basis: m3 @ block_diag(proj_l2(iqrule, facebasis))
            # (pyfr/shapes.py:103-113, :18-19)
            P = l2_projection_1d(self.xs1d, self.nf1)
            fp = np.kron(P, P)
            bd = np.zeros((6*n*n, self.nfpts))
            for f in range(6):
                bd[f*n*n:(f + 1)*n*n, f*fp.shape[1]:(f + 1)*fp.shape[1]] = fp
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

    # -- flux anti-aliasing (pyfr/shapes.py:125-135): quadrature points, M7
    #    (solution -> quadrature points), M8 (L2 projection back), M9
    @cached_property
    def qpts(self):
        x, n = gauss_legendre(self.nq1)[0], self.nq1
        return np.array([(x[i], x[j], x[k])
                         for k in range(n) for j in range(n) for i in range(n)])

    @cached_property
    def m7(self):
        L = lagrange_at(self.x1d, gauss_legendre(self.nq1)[0])  # (nq, n)
        return np.kron(L, np.kron(L, L))

    @cached_property
    def m8(self):
        """The L2 projection onto the order-p space: values at the solution
        points of sum_k psi_k(x) sum_q w_q psi_k(x_q) f_q with the
        orthonormal Legendre products psi_k, a Kronecker product of 1-D
        projections for the tensor-product space."""
        P = l2_projection_1d(self.x1d, self.nq1)
        return np.kron(P, np.kron(P, P))

    @property
    def m9(self):
        m8, (a, b) = self.m8, self.m8.shape
        out = np.zeros((3*a, 3*b))
        for d in range(3):
            out[d*a:(d + 1)*a, d*b:(d + 1)*b] = m8
        return out

    def opmat(self, expr):
        """Evaluate an operator expression ('M0', 'M1 - M3*M2', ...) the way
        ``BaseShape.opmat`` (``pyfr/shapes.py:79-87``) does, flushing
        |x| < 1e-10 to zero.  Products are evaluated with correctly rounded
        sums so the result is the same on every host."""
        e = expr.lower().replace(' ', '')
        if not re.match(r'[m0-9\-+*()]+$', e):
            raise ValueError('Invalid operator matrix expression')
        return flush(_OpExpr(self, e).parse())

    # -- linear element geometry helpers
    @staticmethod
    def linspts():
        """Vertices of the reference hex (``TensorProdShape.std_ele(1)``)."""
        pts1d = np.linspace(-1, 1, 2)
        return np.array([p[::-1] for p in it.product(pts1d, repeat=3)])


class _OpExpr:
    """Recursive-descent evaluator for operator expressions over m0..m9 with
    +, -, * (matrix product) and parentheses."""

    def __init__(self, shape, s):
        self.shape, self.s, self.i = shape, s, 0

    def parse(self):
        v = self._sum()
        if self.i != len(self.s):
            raise ValueError('Invalid operator matrix expression')
        return v

    def _peek(self):
        return self.s[self.i] if self.i < len(self.s) else ''

    def _sum(self):
        v = self._prod()
        while self._peek() in ('+', '-'):
            op = self.s[self.i]
            self.i += 1
            r = self._prod()
            v = v + r if op == '+' else v - r
        return v

    def _prod(self):
        v = self._atom()
        while self._peek() == '*':
            self.i += 1
            v = sparse_matmul(v, self._atom())
        return v

    def _atom(self):
        if self._peek() == '(':
            self.i += 1
            v = self._sum()
            if self._peek() != ')':
                raise ValueError('Invalid operator matrix expression')
            self.i += 1
            return v
        m = re.match(r'm(\d+)', self.s[self.i:])
        if not m:
            raise ValueError('Invalid operator matrix expression')
        self.i += len(m.group(0))
        return np.array(getattr(self.shape, f'm{m.group(1)}'), dtype=float)


SHAPES = {'hex': HexShape}
