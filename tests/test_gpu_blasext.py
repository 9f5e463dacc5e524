"""GPU tests of the BLAS-extension providers against NumPy on the same
arrays: the reduction provider (``pyfr/backends/base/blasext.py:21-46``)
with the expressions the reference's nancheck, residual, integrate and
controller code hand it; scaled axnpby (``blasext.py:9-19``); rkvdh2 with its
error estimate (``rkvdh2.mako``).

Tolerances: reductions accumulate in FP64 whatever the array precision, so
sums agree with NumPy's FP64 sums of the same (FP64-converted) values to
1e-14 of the sum of the terms' magnitudes (the summation order differs, and
sums of mixed-sign terms cancel); maxima are exact.  axnpby and
rkvdh2 are elementwise with the reference's operand order: within a few
roundings of the largest term in the working precision.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NUPTS, NVARS, NELES = 27, 5, 333


def _bank(be, rng, dtype=None, lo=-1.0, hi=1.0, shape=None):
    shape = shape or (NUPTS, NVARS, NELES)
    a = rng.uniform(lo, hi, shape)
    m = be.matrix(shape, a, tags={'align'}, dtype=dtype)
    return m, m.get().astype(np.float64)


@pytest.mark.parametrize('precision', ['double', 'single', 'mixed'])
def test_reduction_plugins(backend_factory, precision):
    """sum of x (nancheck), sum of x*x (residual-style norms), weighted sum
    w*x with (nrow, neles) weights broadcast over variables (integrate),
    max fabs(x); per variable and reduced over variables (retval)."""
    be = backend_factory(precision)
    rng = np.random.default_rng(1)
    dt = be.state_dtype                     # FP64 state banks when mixed
    x, xn = _bank(be, rng, dt)
    wn = rng.uniform(0.1, 1.0, (NUPTS, NELES))
    w = be.const_matrix(wn, dtype=dt, tags={'align'})
    wn = w.get().astype(np.float64)
    ks = {'sum': be.kernel('reduction', 'sum', ['x', 'x*x'], {'x': x}),
          'wsum': be.kernel('reduction', 'sum', ['w*x'], {'x': x, 'w': w}),
          'max': be.kernel('reduction', 'max', ['fabs(x)', 'x'], {'x': x})}
    be.run_kernels(ks.values(), wait=True)
    pv = np.stack([xn.sum(axis=(0, 2)), (xn*xn).sum(axis=(0, 2))])
    mag = np.stack([np.abs(xn).sum(axis=(0, 2)), (xn*xn).sum(axis=(0, 2))])
    assert np.all(np.abs(ks['sum'].result - pv) <= 1e-14*mag)
    assert np.all(np.abs(ks['sum'].retval - pv.sum(axis=1)) <=
                  1e-14*mag.sum(axis=1))
    ref = (wn[:, None, :]*xn).sum(axis=(0, 2))
    wmag = (wn[:, None, :]*np.abs(xn)).sum(axis=(0, 2))
    assert np.all(np.abs(ks['wsum'].result[0] - ref) <= 1e-14*wmag)
    mx = np.stack([np.abs(xn).max(axis=(0, 2)), xn.max(axis=(0, 2))])
    assert np.array_equal(ks['max'].result, mx)
    assert np.array_equal(ks['max'].retval, mx.max(axis=1))


def test_reduction_controller_errest(backend_factory):
    """The PI controller's error norm (pyfr/integrators/controllers.py:
    104-119): svars rtol (bound at run time), pvars atol (one per
    variable), both norms."""
    be = backend_factory('double')
    rng = np.random.default_rng(2)
    cur, curn = _bank(be, rng)
    err, errn = _bank(be, rng, lo=-1e-4, hi=1e-4)
    atol = [1e-6, 2e-6, 3e-6, 4e-6, 5e-6]
    e = 'err / (atol + rtol*fabs(curr))'
    kl2 = be.kernel('reduction', 'sum', [f'({e})*({e})'],
                    {'curr': cur, 'err': err}, svars=['rtol'],
                    pvars={'atol': atol})
    kmax = be.kernel('reduction', 'max', [f'fabs({e})'],
                     {'curr': cur, 'err': err}, svars=['rtol'],
                     pvars={'atol': atol})
    for rtol in (1e-3, 1e-5):
        for k in (kl2, kmax):
            k.bind(rtol)
        be.run_kernels([kl2, kmax], wait=True)
        q = errn/(np.array(atol)[None, :, None] + rtol*np.abs(curn))
        np.testing.assert_allclose(kl2.retval[0], (q*q).sum(), rtol=1e-13)
        assert kmax.retval[0] == np.abs(q).max()


def test_reduction_nan_and_two_arrays(backend_factory):
    be = backend_factory('double')
    rng = np.random.default_rng(3)
    a, an = _bank(be, rng)
    b, bn = _bank(be, rng)
    kd = be.kernel('reduction', 'sum', ['a*b'], {'a': a, 'b': b})
    be.run_kernels([kd], wait=True)
    assert abs(kd.retval[0] - (an*bn).sum()) <= 1e-14*np.abs(an*bn).sum()
    an[3, 2, 17] = np.nan
    a.set(an)
    kn = be.kernel('reduction', 'sum', ['x'], {'x': a})
    be.run_kernels([kn], wait=True)
    assert np.isnan(kn.retval[0])


def test_reduction_in_graph_replays(backend_factory):
    """Captured once, replayed after the data change: every replay equals
    the eager result of the same data."""
    be = backend_factory('double')
    rng = np.random.default_rng(4)
    x, xn = _bank(be, rng)
    k = be.kernel('reduction', 'sum', ['x*x'], {'x': x})
    g = be.graph()
    g.add(k)
    g.commit()
    for it in range(3):
        xn = rng.uniform(-1, 1, xn.shape)
        x.set(xn)
        be.run_graph(g, wait=True)
        got = k.result[0]
        be.run_kernels([k], wait=True)
        assert np.array_equal(got, k.result[0]), it
        np.testing.assert_allclose(got, (xn*xn).sum(axis=(0, 2)),
                                   rtol=1e-13)


def test_reduction_refusals(backend_factory):
    be = backend_factory('double')
    rng = np.random.default_rng(5)
    x, _ = _bank(be, rng)
    y, _ = _bank(be, rng, shape=(NUPTS + 1, NVARS, NELES))
    with pytest.raises(ValueError):
        be.kernel('reduction', 'sum', ['x*y'], {'x': x, 'y': y})
    with pytest.raises(ValueError):
        be.kernel('reduction', 'prod', ['x'], {'x': x})


@pytest.mark.parametrize('precision', ['double', 'single'])
@pytest.mark.parametrize('a0', [0.0, 1.0, 0.7])
def test_axnpby_scaled(backend_factory, precision, a0):
    """axnpby.mako with in_scale/in_scale_idxs/out_scale: per-variable
    scales on the chosen operands and on the sum."""
    be = backend_factory(precision)
    rng = np.random.default_rng(6)
    mats = [_bank(be, rng) for _ in range(3)]
    ins = [0.5, 2.0, 1.5, 0.25, 3.0]
    outs = [1.0, 0.1, 10.0, 2.0, 0.5]
    a = (a0, 0.3, -1.2)
    k = be.kernel('axnpby', *[m for m, _ in mats], in_scale=ins,
                  in_scale_idxs=(0, 2), out_scale=outs)
    k.bind(*a)
    be.run_kernels([k], wait=True)
    x0, x1, x2 = [n for _, n in mats]
    isc = np.array(ins)[None, :, None]
    osc = np.array(outs)[None, :, None]
    if a0 == 0.0:
        ref = osc*(a[1]*x1 + a[2]*isc*x2)
    elif a0 == 1.0:
        ref = x0 + osc*(a[1]*x1 + a[2]*isc*x2)
    else:
        ref = osc*(a[0]*isc*x0 + a[1]*x1 + a[2]*isc*x2)
    # a few roundings of the largest term
    eps = np.finfo(np.float64 if precision == 'double' else np.float32).eps
    mag = np.abs(x0) + osc*(np.abs(a[0]*isc*x0) + np.abs(a[1]*x1) +
                            np.abs(a[2]*isc*x2))
    got = mats[0][0].get()
    assert np.all(np.abs(got - ref) <= 4*eps*mag)


@pytest.mark.parametrize('precision', ['double', 'single'])
def test_rkvdh2_errest(backend_factory, precision):
    """Five RK45 stages of rkvdh2 with the error estimate against the
    template's update written out in NumPy."""
    from pyfr_amd.solvers import RK45Integrator as R
    be = backend_factory(precision)
    rng = np.random.default_rng(7)
    bhat = [1016888040809 / 7410784769900, 11231460423587 / 58533540763752,
            -1563879915014 / 6823010717585, 606302364029 / 971179775848,
            1097981568119 / 3980877426909]
    e = [b - bh for b, bh in zip(R.b, bhat)]
    regs = [_bank(be, rng) for _ in range(4)]
    r1, r2, rold, rerr = [m for m, _ in regs]
    n1, n2, nold, nerr = [n for _, n in regs]
    be.pointwise.register('pyfr.integrators.explicit.kernels.rkvdh2')
    dt = 0.013
    tol = 1e-14 if precision == 'double' else 1e-6
    for stage in range(5):
        tpl = {'a': R.a, 'b': R.b, 'e': e, 'stage': stage, 'nstages': 5,
               'nvars': NVARS, 'errest': True}
        k = be.kernel('rkvdh2', tplargs=tpl, dims=[NUPTS, NELES], r1=r1,
                      r2=r2, rold=rold, rerr=rerr)
        k.bind(dt=dt)
        be.run_kernels([k], wait=True)
        t1, t2 = n1.copy(), n2.copy()
        if stage == 0:
            nerr, nold = dt*e[0]*t2, t1
        else:
            nerr = nerr + dt*e[stage]*t2
        if stage < 4:
            n1, n2 = t1 + dt*R.a[stage]*t2, t1 + dt*R.b[stage]*t2
        else:
            n1 = t1 + dt*R.b[stage]*t2
        for m, ref in ((r1, n1), (r2, n2), (rold, nold), (rerr, nerr)):
            got = m.get()
            assert np.max(np.abs(got - ref)) < tol*(1 + np.abs(ref).max())
        # the next stage's RHS: fresh data in r2, as the stepper supplies
        n2 = rng.uniform(-1, 1, n2.shape)
        r2.set(n2)
        n2 = r2.get().astype(np.float64)
