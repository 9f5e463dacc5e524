"""GPU parity of the dense (tet/prism) path: matrix-core operator kernels
against NumPy, and the Euler RHS / RK4 trajectories on periodic meshes of
tets and prisms (with tet-prism interfaces) against the oracle.

Tolerances: FP64 as for the hex path (operators 1e-13 relative L2, one RHS
max(1e-12, 10 x the oracle's own 1e-16-input-noise floor), trajectories
1e-12); FP32 operators 2e-6 and FP32 RHS / trajectories 1e-4 relative L2
(FP32 rounding amplified by the flux cancellation the FP64 floor measures).
"""
import numpy as np
import pytest

from tests.fr_cases import eval_ics, make_simplex_cfg, mode_ics, rel_l2

pytestmark = pytest.mark.gpu

def _tables(name, p):
    from pyfr_amd.quadtables import TABLES
    nu = ((p + 1)*(p + 2)*(p + 3)//6 if name == 'tet'
          else (p + 1)**2*(p + 2)//2)
    key = (('tet', 'shunn-ham', nu) if name == 'tet' else
           ('pri', 'williams-shunn~gauss-legendre', nu))
    return {'upts': TABLES[key],
            'tri_fpts': TABLES[('tri', 'williams-shunn',
                                (p + 1)*(p + 2)//2)][0],
            'tri_q36': TABLES[('tri', 'williams-shunn', 36)]}


@pytest.mark.parametrize('precision', ['double', 'single'])
@pytest.mark.parametrize('m,k', [(60, 35), (35, 105), (105, 75), (75, 225),
                                 (13, 9), (112, 3)])
@pytest.mark.parametrize('alpha,beta', [(1.0, 0.0), (1.0, 1.0), (0.5, 2.0),
                                        (-2.0, 0.0)])
def test_mul_dense(backend_factory, precision, m, k, alpha, beta):
    be = backend_factory(precision=precision)
    rng = np.random.default_rng(m*1000 + k)
    A = rng.standard_normal((m, k))
    neles = 200
    bnp = rng.standard_normal((k, 5, neles))
    cnp = rng.standard_normal((m, 5, neles))
    a = be.const_matrix(A)
    b = be.matrix((k, 5, neles), bnp, tags={'align'})
    c = be.matrix((m, 5, neles), cnp, tags={'align'})
    print(f'\n[dense mul] {precision} {m}x{k} alpha={alpha} beta={beta}',
          flush=True)
    kern = be.kernel('mul', a, b, out=c, alpha=alpha, beta=beta)
    assert kern.kind == 'dense'
    be.run_kernels([kern], wait=True)
    ref = alpha*(A @ bnp.reshape(k, -1)).reshape(m, 5, neles) + beta*cnp
    tol = 1e-13 if precision == 'double' else 2e-6
    assert rel_l2(c.get(), ref) < tol


class _Plan2:
    """A raw two-input plan through the C ABI (for epilogue tests)."""

    def __init__(self, be, A0, A1):
        import ctypes as ct
        self.lib = be.lib
        self.h0, self.h1 = np.ascontiguousarray(A0), np.ascontiguousarray(A1)
        plan, kind = ct.c_void_p(), ct.c_int()
        dt = 0 if be.fpdtype == np.float64 else 1
        self.lib.mul2_plan_create(ct.byref(plan), A0.shape[0], A0.shape[1],
                                  A1.shape[1], self.h0.ctypes.data,
                                  self.h1.ctypes.data, dt, ct.byref(kind))
        self.plan, self.kind = plan.value, kind.value

    def __del__(self):
        self.lib.mul2_plan_destroy(self.plan)


def _operands(be, m, k0, k1, neles, seed):
    rng = np.random.default_rng(seed)
    A0, A1 = rng.standard_normal((m, k0)), rng.standard_normal((m, k1))
    b0 = be.matrix((k0, 5, neles), rng.standard_normal((k0, 5, neles)),
                   tags={'align'})
    b1 = be.matrix((k1, 5, neles), rng.standard_normal((k1, 5, neles)),
                   tags={'align'})
    c = be.matrix((m, 5, neles), rng.standard_normal((m, 5, neles)),
                  tags={'align'})
    return A0, A1, b0, b1, c, rng


@pytest.mark.parametrize('precision', ['double', 'single'])
@pytest.mark.parametrize('variant', [0, 1, 2, 4])
@pytest.mark.parametrize('shape', [(35, 105, 60), (75, 225, 105),
                                   (10, 30, 24)])
def test_dense_variants_agree(backend_factory, precision, variant, shape):
    """Every kernel variant the tuner may pick gives the same product."""
    be = backend_factory(precision=precision)
    m, k0, k1 = shape
    neles = 256
    A0, A1, b0, b1, c, _ = _operands(be, m, k0, k1, neles, sum(shape))
    p = _Plan2(be, A0, A1)
    assert p.kind == 2
    print(f'\n[dense variant] {precision} {shape} v={variant}', flush=True)
    be.lib.mul2_plan_set_variant(p.plan, variant)
    be.lib.mul2(p.plan, c.ncol, b0.data, b0.leaddim, b1.data, b1.leaddim,
                c.data, c.leaddim, 0, None, 0, None, 0.0, 0.0, 0, be.stream)
    be.wait()
    ref = ((A0 @ b0.get().reshape(k0, -1) + A1 @ b1.get().reshape(k1, -1))
           .reshape(m, 5, neles))
    tol = 1e-13 if precision == 'double' else 2e-6
    assert rel_l2(c.get(), ref) < tol


@pytest.mark.parametrize('precision', ['double', 'single'])
@pytest.mark.parametrize('per_element', [True, False])
@pytest.mark.parametrize('ep,last', [(1, 0), (2, 0), (2, 1)])
def test_dense_epilogues(backend_factory, precision, per_element, ep, last):
    """negdivconf and the RK45 stage epilogues of the dense kernel, with
    1/|J| per point or one per element (row stride 0)."""
    be = backend_factory(precision=precision)
    m, k0, k1, neles = 35, 105, 60, 200
    A0, A1, b0, b1, c, rng = _operands(be, m, k0, k1, neles, 11)
    rc = (rng.uniform(0.5, 2.0, (1, neles)) if per_element else
          rng.uniform(0.5, 2.0, (m, neles)))
    rcp = be.const_matrix(rc, tags={'align'})
    xa = be.matrix((m, 5, neles), rng.standard_normal((m, 5, neles)),
                   tags={'align'})
    c0, x0 = c.get(), xa.get()
    p = _Plan2(be, A0, A1)
    print(f'\n[dense epilogue] {precision} ep={ep} last={last} '
          f'per_element={per_element}', flush=True)
    dta, dtb = 0.3, -0.7
    be.lib.mul2(p.plan, c.ncol, b0.data, b0.leaddim, b1.data, b1.leaddim,
                c.data, c.leaddim, ep, rcp.data,
                0 if per_element else rcp.leaddim, xa.data, dta, dtb, last,
                be.stream)
    be.wait()
    acc = ((A0 @ b0.get().reshape(k0, -1) + A1 @ b1.get().reshape(k1, -1))
           .reshape(m, 5, neles))
    r = np.broadcast_to(rc, (m, neles))[:, None, :]
    t = -r*acc
    tol = 1e-13 if precision == 'double' else 2e-6
    if ep == 1:
        assert rel_l2(c.get(), t) < tol
    elif not last:
        assert rel_l2(xa.get(), x0 + dta*t) < tol
        assert rel_l2(c.get(), x0 + dtb*t) < tol
    else:
        assert rel_l2(xa.get(), x0 + dtb*t) < tol
        assert np.array_equal(c.get(), c0)


# ---------------------------------------------------------------------------
# RHS and trajectories on tet/prism meshes
# ---------------------------------------------------------------------------

def _mixed(be, n, nt, npr, order, scheme='rk4', dt=1e-3, precision=None):
    from oracle.simplex import OracleMixedSystem
    from pyfr_amd.mesh import SimplexBoxMesh
    from pyfr_amd.solvers import get_integrator
    ics = mode_ics(seed=3)
    cfg = make_simplex_cfg(order, precision or be.precision, scheme, dt, ics)
    cfg.set('backend-hip', 'fused', str(be.fused))
    mesh = SimplexBoxMesh.stacked(n, nt, npr)
    integ = get_integrator(be, mesh, cfg)
    L = 2*np.pi
    Lz = (nt + npr)*L/n
    verts = {et: v.transpose(0, 2, 1) for et, v in mesh.spts.items()}
    idx = {et: np.arange(v.shape[1]) for et, v in mesh.spts.items()}
    orc = OracleMixedSystem(verts, order, {et: _tables(et, order)
                                           for et in verts},
                            [((L, L, Lz), idx)])
    return integ, orc, orc.ics(eval_ics(ics))


def _cat(d, names):
    return np.concatenate([d[n].ravel() for n in names])


def _noise_floor(orc, u, seed=0):
    rng = np.random.default_rng(seed)
    up = {k: v*(1 + 1e-16*rng.standard_normal(v.shape)) for k, v in u.items()}
    names = list(u)
    return rel_l2(_cat(orc.rhs(up), names), _cat(orc.rhs(u), names))


@pytest.mark.parametrize('fused', ['flux', 'div', 'none'])
@pytest.mark.parametrize('n,nt,npr,order', [(3, 2, 1, 1), (3, 2, 1, 2),
                                            (3, 2, 1, 3), (3, 2, 1, 4),
                                            (2, 2, 0, 4), (2, 0, 2, 3)])
def test_simplex_rhs_parity(backend_factory, n, nt, npr, order, fused):
    """fused: 'flux' = tdisf folded into the divergence kernel, 'div' =
    the fused divergence on a stored flux, 'none' = the reference's kernel
    sequence."""
    be = backend_factory(fused=fused != 'none', fuse_flux=fused == 'flux')
    print(f'\n[simplex rhs] n={n} tets={nt} prisms={npr} p={order} '
          f'fused={fused}', flush=True)
    integ, orc, u0 = _mixed(be, n, nt, npr, order)
    sysm = integ.system
    names = sysm.ele_types
    got0 = dict(zip(names, sysm.ele_scal_upts(0)))
    assert rel_l2(_cat(got0, names), _cat(u0, names)) < 1e-14
    sysm.rhs(0.0, 0, 1)
    got = dict(zip(names, sysm.ele_scal_upts(1)))
    ref = orc.rhs(u0)
    tol = max(1e-12, 10*_noise_floor(orc, u0))
    err = rel_l2(_cat(got, names), _cat(ref, names))
    print(f'[simplex rhs] rel L2 {err:.3e} (tol {tol:.3e})', flush=True)
    assert err < tol


@pytest.mark.parametrize('precision', ['double', 'single'])
@pytest.mark.parametrize('variant', [1, 2, 4, 17, 18, 20, 32, 64, 65])
def test_flux_variants_parity(backend_factory, precision, variant):
    """Every variant of the flux-fused divergence kernel the tuner may pick
    (1/2/4 waves per block, +16 LDS-staged operator, 32 block-cooperative
    flux through LDS, 64 the FP32 4x4x1 MFMA form, 65 the same with an
    element block's five wavefronts in one workgroup) against the oracle,
    at p = 3 (tet 20 / prism 40 solution points) and p = 4 (35 / 75)."""
    from pyfr_amd._lib import FrHipError
    from pyfr_amd.backend import MulProvider
    # no graphs: a kernel reads its plan's variant at each launch
    be = backend_factory(precision=precision, graphs=False, fused=True,
                         fuse_flux=True)
    mp = next(p for p in be._providers if isinstance(p, MulProvider))
    for order in (3, 4):
        print(f'\n[flux variant] {precision} v={variant} p={order}',
              flush=True)
        integ, orc, u0 = _mixed(be, 3, 2, 1, order)
        sysm = integ.system
        names = sysm.ele_types
        sysm.rhs(0.0, 0, 1)       # builds (and tunes) the kernels
        plans = [p for p, _, _ in mp._plans3.values()]
        assert plans
        taken = 0
        for plan in plans:
            try:
                be.lib.tdivf_plan_set_variant(plan, variant)
                taken += 1
            except FrHipError:
                # LDS-staged operators larger than the LDS budget are
                # refused (the FP64 ones at p = 4), and the 4x4x1 form
                # exists in FP32 only; the rest must run
                assert variant & 16 or (variant in (64, 65) and
                                        precision == 'double'), variant
        if not taken:
            continue
        for b in sysm.ele_banks:      # the second RHS must rewrite it all
            b[1].set(np.full(b[1].ioshape, np.nan))
        sysm.rhs(0.0, 0, 1)
        got = dict(zip(names, sysm.ele_scal_upts(1)))
        ref = orc.rhs(u0)
        tol = (max(1e-12, 10*_noise_floor(orc, u0))
               if precision == 'double' else 1e-4)
        err = rel_l2(_cat(got, names), _cat(ref, names))
        print(f'[flux variant] rel L2 {err:.3e} (tol {tol:.3e})', flush=True)
        assert err < tol
        del integ, sysm


def test_simplex_rk4_trajectory(backend_factory):
    be = backend_factory()
    integ, orc, u0 = _mixed(be, 3, 2, 1, 3, dt=2e-3)
    names = integ.system.ele_types
    integ.run_steps(10, 2e-3)
    u = u0
    for _ in range(10):
        u = _rk4_dict(orc.rhs, u, 2e-3)
    err = rel_l2(_cat(dict(zip(names, integ.soln)), names), _cat(u, names))
    print(f'\n[simplex rk4] 10 steps rel L2 {err:.3e}', flush=True)
    assert err < 1e-12


def _rk4_dict(rhs, u, dt, step=None):
    """oracle.rhs.rk4_step (or step) over dicts of per-type arrays."""
    from oracle.rhs import rk4_step
    step = step or rk4_step
    names = list(u)
    shapes = [u[n].shape for n in names]
    sizes = [u[n].size for n in names]

    def flat(d):
        return np.concatenate([d[n].ravel() for n in names])

    def unflat(v):
        out, o = {}, 0
        for n, sh, sz in zip(names, shapes, sizes):
            out[n] = v[o:o + sz].reshape(sh)
            o += sz
        return out
    return unflat(step(lambda v: flat(rhs(unflat(v))), flat(u), dt))


@pytest.mark.parametrize('precision,variant,order', [
    ('double', None, 3), ('single', None, 3), ('single', 64, 3),
    ('single', 65, 3), ('single', 64, 4), ('single', 65, 4)])
def test_simplex_rk4_fused_stages(backend_factory, precision, variant,
                                  order):
    """RK4 with each stage's register updates in the flux-fused divergence
    epilogue against the reference kernel sequence (separate axnpby),
    10 steps on a mixed mesh; the tuner's kernel variants, or the 4x4x1
    MFMA form (64 / 65, its RK4 epilogue batching the rows' loads ahead
    of their stores) forced on every flux-fused plan."""
    from pyfr_amd.backend import MulProvider
    # forced variants: no graphs, a kernel reads its plan's variant at
    # each launch
    be = backend_factory(precision=precision, graphs=variant is None)
    integ, _, _ = _mixed(be, 3, 2, 1, order, dt=2e-3)
    assert integ.system.fused_rk4
    if variant is not None:
        integ.system.rhs(0.0, 0, 1)       # builds (and tunes) the plans
        mp = next(p for p in be._providers if isinstance(p, MulProvider))
        plans = [p for p, _, _ in mp._plans3.values()]
        assert plans
        for plan in plans:
            be.lib.tdivf_plan_set_variant(plan, variant)
    integ.run_steps(10, 2e-3)
    ref_be = backend_factory(precision=precision, fused=False)
    ref, _, _ = _mixed(ref_be, 3, 2, 1, order, dt=2e-3)
    assert not ref.system.fused_rk4
    ref.run_steps(10, 2e-3)
    names = integ.system.ele_types
    err = rel_l2(_cat(dict(zip(names, integ.soln)), names),
                 _cat(dict(zip(names, ref.soln)), names))
    print(f'\n[rk4 fused stages] {precision} rel L2 {err:.3e}', flush=True)
    assert err < (1e-14 if precision == 'double' else 1e-5)
    assert integ.nrhsevals == ref.nrhsevals == 40


@pytest.mark.parametrize('fuse_flux', [True, False])
def test_simplex_rk45_trajectory(backend_factory, fuse_flux):
    """RK45 with the stage update fused into each element type's
    divergence kernel on a mixed tet/prism mesh: every type's kernel takes
    the stage arguments."""
    from oracle.rhs import rk45_step
    be = backend_factory(fuse_flux=fuse_flux)
    integ, orc, u0 = _mixed(be, 3, 2, 1, 3, scheme='rk45', dt=2e-3)
    assert integ.system.fused_stages and len(integ.system.ele_types) == 2
    names = integ.system.ele_types
    integ.run_steps(10, 2e-3)
    u = u0
    for _ in range(10):
        u = _rk4_dict(orc.rhs, u, 2e-3, rk45_step)
    err = rel_l2(_cat(dict(zip(names, integ.soln)), names), _cat(u, names))
    print(f'\n[simplex rk45] fuse_flux={fuse_flux} 10 steps rel L2 '
          f'{err:.3e}', flush=True)
    assert err < 1e-12


def test_simplex_fp32_rhs_and_trajectory(backend_factory):
    be = backend_factory(precision='single')
    integ, orc, u0 = _mixed(be, 3, 2, 1, 4, dt=2e-3)
    sysm = integ.system
    names = sysm.ele_types
    sysm.rhs(0.0, 0, 1)
    got = dict(zip(names, sysm.ele_scal_upts(1)))
    err = rel_l2(_cat(got, names), _cat(orc.rhs(u0), names))
    print(f'\n[simplex fp32 rhs] rel L2 {err:.3e}', flush=True)
    assert err < 1e-4
    integ.run_steps(10, 2e-3)
    u = u0
    for _ in range(10):
        u = _rk4_dict(orc.rhs, u, 2e-3)
    err = rel_l2(_cat(dict(zip(names, integ.soln)), names), _cat(u, names))
    print(f'[simplex fp32 rk4] 10 steps rel L2 {err:.3e}', flush=True)
    assert err < 1e-5


def _bf16(x):
    """Round to bfloat16 (nearest even), returned as float64."""
    a = np.asarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    a = (a + 0x7fff + ((a >> 16) & 1)) >> 16 << 16
    return a.astype(np.uint32).view(np.float32).astype(np.float64)


@pytest.mark.parametrize('shape', [(64, 192, 96), (192, 64, 96),
                                   (35, 105, 60)])
def test_bf16_operand_plans(backend_factory, shape):
    """BF16-operand matrix-core plans (the precision form's BF16 run): the
    product equals NumPy's on the bfloat16-rounded operands to FP32
    accumulation error, and differs from the unrounded product by about
    the bfloat16 rounding."""
    import ctypes as ct
    be = backend_factory(precision='single')
    m, k0, k1 = shape
    neles = 128
    A0, A1, b0, b1, c, _ = _operands(be, m, k0, k1, neles, 7)
    h0, h1 = np.ascontiguousarray(A0), np.ascontiguousarray(A1)
    plan, kind = ct.c_void_p(), ct.c_int()
    be.lib.mul2_plan_create(ct.byref(plan), m, k0, k1, h0.ctypes.data,
                            h1.ctypes.data, 2, ct.byref(kind))
    try:
        assert kind.value == 3
        print(f'\n[bf16 plan] {shape}', flush=True)
        be.lib.mul2(plan.value, c.ncol, b0.data, b0.leaddim, b1.data,
                    b1.leaddim, c.data, c.leaddim, 0, None, 0, None, 0.0,
                    0.0, 0, be.stream)
        be.wait()
    finally:
        be.lib.mul2_plan_destroy(plan.value)
    B0, B1 = b0.get().reshape(k0, -1), b1.get().reshape(k1, -1)
    exact = (A0 @ B0 + A1 @ B1).reshape(m, 5, neles)
    rounded = (_bf16(A0) @ _bf16(B0) + _bf16(A1) @ _bf16(B1)).reshape(
        m, 5, neles)
    got = c.get()
    assert rel_l2(got, rounded) < 2e-6
    assert 1e-4 < rel_l2(got, exact) < 2e-2


# -- Navier-Stokes and boundary conditions on tets / prisms -----------------

# side -> (type, {option: expression}) on the p = 1/1.4, sound speed 1
# state of mode_ics
FREE = {'rho': '1 + 0.05*sin(z)', 'u': '0.2', 'v': '0.1*cos(z)',
        'w': '-0.05', 'p': '1/1.4 + 0.02*cos(y)'}
SIMPLEX_BCS = {
    'euler': {'x0': ('char-riem-inv', FREE), 'x1': ('slp-adia-wall', {}),
              'z0': ('slp-adia-wall', {}), 'z1': ('sup-out-fn', {})},
    'navier-stokes': {
        'x0': ('no-slp-isot-wall', {'u': '0.05*sin(z)', 'v': '0',
                                    'w': '0', 'cpTw': '3.4'}),
        'x1': ('no-slp-adia-wall', {}), 'z0': ('char-riem-inv', FREE),
        'z1': ('sub-out-fp', {'p': '0.7'})},
}
MU = 0.02


def _mixed_ns_bc(be, n, nt, npr, order, system, bcs=False, rsolver='rusanov',
                 beta=0.0, tau=0.0, aa=(), sorder=1, warp=None):
    from oracle.simplex import OracleMixedSystem
    from pyfr_amd.mesh import SimplexBoxMesh
    from pyfr_amd.solvers import get_integrator
    ics = mode_ics(seed=5)
    cfg = make_simplex_cfg(order, be.precision, ics=ics, rsolver=rsolver,
                           system=system, beta=beta, tau=tau)
    cfg.set('backend-hip', 'fused', str(be.fused))
    if aa:
        cfg.set('solver', 'anti-alias', ', '.join(aa))
    c = {'gamma': 1.4}
    if system == 'navier-stokes':
        cfg.set('constants', 'mu', repr(MU))
        c |= {'mu': MU, 'Pr': 0.71, 'ldg-beta': beta, 'ldg-tau': tau}
    periodic = (False, True, False) if bcs else (True, True, True)
    obcs = {}
    if bcs:
        for side, (bctype, opts) in SIMPLEX_BCS[system].items():
            sect = f'soln-bcs-{side}'
            cfg.set(sect, 'type', bctype)
            vals = {}
            for k, v in opts.items():
                cfg.set(sect, k, v)
                if k in ('cpTw',):
                    c[k] = float(eval(v))
                else:
                    vals[k] = eval_ics({k: v})[k]
            obcs[side] = (bctype, vals)
    mesh = SimplexBoxMesh.stacked(n, nt, npr, periodic=periodic,
                                  sorder=sorder, warp=warp)
    integ = get_integrator(be, mesh, cfg)
    L = 2*np.pi
    verts = {et: v.transpose(0, 2, 1) for et, v in mesh.spts.items()}
    idx = {et: np.arange(v.shape[1]) for et, v in mesh.spts.items()}
    tabs = {et: _tables(et, order) for et in verts}
    from pyfr_amd.quadtables import TABLES
    from pyfr_amd.shapes import gauss_legendre
    from pyfr_amd.simplex import AA_RULES, SF_TRI_RULES
    for et in tabs:
        if 'flux' in aa:
            tabs[et]['qrule'] = TABLES[(et, *AA_RULES[et][order])]
        if 'surf-flux' in aa:
            g, w = gauss_legendre(order + 2)
            r = range(order + 2)
            tabs[et]['iqrules'] = {
                'tri': TABLES[('tri', *SF_TRI_RULES[order])],
                'quad': (np.array([(g[i], g[j]) for j in r for i in r]),
                         np.array([w[i]*w[j] for j in r for i in r]))}
    orc = OracleMixedSystem(verts, order, tabs,
                            [((L, L, (nt + npr)*L/n), idx)], rsolver=rsolver,
                            system=system, constants=c, periodic=periodic,
                            bcs=obcs, antialias=tuple(a for a in aa
                                                      if a == 'flux'))
    return integ, orc, orc.ics(eval_ics(ics))


def _rhs_check(integ, orc, u0, tag):
    sysm = integ.system
    names = sysm.ele_types
    got0 = dict(zip(names, sysm.ele_scal_upts(0)))
    assert rel_l2(_cat(got0, names), _cat(u0, names)) < 1e-14
    sysm.rhs(0.0, 0, 1)
    got = dict(zip(names, sysm.ele_scal_upts(1)))
    ref = orc.rhs(u0)
    tol = max(1e-12, 10*_noise_floor(orc, u0))
    err = rel_l2(_cat(got, names), _cat(ref, names))
    print(f'[{tag}] rel L2 {err:.3e} (tol {tol:.3e})', flush=True)
    assert err < tol


@pytest.mark.parametrize('fused', [True, False])
@pytest.mark.parametrize('n,nt,npr,order,beta,tau', [
    (2, 1, 1, 1, 0.0, 0.0), (2, 1, 1, 2, 0.0, 0.0), (3, 2, 1, 3, 0.0, 0.0),
    (2, 1, 1, 4, 0.0, 0.0), (2, 1, 1, 3, 0.25, 0.1), (2, 2, 0, 3, 0.5, 0.0),
    (2, 1, 1, 2, -0.5, 0.1)])
def test_simplex_ns_rhs_parity(backend_factory, n, nt, npr, order, beta,
                               tau, fused):
    """NS on periodic tet/prism meshes: LDG gradient (M4 - M6 M0, M6), the
    affine fused NS flux, gradients at the flux points and the NS interface
    flux, with mu large enough that the viscous terms matter."""
    be = backend_factory(fused=fused)
    print(f'\n[simplex ns] n={n} tets={nt} prisms={npr} p={order} '
          f'beta={beta} tau={tau} fused={fused}', flush=True)
    integ, orc, u0 = _mixed_ns_bc(be, n, nt, npr, order, 'navier-stokes',
                                  beta=beta, tau=tau)
    _rhs_check(integ, orc, u0, 'simplex ns')


@pytest.mark.parametrize('fused', [True, False])
@pytest.mark.parametrize('system,order,rsolver', [
    ('euler', 2, 'rusanov'), ('euler', 3, 'hllc'), ('euler', 4, 'roe'),
    ('navier-stokes', 2, 'rusanov'), ('navier-stokes', 3, 'hll'),
    ('navier-stokes', 4, 'rusanov')])
def test_simplex_bc_rhs_parity(backend_factory, system, order, rsolver,
                               fused):
    """Boundary conditions on a tet/prism mesh periodic in y only: x sides
    and the tet bottom / prism top are boundaries (triangles and quads)."""
    be = backend_factory(fused=fused)
    print(f'\n[simplex bc] {system} p={order} {rsolver} fused={fused}',
          flush=True)
    integ, orc, u0 = _mixed_ns_bc(be, 3, 2, 1, order, system, bcs=True,
                                  rsolver=rsolver)
    assert len(integ.system._bc_inters) == 4
    _rhs_check(integ, orc, u0, 'simplex bc')


@pytest.mark.parametrize('fused', ['flux', 'div', 'none'])
@pytest.mark.parametrize('system,n,nt,npr,order', [
    ('euler', 2, 1, 1, 1), ('euler', 2, 1, 1, 2), ('euler', 3, 2, 1, 3),
    ('euler', 2, 1, 1, 4), ('navier-stokes', 2, 1, 1, 2),
    ('navier-stokes', 2, 1, 1, 3), ('navier-stokes', 2, 2, 0, 4)])
def test_simplex_flux_aa_rhs_parity(backend_factory, system, n, nt, npr,
                                    order, fused):
    """Flux anti-aliasing on tets and prisms: u (and, for NS, the physical
    gradient) at the element quadrature rule, the flux formed there with
    the affine metric, (M1 - M3 M2) M9 divergence (dense operators);
    'flux' also checks that the flux-folded kernel steps aside."""
    be = backend_factory(fused=fused != 'none', fuse_flux=fused == 'flux')
    print(f'\n[simplex aa] {system} n={n} tets={nt} prisms={npr} '
          f'p={order} fused={fused}', flush=True)
    integ, orc, u0 = _mixed_ns_bc(be, n, nt, npr, order, system,
                                  aa=('flux',))
    for e in integ.system.ele_map.values():
        assert e.fluxaa and 'tdivf_flux' not in e.kernels
    _rhs_check(integ, orc, u0, 'simplex aa')


@pytest.mark.parametrize('fused', [True, False])
@pytest.mark.parametrize('system,order,aa,bcs', [
    ('euler', 1, ('surf-flux',), False), ('euler', 2, ('surf-flux',), False),
    ('euler', 3, ('surf-flux',), False), ('euler', 4, ('surf-flux',), False),
    ('navier-stokes', 2, ('surf-flux',), False),
    ('navier-stokes', 3, ('flux', 'surf-flux'), False),
    ('euler', 2, ('flux', 'surf-flux'), True),
    ('navier-stokes', 2, ('surf-flux',), True)])
def test_simplex_surf_flux_rhs_parity(backend_factory, system, order, aa,
                                      bcs, fused):
    """Surface-flux anti-aliasing on tets and prisms: flux points at the
    face rules (triangles and quads), dense M0 / M3, interfaces and
    boundaries on the new points; alone and with flux anti-aliasing."""
    be = backend_factory(fused=fused)
    print(f'\n[simplex sf] {system} p={order} aa={aa} bcs={bcs} '
          f'fused={fused}', flush=True)
    integ, orc, u0 = _mixed_ns_bc(be, 3 if bcs else 2, 2 if bcs else 1, 1,
                                  order, system, bcs=bcs, aa=aa)
    _rhs_check(integ, orc, u0, 'simplex sf')


@pytest.mark.parametrize('precision', ['single', 'double'])
def test_flux_fused_in_place(backend_factory, precision):
    """An RHS in place (u and the output the same bank, as the reference's
    RK steppers call it): the flux-fused kernel reads u and writes its
    output in one launch, so an in-place call runs the block-cooperative
    form, whose barriers order all reads of u before any write, whatever
    the plan's variant; so does an RK45 stage whose accumulator is the
    input (xa == u).  Checked on 8,192 element blocks (more blocks than
    the device holds at once) for every variant: the in-place result
    equals, bit for bit, that form's out-of-place result."""
    import ctypes as ct
    from pyfr_amd._lib import FrHipError
    be = backend_factory(precision=precision)
    lib = be.lib
    dt = np.float32 if precision == 'single' else np.float64
    code = 1 if precision == 'single' else 0
    m, nu, nf = 20, 20, 40                   # tet p = 3 shapes
    neles = 64*8192
    ncol = 5*neles
    rng = np.random.default_rng(7)
    a0 = rng.standard_normal((m, 3*nu))
    a1 = rng.standard_normal((m, nf))
    plan = ct.c_void_p()
    lib.tdivf_plan_create(ct.byref(plan), m, nu, nf, a0.ctypes.data,
                          a1.ctypes.data, code)
    u = np.empty((nu, neles//64, 5, 64), dtype=dt)
    u[:, :, 0] = 1 + 0.1*rng.random((nu, neles//64, 64))
    u[:, :, 1:4] = 0.1*rng.standard_normal((nu, neles//64, 3, 64))
    u[:, :, 4] = 2.5 + 0.1*rng.random((nu, neles//64, 64))
    u = u.reshape(nu, ncol)
    sm = np.zeros((9, neles), dtype=dt)
    sm[0] = sm[4] = sm[8] = 1.0
    sm += (0.01*rng.standard_normal((9, neles))).astype(dt)
    fn = rng.standard_normal((nf, ncol)).astype(dt)
    rcp = (1 + 0.1*rng.random((1, neles))).astype(dt)
    sz = u.nbytes
    bufs = {}

    def dev(name, arr):
        p = ct.c_void_p()
        lib.malloc(ct.byref(p), arr.nbytes)
        lib.memcpy_htod(p, arr.ctypes.data, arr.nbytes, None)
        bufs[name] = p
        return p

    du, dsm, dfn, drcp = (dev('u', u), dev('sm', sm), dev('fn', fn),
                          dev('rcp', rcp))
    dc = ct.c_void_p()
    lib.malloc(ct.byref(dc), sz)
    bufs['c'] = dc
    dw = ct.c_void_p()
    lib.malloc(ct.byref(dw), sz)
    bufs['w'] = dw

    def run(src, dst):
        lib.tdivf_flux_euler(plan, ncol, src, ncol, dsm, neles, dfn, ncol,
                             dst, ncol, 1, drcp, 0, None, 0.0, 0.0, 0, 1.4,
                             None)
        lib.stream_sync(None)
        out = np.empty_like(u)
        lib.memcpy_dtoh(out.ctypes.data, dst, sz, None)
        return out

    variants = (1, 2, 4, 32) + ((64, 65) if precision == 'single' else ())
    try:
        lib.tdivf_plan_set_variant(plan, 32)
        want = run(du, dc)
        assert np.isfinite(want).all()
        for v in variants:
            print(f'\n[in place] {precision} variant {v}', flush=True)
            lib.tdivf_plan_set_variant(plan, v)
            lib.memcpy_dtod(dw, du, sz, None)
            got = run(dw, dw)
            assert np.array_equal(got, want), v

        # RK45 stage epilogues with the accumulator the input (xa == u,
        # the first stage): against the same form with xa a copy of u
        dx = ct.c_void_p()
        lib.malloc(ct.byref(dx), sz)
        bufs['x'] = dx

        def run_rk(src, xa, dst, last):
            lib.tdivf_flux_euler(plan, ncol, src, ncol, dsm, neles, dfn,
                                 ncol, dst, ncol, 2, drcp, 0, xa, 0.3, -0.7,
                                 last, 1.4, None)
            lib.stream_sync(None)
            outs = []
            for q in (dst, xa):
                o = np.empty_like(u)
                lib.memcpy_dtoh(o.ctypes.data, q, sz, None)
                outs.append(o)
            return outs

        for last in (0, 1):
            lib.tdivf_plan_set_variant(plan, 32)
            lib.memcpy_dtod(dx, du, sz, None)
            want_c, want_x = run_rk(du, dx, dc, last)
            for v in variants:
                print(f'[xa == u] {precision} variant {v} last={last}',
                      flush=True)
                lib.tdivf_plan_set_variant(plan, v)
                lib.memcpy_dtod(dw, du, sz, None)
                lib.memcpy_dtod(dx, du, sz, None)
                got_c, got_x = run_rk(dw, dw, dx, last)
                assert np.array_equal(got_x, want_x), v
                if not last:
                    assert np.array_equal(got_c, want_c), v
    finally:
        for p in bufs.values():
            lib.free(p)
        lib.tdivf_plan_destroy(plan)


def _warp(x, y, z):
    return (x + 0.12*np.sin(y)*np.sin(z), y + 0.08*np.sin(x)*np.cos(z),
            z + 0.05*np.sin(x + y))


def _warp_bc(x, y, z):
    """Moves no point off the non-periodic x and z sides."""
    return (x, y + 0.08*np.sin(x)*np.sin(z), z)


@pytest.mark.parametrize('fused', [True, False])
@pytest.mark.parametrize('system,order,sorder,extra', [
    ('euler', 1, 2, None), ('euler', 2, 2, None), ('euler', 3, 2, None),
    ('euler', 2, 3, None), ('navier-stokes', 2, 2, None),
    ('navier-stokes', 3, 2, None), ('euler', 2, 2, 'bcs'),
    ('navier-stokes', 2, 2, 'bcs'), ('euler', 2, 2, 'flux'),
    ('navier-stokes', 2, 2, 'flux')])
def test_simplex_curved_rhs_parity(backend_factory, system, order, sorder,
                                   extra, fused):
    """Curved tets and prisms (order-k shape points of a warped box): the
    stored invariant-form metric at every point, curved tflux / NS flux,
    curved normals; with boundaries and with flux anti-aliasing."""
    be = backend_factory(fused=fused)
    print(f'\n[simplex curved] {system} p={order} k={sorder} {extra} '
          f'fused={fused}', flush=True)
    bcs = extra == 'bcs'
    integ, orc, u0 = _mixed_ns_bc(
        be, 3 if bcs else 2, 2 if bcs else 1, 1, order, system, bcs=bcs,
        aa=('flux',) if extra == 'flux' else (), sorder=sorder,
        warp=_warp_bc if bcs else _warp)
    for e in integ.system.ele_map.values():
        assert e.curved_simplex and e.mesh_regions == {'curved': e.neles}
    _rhs_check(integ, orc, u0, 'simplex curved')


def test_simplex_curved_free_stream(backend_factory):
    """A uniform state stays steady on the warped mesh (the product's own
    metric, no oracle)."""
    from pyfr_amd.mesh import SimplexBoxMesh
    from pyfr_amd.solvers import get_integrator
    be = backend_factory()
    ics = {'rho': '1.1', 'u': '0.3', 'v': '-0.2', 'w': '0.1', 'p': '0.9'}
    cfg = make_simplex_cfg(3, be.precision, ics=ics)
    integ = get_integrator(be, SimplexBoxMesh.stacked(2, 1, 1, sorder=2,
                                                      warp=_warp), cfg)
    integ.system.rhs(0.0, 0, 1)
    r = integ.system.ele_scal_upts(1)
    assert max(np.abs(a).max() for a in r) < 1e-10


@pytest.mark.parametrize('sorder', [2, 3])
def test_simplex_curved_gmsh_import(backend_factory, sorder):
    """Curved tets and prisms read from a high-order .msh file (Gmsh types
    11 / 13, 29 / 90) run the RHS of the generated mesh they were written
    from (which the oracle pins above): same shape points, same pairs."""
    import io
    from pyfr_amd.gmsh import GmshMesh
    from pyfr_amd.mesh import SimplexBoxMesh
    from pyfr_amd.solvers import get_integrator
    from tests.msh_files import mesh_to_msh
    from tests.test_gmsh import _sim_sides
    n, nt, npr, order = 2, 1, 1, 2
    L = 2*np.pi
    flat = SimplexBoxMesh.stacked(n, nt, npr, sorder=sorder)
    sides = _sim_sides(flat, (L, L, (nt + npr)*L/n))
    gen = SimplexBoxMesh.stacked(n, nt, npr, sorder=sorder, warp=_warp)
    mesh = GmshMesh(io.StringIO(mesh_to_msh(gen, sides, '2.2')))
    print(f'\n[simplex gmsh] k={sorder}', flush=True)
    out = []
    for m in (gen, mesh):
        be = backend_factory()
        cfg = make_simplex_cfg(order, be.precision, ics=mode_ics(seed=3))
        integ = get_integrator(be, m, cfg)
        for e in integ.system.ele_map.values():
            assert e.curved_simplex
        integ.system.rhs(0.0, 0, 1)
        # by name: the mesh read from a file lists its types in name order
        out.append(_cat(dict(zip(integ.system.ele_types,
                                 integ.system.ele_scal_upts(1))),
                        ['tet', 'pri']))
    err = rel_l2(out[1], out[0])
    print(f'[simplex gmsh] rel L2 {err:.3e}', flush=True)
    assert err < 1e-13


def test_concurrent_graph_branches_tet_prism(backend_factory):
    """Tets and prisms (the dense path, RK4 stages fused): the graphs with
    per-type branches against the eager one-stream run, bit for bit."""
    print('\n[concurrent graph] tet/prism rk4', flush=True)
    runs = []
    for graphs in (True, False):
        be = backend_factory(graphs=graphs)
        integ, _, _ = _mixed(be, 3, 2, 1, 3, dt=2e-3)
        assert integ.system.fused_rk4
        out = []
        for _ in range(2):
            integ.run_steps(5, 2e-3)
            out.append(np.concatenate([a.ravel() for a in integ.soln]))
        runs.append(out)
    for a, b in zip(*runs):
        assert np.array_equal(a, b)


@pytest.mark.parametrize('scheme', ['euler', 'tvd-rk3', 'rk34'])
@pytest.mark.parametrize('precision', ['double', 'single'])
def test_simplex_other_steppers(backend_factory, scheme, precision):
    """Forward Euler, TVD-RK3 and RK34 on tets and prisms (flux-fused
    kernels; Euler and RK34 with the stage update in the epilogue, whose
    accumulator is the RHS input), 10 steps against the oracle (FP64) and
    the unfused kernel sequence (both precisions)."""
    from oracle.rhs import STEPS
    be = backend_factory(precision=precision)
    integ, orc, u0 = _mixed(be, 3, 2, 1, 3, scheme=scheme, dt=2e-3)
    names = integ.system.ele_types
    print(f'\n[simplex {scheme}] {precision}', flush=True)
    integ.run_steps(10, 2e-3)
    got = _cat(dict(zip(names, integ.soln)), names)
    ref_be = backend_factory(precision=precision, fused=False)
    ref, _, _ = _mixed(ref_be, 3, 2, 1, 3, scheme=scheme, dt=2e-3)
    ref.run_steps(10, 2e-3)
    err = rel_l2(got, _cat(dict(zip(names, ref.soln)), names))
    print(f'[simplex {scheme}] fused vs unfused rel L2 {err:.3e}',
          flush=True)
    assert err < (1e-14 if precision == 'double' else 1e-5)
    if precision == 'double':
        u = u0
        for _ in range(10):
            u = _rk4_dict(orc.rhs, u, 2e-3, STEPS[scheme])
        err = rel_l2(got, _cat(u, names))
        print(f'[simplex {scheme}] vs oracle rel L2 {err:.3e}', flush=True)
        assert err < 1e-12
