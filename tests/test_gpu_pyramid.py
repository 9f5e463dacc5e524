"""Pyramids on the GPU against the oracle: Euler and Navier-Stokes RHS on
meshes of pyramids alone (6 per cube, quad faces between cubes, triangles
inside), pyramids with hexes (quad interfaces) and pyramids with tets
(triangle interfaces), with boundary conditions and anti-aliasing; every
flux-kernel variant in FP32 (the 4x4x1 MFMA form at 5, 14, 30 and 55
solution points); RK4 trajectories.

Tolerances: max(1e-12, 10 x the oracle's input-noise floor) up to p = 3;
5e-11 at p = 4, where the oracle's operators (the reference's Jacobi form)
carry about 3e-11 of rounding (tests/test_pyramid.py shows the product's
are the more accurate); FP32 2e-5."""
import numpy as np
import pytest

from tests.fr_cases import eval_ics, make_simplex_cfg, mode_ics, rel_l2
from tests.test_gpu_simplex import MU, SIMPLEX_BCS, _cat, _noise_floor, \
    _rk4_dict

pytestmark = pytest.mark.gpu

L = 2*np.pi


def _tables(et, p, aa=(), fams=None):
    """Oracle tables; fams: {etype: soln-pts, 'tri': flux-pts}."""
    from pyfr_amd.quadtables import TABLES
    from pyfr_amd.shapes import gauss_legendre
    from pyfr_amd.simplex import AA_RULES, SF_TRI_RULES, SIMPLEX_SHAPES
    fams = fams or {}
    if et == 'hex':
        return ({'nq1': p + 2} if 'flux' in aa else {}) | (
            {'nf1': p + 2} if 'surf-flux' in aa else {}) | (
            {'upts': fams['hex']} if 'hex' in fams else {})
    sh = SIMPLEX_SHAPES[et](p, upts_name=fams.get(et))
    t = {'upts': TABLES[(et, sh.upts_name, sh.nupts)],
         'tri_fpts': TABLES[('tri', fams.get('tri', 'williams-shunn'),
                             (p + 1)*(p + 2)//2)][0],
         'tri_q36': TABLES[('tri', 'williams-shunn', 36)]}
    if 'flux' in aa:
        t['qrule'] = TABLES[(et, *AA_RULES[et][p])]
    if 'surf-flux' in aa:
        g, w = gauss_legendre(p + 2)
        r = range(p + 2)
        t['iqrules'] = {
            'tri': TABLES[('tri', *SF_TRI_RULES[p])],
            'quad': (np.array([(g[i], g[j]) for j in r for i in r]),
                     np.array([w[i]*w[j] for j in r for i in r]))}
    return t


def _case(be, layers, order, system='euler', bcs=False, rsolver='rusanov',
          scheme='rk4', dt=1e-3, aa=(), n=2, beta=0.0, tau=0.0, sorder=1,
          warp=None, mesh=None, fams=None):
    """The integrator and the oracle on a box of the layers (or on mesh,
    whose elements the oracle takes as they are)."""
    from oracle.simplex import OracleMixedSystem
    from pyfr_amd.mesh import SimplexBoxMesh
    from pyfr_amd.solvers import get_integrator
    ics = mode_ics(seed=7)
    cfg = make_simplex_cfg(order, be.precision, scheme, dt, ics=ics,
                           rsolver=rsolver, system=system, beta=beta,
                           tau=tau)
    cfg.set('backend-hip', 'fused', str(be.fused))
    if aa:
        cfg.set('solver', 'anti-alias', ', '.join(aa))
    for et, fam in (fams or {}).items():
        if et == 'tri':
            cfg.set('solver-interfaces-tri', 'flux-pts', fam)
        else:
            cfg.set(f'solver-elements-{et}', 'soln-pts', fam)
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
                if k == 'cpTw':
                    c[k] = float(eval(v))
                else:
                    vals[k] = eval_ics({k: v})[k]
            obcs[side] = (bctype, vals)
    nz = sum(k for _, k in layers)
    if mesh is None:
        mesh = SimplexBoxMesh([((n, n, nz), L/n, layers)], periodic,
                              sorder=sorder, warp=warp)
    integ = get_integrator(be, mesh, cfg)
    verts = {et: v.transpose(0, 2, 1) for et, v in mesh.spts.items()}
    idx = {et: np.arange(v.shape[1]) for et, v in mesh.spts.items()}
    orc = OracleMixedSystem(verts, order, {et: _tables(et, order, aa, fams)
                                           for et in verts},
                            [((L, L, nz*L/n), idx)], rsolver=rsolver,
                            system=system, constants=c, periodic=periodic,
                            bcs=obcs, antialias=tuple(a for a in aa
                                                      if a == 'flux'))
    return integ, orc, orc.ics(eval_ics(ics))


def _tol(orc, u0, order):
    return max(1e-12 if order <= 3 else 5e-11, 10*_noise_floor(orc, u0))


def _check(integ, orc, u0, order, tag, tol=None):
    sysm = integ.system
    names = sysm.ele_types
    assert 'pyr' in names
    got0 = dict(zip(names, sysm.ele_scal_upts(0)))
    assert rel_l2(_cat(got0, names), _cat(u0, names)) < 1e-7
    sysm.rhs(0.0, 0, 1)
    got = dict(zip(names, sysm.ele_scal_upts(1)))
    ref = orc.rhs(u0)
    tol = tol or _tol(orc, u0, order)
    err = rel_l2(_cat(got, names), _cat(ref, names))
    print(f'[{tag}] rel L2 {err:.3e} (tol {tol:.3e})', flush=True)
    assert err < tol


LAYOUTS = {'pyr': [('pyr', 2)], 'hex-pyr': [('hex', 1), ('pyr', 1)],
           'tet-pyt': [('tet', 1), ('pyt', 1)],
           'pyt-pri': [('pyt', 1), ('pri', 1)]}


@pytest.mark.parametrize('fused', ['flux', 'div', 'none'])
@pytest.mark.parametrize('layout,order', [
    ('pyr', 1), ('pyr', 2), ('pyr', 3), ('pyr', 4), ('hex-pyr', 2),
    ('hex-pyr', 3), ('tet-pyt', 2), ('tet-pyt', 4), ('pyt-pri', 3)])
def test_pyramid_euler_rhs_parity(backend_factory, layout, order, fused):
    be = backend_factory(fused=fused != 'none', fuse_flux=fused == 'flux')
    tag = f'pyr euler {layout} p={order} fused={fused}'
    print(f'\n[{tag}]', flush=True)
    integ, orc, u0 = _case(be, LAYOUTS[layout], order)
    _check(integ, orc, u0, order, tag)


@pytest.mark.parametrize('layout,order,beta,tau,bcs', [
    ('pyr', 2, 0.0, 0.0, False), ('pyr', 3, 0.25, 0.1, False),
    ('hex-pyr', 2, 0.0, 0.0, False), ('tet-pyt', 3, 0.5, 0.0, False),
    ('pyr', 2, 0.0, 0.0, True), ('tet-pyt', 2, 0.0, 0.0, True)])
def test_pyramid_ns_rhs_parity(backend_factory, layout, order, beta, tau,
                               bcs):
    be = backend_factory()
    tag = f'pyr ns {layout} p={order} beta={beta} tau={tau} bcs={bcs}'
    print(f'\n[{tag}]', flush=True)
    integ, orc, u0 = _case(be, LAYOUTS[layout], order, 'navier-stokes',
                           bcs=bcs, beta=beta, tau=tau)
    _check(integ, orc, u0, order, tag)


@pytest.mark.parametrize('system,layout,order,aa', [
    ('euler', 'pyr', 2, ('flux',)), ('navier-stokes', 'pyr', 2, ('flux',)),
    ('euler', 'hex-pyr', 2, ('surf-flux',)),
    ('navier-stokes', 'tet-pyt', 2, ('flux', 'surf-flux'))])
def test_pyramid_antialias_rhs_parity(backend_factory, system, layout, order,
                                      aa):
    be = backend_factory()
    tag = f'pyr aa {system} {layout} p={order} {aa}'
    print(f'\n[{tag}]', flush=True)
    integ, orc, u0 = _case(be, LAYOUTS[layout], order, system, aa=aa)
    _check(integ, orc, u0, order, tag)


def test_pyramid_euler_bcs(backend_factory):
    be = backend_factory()
    print('\n[pyr euler bcs] hex-pyr p=3', flush=True)
    integ, orc, u0 = _case(be, LAYOUTS['hex-pyr'], 3, bcs=True, rsolver='hll')
    _check(integ, orc, u0, 3, 'pyr euler bcs')


@pytest.mark.parametrize('variant', [1, 4, 32, 64, 65])
def test_pyramid_fp32_flux_variants(backend_factory, variant):
    """Every FP32 variant of the flux-fused divergence kernel on pyramids
    (5, 14, 30, 55 solution points: 2, 4, 8 and 14 4-row groups of the
    4x4x1 MFMA form) against the oracle."""
    from pyfr_amd.backend import MulProvider
    be = backend_factory(precision='single', graphs=False, fused=True,
                         fuse_flux=True)
    mp = next(p for p in be._providers if isinstance(p, MulProvider))
    for order in (1, 2, 3, 4):
        print(f'\n[pyr fp32 variant] v={variant} p={order}', flush=True)
        integ, orc, u0 = _case(be, LAYOUTS['pyr'], order)
        sysm = integ.system
        names = sysm.ele_types
        sysm.rhs(0.0, 0, 1)
        n0 = len(mp._plans3)
        for plan, _, _ in mp._plans3.values():
            be.lib.tdivf_plan_set_variant(plan, variant)
        assert n0
        for b in sysm.ele_banks:
            b[1].set(np.full(b[1].ioshape, np.nan))
        sysm.rhs(0.0, 0, 1)
        got = dict(zip(names, sysm.ele_scal_upts(1)))
        err = rel_l2(_cat(got, names), _cat(orc.rhs(u0), names))
        print(f'[pyr fp32 variant] rel L2 {err:.3e}', flush=True)
        assert err < 2e-5
        del integ, sysm


@pytest.mark.parametrize('layout,precision', [('hex-pyr', 'double'),
                                              ('tet-pyt', 'double'),
                                              ('pyr', 'single')])
def test_pyramid_rk4_trajectory(backend_factory, layout, precision):
    be = backend_factory(precision=precision)
    dt = 2e-3
    print(f'\n[pyr rk4] {layout} {precision}', flush=True)
    integ, orc, u0 = _case(be, LAYOUTS[layout], 3, dt=dt)
    names = integ.system.ele_types
    integ.run_steps(10, dt)
    u = u0
    for _ in range(10):
        u = _rk4_dict(orc.rhs, u, dt)
    err = rel_l2(_cat(dict(zip(names, integ.soln)), names), _cat(u, names))
    print(f'[pyr rk4] 10 steps rel L2 {err:.3e}', flush=True)
    assert err < (1e-12 if precision == 'double' else 2e-5)


@pytest.mark.parametrize('system,layout,order,sorder', [
    ('euler', 'pyr', 2, 2), ('euler', 'tet-pyt', 3, 3),
    ('navier-stokes', 'pyr', 3, 2), ('navier-stokes', 'hex-pyr', 2, 2)])
def test_pyramid_curved_rhs_parity(backend_factory, system, layout, order,
                                   sorder):
    """Pyramids given by order-k shape points take the curved path (the
    stored invariant-form metric at every point).  The box is not warped:
    warped pyramids do not meet their neighbours exactly (their shape on a
    triangular face depends on shape points off it), so the oracle could
    not pair their faces by coordinates; tests/test_pyramid.py checks the
    warped metric itself against the oracle."""
    be = backend_factory()
    tag = f'pyr curved {system} {layout} p={order} k={sorder}'
    print(f'\n[{tag}]', flush=True)
    integ, orc, u0 = _case(be, LAYOUTS[layout], order, system,
                           sorder=sorder)
    e = integ.system.ele_map['pyr']
    assert e.curved_simplex and e.mesh_regions == {'curved': e.neles}
    _check(integ, orc, u0, order, tag)


def test_pyramid_gmsh_pyfrm_rhs_parity(backend_factory, tmp_path):
    """A pyramid-tet mesh written as Gmsh, imported, written as .pyfrm and
    read back runs the RHS of the oracle on the file's elements."""
    import io
    from pyfr_amd.gmsh import GmshMesh
    from pyfr_amd.mesh import SimplexBoxMesh
    from pyfr_amd.native import NativeMesh
    from tests.msh_files import mesh_to_msh
    from tests.test_gmsh import _sim_sides
    blk = [((2, 2, 2), L/2, LAYOUTS['tet-pyt'])]
    gen = SimplexBoxMesh(blk)
    fn = str(tmp_path/'pyt.pyfrm')
    GmshMesh(io.StringIO(mesh_to_msh(gen, _sim_sides(gen, (L, L, L))))) \
        .write_pyfrm(fn)
    mesh = NativeMesh(fn)
    assert sorted(mesh.etypes) == ['pyr', 'tet']
    be = backend_factory()
    print('\n[pyr pyfrm]', flush=True)
    integ, orc, u0 = _case(be, LAYOUTS['tet-pyt'], 3, mesh=mesh)
    _check(integ, orc, u0, 3, 'pyr pyfrm')


@pytest.mark.parametrize('system,layout,order,fams', [
    ('euler', 'tet-pyt', 3, {'tet': 'alpha-opt',
                             'pyr': 'gauss-legendre-lobatto',
                             'tri': 'alpha-opt'}),
    ('navier-stokes', 'pyt-pri', 2,
     {'pri': 'williams-shunn~gauss-legendre-lobatto',
      'pyr': 'gauss-legendre-lobatto'}),
    ('euler', 'pyt-pri', 4, {'pri': 'alpha-opt~gauss-legendre-lobatto',
                             'tet': 'alpha-opt', 'tri': 'alpha-opt'}),
    ('euler', 'hex-pyr', 3, {'hex': 'gauss-legendre-lobatto'}),
    ('navier-stokes', 'hex-pyr', 2, {'hex': 'gauss-legendre-lobatto',
                                     'pyr': 'gauss-legendre-lobatto'})])
def test_point_families_rhs_parity(backend_factory, system, layout, order,
                                   fams):
    """[solver-elements-<type>] soln-pts and [solver-interfaces-tri]
    flux-pts other than the defaults, on tet/pyramid/prism meshes, and
    Gauss-Legendre-Lobatto hexes (dense operator plans: the generated hex
    kernels are for Gauss-Legendre points)."""
    be = backend_factory()
    tag = f'families {system} {layout} p={order} {fams}'
    print(f'\n[{tag}]', flush=True)
    integ, orc, u0 = _case(be, LAYOUTS[layout], order, system, fams=fams)
    for et, fam in fams.items():
        if et in integ.system.ele_map:
            assert integ.system.ele_map[et].basis.upts_name == fam
    _check(integ, orc, u0, order, tag)
