"""Benchmark: GDOF.RHS/s of the FR right-hand side on one MI355X.

Workload (the headline configuration that fits one GPU): compressible
Navier-Stokes on the periodic cube [0, 2pi]^3 meshed as 64^3 linear hexes,
FR order p=3 (64 Gauss-Legendre solution points per element, 83.9M DOF),
FP64, Rusanov + LDG (beta=0, tau=0) interfaces, fixed-dt RK45 (dt=2e-4),
Taylor-Green initial field at Mach 0.1, Re 1600.  A step is one RK45 step
(5 RHS evaluations + 5 stage updates) with all data resident in HBM.

  python bench.py [--gpus 1] [--steps K] [--warmup W]
                  [--config cfg1|cfg0|cfg3|hex64rk4|pyr]

Prints one JSON line: `value` is the FP64 plain GDOF*RHS/s; a roofline
object for the dominant kernel (algorithmic bytes per launch / mean launch
time from HIP events on the solver's stream); `precision_form` (FP32 RHS,
FP64 RK state) with its speed-up; `serving` (continuous batching of a seeded
Poisson stream of independent solves: p50/p99 latency, aggregate
GDOF*RHS/s and the ratio to one-solve-at-a-time); and a CPU baseline (the
oracle on a bounded sample).
"""
import argparse
import ctypes as ct
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0

CONFIGS = {
    'cfg1': dict(system='navier-stokes', n=64, order=3, scheme='rk45',
                 dt=2e-4, workload='ns-hex64-p3-rk45-fp64'),
    'cfg0': dict(system='euler', n=8, order=2, scheme='rk4', dt=1e-3,
                 workload='euler-hex8-p2-rk4-fp64'),
    # cfg0's scheme at cfg1's size: Euler RK4 with the fused stage updates
    'hex64rk4': dict(system='euler', n=64, order=2, scheme='rk4', dt=1e-4,
                     workload='euler-hex64-p2-rk4-fp64'),
    # dense-operator mixed mesh: a periodic box of 48^3 cells x 6 tets and a
    # periodic 16^3-cell x 2-prism slab, p=4, FP32, 8 random-phase modes
    'cfg3': dict(system='euler', ntet=48, npri=16, order=4, scheme='rk4',
                 dt=1e-3, precision='single',
                 workload='euler-tet48^3x6+pri16^3x2-p4-rk4-fp32'),
    # pyramids on the same dense path: a periodic box of 32^3 cells x 6
    # pyramids (apex at the cell centre), p=4 (55 solution points), FP32
    'pyr': dict(system='euler', npyr=32, order=4, scheme='rk4', dt=1e-3,
                precision='single', workload='euler-pyr32^3x6-p4-rk4-fp32'),
}

MFMA_PEAK_TFS = {'f32': 157.3, 'f64': 78.6}


FUSED = True
# [backend-hip] row-pad of the measured runs (elements; None: the
# backend's default)
ROW_PAD = None


def make_cfg(c, precision='double'):
    from pyfr_amd.cases import make_cfg as mk, make_simplex_cfg, mode_ics
    if 'ntet' in c or 'npyr' in c:
        cfg = make_simplex_cfg(c['order'], precision, c['scheme'], c['dt'],
                               mode_ics(seed=0, nmodes=8, amp=0.05))
    else:
        cfg = mk(c['system'], c['order'], 'rusanov', precision=precision,
                 scheme=c['scheme'], dt=c['dt'])
    cfg.set('backend-hip', 'fused', str(FUSED))
    if ROW_PAD is not None:
        cfg.set('backend-hip', 'row-pad', str(ROW_PAD))
    return cfg


def make_mesh(c):
    from pyfr_amd.mesh import BoxMesh, SimplexBoxMesh
    if 'ntet' in c:
        return SimplexBoxMesh.disjoint(c['ntet'], c['npri'])
    if 'npyr' in c:
        n = c['npyr']
        return SimplexBoxMesh([((n, n, n), 2*np.pi/n, [('pyr', n)])])
    return BoxMesh(c['n'])


# -- algorithmic bytes per launch (every input read once, every output
#    written once, read-modify-write counted twice)
def kernel_bytes(k, isz, nvars=5):
    name = k.name
    if name == 'mul':
        a, b, c = k.mats
        beta = k.rtargs.get('beta', 0.0)
        return (b.nrow + c.nrow*(1 + (beta != 0)))*b.ncol*isz + a.nrow*a.ncol*0
    if name == 'tflux':
        u, f, geo = k.mats
        return (u.nrow*u.ncol + f.nrow*f.ncol + geo.nrow*geo.ncol)*isz
    if name == 'tflux_ns':
        u, f, g, geo = k.mats
        return (u.nrow*u.ncol + f.nrow*f.ncol + 2*g.nrow*g.ncol +
                geo.nrow*geo.ncol)*isz
    if name == 'negdivconf':
        t, r = k.mats
        return (2*t.nrow*t.ncol + r.nrow*r.ncol)*isz
    if name == 'intcflux':
        ul = k.views[0]
        return ul.n*(4*nvars*isz + 3*isz + 2*4)
    if name == 'intcflux_ns':
        ul = k.views[0]
        return ul.n*(4*nvars*isz + 2*3*nvars*isz + 3*isz + 4*4)
    if name == 'intconu':
        ul = k.views[0]
        return ul.n*(4*nvars*isz + 4*4)
    if name == 'rkvdh2':
        r1, r2 = k.mats
        return 4*r1.nrow*r1.ncol*isz
    if name == 'rkvdh2_mixed':
        acc, w = k.mats
        return acc.nrow*acc.ncol*(2*8 + 2*4)
    if name.startswith('mul2_ep'):
        ep = int(name[-1])
        a0, a1, b0, b1, out = k.mats[:5]
        n = out.ncol
        by = (b0.nrow + b1.nrow + out.nrow)*n*isz
        if ep >= 1:
            rcp = k.mats[5]                          # 1/|J| per point or
            by += rcp.nrow*(n // nvars)*isz          # per element
        if ep == 2:
            by += 2*out.nrow*n*isz                   # r1 read + write
        if ep == 3:
            by += 2*out.nrow*n*8                     # FP64 solution r/w
        return by
    if name.startswith('tdivf_flux_ep'):
        ep = int(name[-1])
        a0, a1, u, sm, fn, out, rcp = k.mats
        n = out.ncol
        by = (u.nrow + fn.nrow + out.nrow)*n*isz
        by += (sm.nrow + rcp.nrow)*(n // nvars)*isz   # S, 1/|J| per element
        if ep == 2:
            by += 2*out.nrow*n*isz
        if ep == 3:
            by += 2*out.nrow*n*8
        return by
    if name == 'axnpby':
        x = k.mats
        return (len(x) + 1)*x[0].nrow*x[0].ncol*isz
    return 0


def rhs_kernel_list(be, integ, c, precision):
    """The kernels of one RHS evaluation plus its stage update, in launch
    order, as the timed steps run them (after at least one step)."""
    sysm, dt = integ.system, c['dt']
    mixed = precision == 'mixed'
    if c['scheme'] == 'rk45' and sysm.fused_stages:
        # the graph of a middle RK45 stage (RHS fused with its update)
        key = [k for k in sysm._kcache if len(k) == 5 and k[3] == 1][0]
        return flatten(sysm._kcache[key][0].klist)
    g = sysm._rhs_graphs(*((1, 1) if mixed else (0, 1)))[0]
    kerns = flatten(g.klist)
    if c['scheme'] == 'rk45' and not mixed:
        rk = integ._rkvdh2_kerns(0, 0, 1)
        for k in rk:
            k.bind(dt=dt)
        kerns += rk
    elif mixed:
        kerns += [mixed_stage_kernel(be, integ, dt)]
    return kerns


def measured_traffic(c, precision, label):
    """HBM bytes per launch of a kernel from the committed rocprofv3
    FETCH_SIZE/WRITE_SIZE summary of the same workload (None if it holds no
    figure for it)."""
    fn = os.path.join(ROOT, 'profiles', 'pmc-traffic_rhs-kernels.json')
    try:
        data = json.load(open(fn))['workloads']
    except (OSError, KeyError, ValueError):
        return None
    wl = c['workload'].replace('-fp64', '-' + {
        'double': 'fp64', 'single': 'fp32', 'mixed': 'mixed'}[precision])
    ent = data.get(wl, {}).get(label)
    return ent['traffic_bytes'] if ent else None


def kernel_label(k):
    """Timing label: the kernel name plus the operator shape for operator
    applications (different element types stay apart)."""
    if k.name == 'mul':
        return f'mul{k.mats[0].nrow}x{k.mats[0].ncol}'
    if k.name.startswith('mul2_ep') or k.name.startswith('tdivf_flux'):
        a0, a1 = k.mats[:2]
        return f'{k.name}_{a0.nrow}x{a0.ncol}+{a1.ncol}'
    return k.name


def kernel_flops(k):
    """Algorithmic flops of dense operator applications (2 m k per
    column); 0 for everything else."""
    if getattr(k, 'kind', None) != 'dense':
        return 0
    if k.name == 'mul':
        a, b = k.mats[:2]
        return 2*a.nrow*a.ncol*b.ncol
    a0, a1, b0 = k.mats[:3]
    return 2*a0.nrow*(a0.ncol + a1.ncol)*b0.ncol   # mul2 / tdivf_flux


def mesh_label(c):
    if 'ntet' in c:
        return f'{c["ntet"]}^3x6 tet + {c["npri"]}^3x2 pri (periodic boxes)'
    if 'npyr' in c:
        return f'{c["npyr"]}^3x6 pyr (periodic box)'
    return f'{c["n"]}^3 hex'


def mixed_stage_kernel(be, integ, dt):
    """The FP64/FP32 stage update of the precision form as a timed kernel."""
    from pyfr_amd.backend import Kernel
    acc, w = integ.system.ele_banks[0][0], integ.system.ele_banks[0][1]
    a, b = dt*integ.a[0], dt*integ.b[0]
    return Kernel(lambda stream: be.lib.rkvdh2_mixed(
        acc.nrow, acc.ncol, acc.leaddim, acc.data, w.data, a, b, 0, stream),
        mats=(acc, w), name='rkvdh2_mixed')


def flatten(kerns):
    out = []
    for k in kerns:
        if hasattr(k, 'kernels'):
            out += flatten(k.kernels)
        else:
            out.append(k)
    return out


def time_kernels(be, kerns, reps):
    """Mean duration (ms) of each kernel over reps launches, with HIP events
    recorded on the stream the kernels run on."""
    lib = be.lib
    e0, e1 = ct.c_void_p(), ct.c_void_p()
    lib.event_create(ct.byref(e0))
    lib.event_create(ct.byref(e1))
    out = []
    for k in kerns:
        k.run(be.stream)
        lib.event_record(e0, be.stream)
        for _ in range(reps):
            k.run(be.stream)
        lib.event_record(e1, be.stream)
        lib.event_sync(e1)
        ms = ct.c_float()
        lib.event_elapsed(ct.byref(ms), e0, e1)
        out.append(ms.value/reps)
    lib.event_destroy(e0)
    lib.event_destroy(e1)
    return out


def cpu_baseline(c, budget_s=12.0):
    """The oracle's C/OpenMP restatement of the reference OpenMP backend's
    RHS (oracle/crhs.c, blocked 8-wide SoA layout, sparse operators),
    built with that backend's flags (-O3 -ffast-math, the host's widest ISA:
    pyfr/backends/openmp/compiler.py:76-96), on all host cores this process
    may use, on a bounded sample of the same workload: same system, order
    and field on a 24^3 box (12^3 for cfg0's order).  The full 64^3 box is
    not used because the NumPy oracle that builds the operators, metric
    terms and interface lists takes minutes and ~18 GB there; the per-DOF
    rate is size-independent once the working set (~1.5 GB at 24^3) is far
    beyond the host caches."""
    ncores = len(os.sched_getaffinity(0))
    os.environ['OMP_NUM_THREADS'] = str(ncores)
    os.environ.setdefault('OMP_PROC_BIND', 'close')
    from oracle.crhs import CRhs, cpu_model, host_isa
    from oracle.rhs import OracleHexSystem
    from pyfr_amd.cases import NS_CONSTS
    n = 24 if c['order'] >= 3 else 12
    orc = OracleHexSystem(n, c['order'], c['system'],
                          constants=({'gamma': 1.4} if c['system'] == 'euler'
                                     else {**NS_CONSTS, 'ldg-beta': 0.0,
                                           'ldg-tau': 0.0}))
    crhs = CRhs(orc, fast=True)
    crhs.set_u(orc.ics())
    crhs.rhs_inplace()
    t0, nrhs = time.perf_counter(), 0
    while time.perf_counter() - t0 < budget_s:
        crhs.rhs_inplace()
        nrhs += 1
    dt = time.perf_counter() - t0
    isa = {'v3': 'x86-64-v3 (AVX2)', 'v4': 'x86-64-v4 (AVX-512)'}[host_isa()]
    return {'value': orc.ndofs*nrhs/dt/1e9, 'unit': 'GDOF*RHS/s',
            'cores': ncores, 'kind': 'port',
            'sample': f'oracle C/OpenMP FP64 {c["system"]} RHS (restatement '
                      f'of the reference OpenMP kernels), {n}^3 hex '
                      f'p={c["order"]} ({orc.ndofs} DOF), {nrhs} RHS in '
                      f'{dt:.1f} s on {ncores} threads (OMP_PROC_BIND='
                      f'{os.environ["OMP_PROC_BIND"]}); gcc -O3 -ffast-math '
                      f'-march={isa} -fopenmp; CPU {cpu_model()}'}


def serving_report(nreq, seed=0, rate=50.0):
    """Serving form: a seeded Poisson stream (mean 50 req/s) of independent
    periodic-box solves (N ~ U{4..24}, p ~ U{1..4}, Mach ~ U[0.05, 0.3],
    Re ~ U[100, 1600] or inviscid w.p. 0.3; 200 RK4 steps each), served with
    continuous batching and, as the baseline, one solve at a time."""
    from pyfr_amd.backend import AMDBackend
    from pyfr_amd.inifile import Inifile
    from pyfr_amd.serving import (ServingEngine, poisson_requests,
                                  sequential_bound, summarize)
    cfg = Inifile.from_dict({'backend': {'precision': 'double'}})
    be = AMDBackend(cfg)
    t0 = time.perf_counter()
    eng = ServingEngine(be, fused=FUSED)
    reqs = poisson_requests(seed, rate, nreq)
    for r in reqs:
        eng.box(r.N, r.order)
    tsetup = time.perf_counter() - t0
    print(f'# serving: {len(reqs)} requests, pools {eng.hbm_bytes/2**30:.1f}'
          f' GiB, host set-up {tsetup:.1f} s', file=sys.stderr, flush=True)
    # warm-up: capture/compile paths for every class once
    warm = poisson_requests(seed + 1, rate, 8, nsteps=2)
    eng.run([type(r)(r.rid, 4, p, 0.1, re, 2, 0.0) for p in (1, 2, 3, 4)
             for re in (None, 500.0) for r in warm[:1]])
    out = {}
    for mode in ('batched', 'sequential'):
        for r in reqs:
            r.admit_time = r.finish_time = r.result = None
        for p in eng.pools.values():
            p.ncaptures = p.nsteps_run = 0
            p.host_s = 0.0
        done = eng.run(reqs, mode=mode)
        out[mode] = summarize(done)
        steps = sum(p.nsteps_run for p in eng.pools.values())
        out[mode]['graph_captures'] = sum(p.ncaptures
                                          for p in eng.pools.values())
        out[mode]['host_ms_per_pool_step'] = (
            1e3*sum(p.host_s for p in eng.pools.values())/max(steps, 1))
        if mode == 'sequential':
            out['bound'] = sequential_bound(done)
        print(f'# {mode}: {out[mode]}', file=sys.stderr, flush=True)
    print(f'# bound: {out["bound"]}', file=sys.stderr, flush=True)
    out['throughput_ratio'] = (out['batched']['gdof_rhs_per_s'] /
                               out['sequential']['gdof_rhs_per_s'])
    out['workload'] = (f'poisson-{rate:g}req/s-{nreq}solves-hex-N4..24-'
                       'p1..4-rk4x200-fp64')
    out['kernels'] = 'fused' if FUSED else 'unfused'
    return out


def accuracy_report(c, checkpoints=(100, 500)):
    """Per-field relative L2 and max errors of the precision form against
    the FP64 run from the same initial field at the checkpoints, plus the
    FP64 conservation sums: RK45 (the workload's scheme) in the base-state
    form, the full-state form and with BF16 operator inputs, and RK4 in the
    base-state form.  Momenta are also normalised by the momentum vector's
    L2 norm (the w-momentum of the Taylor-Green field starts at zero)."""
    from pyfr_amd.backend import AMDBackend
    from pyfr_amd.mesh import BoxMesh
    from pyfr_amd.solvers import get_integrator
    runs = {'rk45': ('double', 'mixed', 'mixed-fullstate', 'mixed-bf16'),
            'rk4': ('double', 'mixed')}
    names = ['rho', 'rhou', 'rhov', 'rhow', 'E']
    out = {'workload': c['workload'].replace('-fp64', '-mixed-vs-fp64'),
           'checkpoints': {}}
    for scheme, forms in runs.items():
        res = {}
        for form in forms:
            cfg = make_cfg(dict(c, scheme=scheme), form.split('-')[0])
            if form == 'mixed-bf16':
                cfg.set('backend-hip', 'operator-inputs', 'bf16')
            if form == 'mixed-fullstate':
                cfg.set('backend-hip', 'base-state', 'off')
            be = AMDBackend(cfg)
            integ = get_integrator(be, BoxMesh(c['n']), cfg)
            res[form] = {'sums0': integ.system.conservation_sums(0).tolist()}
            done = 0
            t0 = time.perf_counter()
            for cp in checkpoints:
                integ.run_steps(cp - done, c['dt'])
                done = cp
                res[form][cp] = integ.soln[0]
                res[form][f'sums{cp}'] = integ.system.conservation_sums(
                    integ.idxcurr).tolist()
                print(f'# {scheme} {form}: step {cp} at '
                      f'{time.perf_counter() - t0:.1f} s', file=sys.stderr,
                      flush=True)
            del integ, be
        for form in forms[1:]:
            for cp in checkpoints:
                a, b = res[form][cp], res['double'][cp]
                mom = float(np.linalg.norm(b[:, 1:4]))
                d = {}
                for v, nm in enumerate(names):
                    da, db = a[:, v], b[:, v]
                    d[nm] = {'rel_l2': float(np.linalg.norm(da - db) /
                                             np.linalg.norm(db)),
                             'max_abs': float(np.abs(da - db).max()),
                             'ref_max_abs': float(np.abs(db).max())}
                    if 1 <= v <= 3:
                        d[nm]['rel_l2_mom'] = float(
                            np.linalg.norm(da - db)/mom)
                d['state_rel_l2'] = float(np.linalg.norm(a - b) /
                                          np.linalg.norm(b))
                d['sums_fp64'] = res['double'][f'sums{cp}']
                d['sums_form'] = res[form][f'sums{cp}']
                out['checkpoints'][f'{scheme}:{form}@{cp}'] = d
        out[f'sums_initial_{scheme}'] = res['double']['sums0']
    print(json.dumps(out))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=20)
    ap.add_argument('--warmup', type=int, default=3)
    ap.add_argument('--config', default='cfg1', choices=sorted(CONFIGS))
    ap.add_argument('--precision', default='double',
                    choices=['double', 'mixed', 'single'])
    ap.add_argument('--no-cpu-baseline', action='store_true')
    ap.add_argument('--unfused', action='store_true',
                    help="run the reference's unfused kernel decomposition")
    ap.add_argument('--no-mixed', action='store_true',
                    help='skip the precision-form measurement')
    ap.add_argument('--no-serving', action='store_true',
                    help='skip the serving-form measurement')
    ap.add_argument('--no-dense', action='store_true',
                    help='skip the dense tet/prism mixed-mesh measurement')
    ap.add_argument('--kernel-reps', type=int, default=5)
    ap.add_argument('--row-pad', type=int, default=None,
                    help='[backend-hip] row-pad of the arrays (elements)')
    ap.add_argument('--serving', action='store_true',
                    help='serving form: continuous batching of a Poisson '
                    'stream of independent solves vs one at a time')
    ap.add_argument('--serving-requests', type=int, default=200)
    ap.add_argument('--counter-probe', action='store_true',
                    help='for rocprofv3 --pmc: one launch of each kernel of '
                    'one RHS (+ stage), then a 1 GiB device copy as the '
                    'traffic calibration')
    ap.add_argument('--accuracy', action='store_true',
                    help='precision form: compare mixed against FP64 at '
                    'steps 100 and 500 instead of timing')
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get('WORLD_SIZE', '1')) > 1:
        print('bench.py measures one GPU only: this build has no multi-GPU '
              'path (the domain-decomposed design is described in '
              'DESIGN.md).', file=sys.stderr)
        sys.exit(2)

    global FUSED, ROW_PAD
    FUSED = not args.unfused
    ROW_PAD = args.row_pad
    c = CONFIGS[args.config]
    if args.accuracy:
        return accuracy_report(c)
    if args.counter_probe:
        return counter_probe(c, c.get('precision', args.precision))
    if args.serving:
        print(json.dumps(serving_report(args.serving_requests)))
        return

    prec = c.get('precision', args.precision)
    res = measure(c, prec, args.steps, args.warmup, args.kernel_reps)
    line = {
        'metric': 'GDOF*RHS/s', 'value': round(res['gdofs'], 4),
        'unit': 'GDOF*RHS/s', 'n_gpus': 1, 'steps': args.steps,
        'warmup': args.warmup, 'ms_per_step': round(res['ms_per_step'], 3),
        'higher_is_better': True, 'scaling': 'weak', 'vs_baseline': None,
        'dtype': res['dtype'],
        'data': ('synthetic (8 seeded random-phase modes)'
                 if 'ntet' in c or 'npyr' in c
                 else 'synthetic (analytic Taylor-Green field)'),
        'config': res['config'], 'roofline': res['roofline'],
        'rhs_kernels': res['rhs_kernels'],
        'rhs_alg_bytes_per_dof': res['rhs_alg_bytes_per_dof'],
        'rhs_kernel_ms_sum': res['rhs_kernel_ms_sum'],
    }
    if prec == 'double' and not args.no_mixed:
        mx = measure(c, 'mixed', args.steps, args.warmup, args.kernel_reps)
        line['precision_form'] = {
            'value': round(mx['gdofs'], 4), 'unit': 'GDOF*RHS/s',
            'dtype': mx['dtype'], 'ms_per_step': round(mx['ms_per_step'], 3),
            'speedup_vs_fp64': round(mx['gdofs']/res['gdofs'], 3),
            'roofline': mx['roofline'],
            'rhs_alg_bytes_per_dof': mx['rhs_alg_bytes_per_dof'],
            'accuracy': 'profiles/accuracy_ns-hex64-p3-rk45_mixed-vs-fp64.md'}
    if not args.no_serving and args.config == 'cfg1':
        sv = serving_report(args.serving_requests)
        line['serving'] = {
            'workload': sv['workload'],
            'p50_latency_s': round(sv['batched']['p50_s'], 3),
            'p99_latency_s': round(sv['batched']['p99_s'], 3),
            'value': round(sv['batched']['gdof_rhs_per_s'], 4),
            'unit': 'GDOF*RHS/s',
            'one_at_a_time': {
                'value': round(sv['sequential']['gdof_rhs_per_s'], 4),
                'p50_latency_s': round(sv['sequential']['p50_s'], 3),
                'p99_latency_s': round(sv['sequential']['p99_s'], 3)},
            'throughput_ratio': round(sv['throughput_ratio'], 3),
            'kernels': sv['kernels'],
            'graph_captures': sv['batched']['graph_captures'],
            'host_ms_per_pool_step': round(
                sv['batched']['host_ms_per_pool_step'], 3),
            'bound_from_one_at_a_time': {
                k: (round(v, 3) if isinstance(v, float) else
                    {kk: round(vv, 3) for kk, vv in v.items()})
                for k, v in sv['bound'].items()}}
    if not args.no_dense and args.config == 'cfg1':
        line['dense_mixed_mesh'] = dense_report(args)
    line['cpu_baseline'] = (None if args.no_cpu_baseline or 'ntet' in c
                            or 'npyr' in c
                            else cpu_baseline(c))
    print(json.dumps(line))


def counter_probe(c, precision):
    """Launch each kernel of one RHS (+ stage update) once, in order, then
    copy 1 GiB device to device: the per-dispatch FETCH_SIZE / WRITE_SIZE
    of the copy calibrate the counters against a known byte count."""
    from pyfr_amd.backend import AMDBackend
    from pyfr_amd.solvers import get_integrator
    cfg = make_cfg(c, precision)
    be = AMDBackend(cfg)
    integ = get_integrator(be, make_mesh(c), cfg)
    integ.run_steps(1, c['dt'])
    be.wait()
    kerns = rhs_kernel_list(be, integ, c, precision)
    tuned = [t for p in be._providers if hasattr(p, 'tuning_report')
             for t in p.tuning_report()]
    print('# tuned variants: ' + ', '.join(
        f'{t["kernel"]} {t["operator"]} -> {t["variant"]}' for t in tuned),
        file=sys.stderr, flush=True)
    for k in kerns:
        print(f'# dispatch {k.name}', file=sys.stderr, flush=True)
        k.run(be.stream)
    be.wait()
    n = 1 << 30
    src, dst = ct.c_void_p(), ct.c_void_p()
    be.lib.malloc(ct.byref(src), n)
    be.lib.malloc(ct.byref(dst), n)
    be.lib.memset(src.value, 1, n, be.stream)
    be.wait()
    print(f'# dispatch copy {n} bytes', file=sys.stderr, flush=True)
    be.lib.memcpy_dtod(dst.value, src.value, n, be.stream)
    be.wait()
    be.lib.free(src.value)
    be.lib.free(dst.value)
    print(json.dumps({'kernels': [k.name for k in kerns],
                      'labels': [kernel_label(k) for k in kerns],
                      'alg_bytes': [kernel_bytes(
                          k, np.dtype(be.fpdtype).itemsize) for k in kerns],
                      'copy_bytes': n}))


def dense_report(args):
    """The dense-operator mixed mesh (tets + prisms, p=4, FP32) as a
    secondary measurement of the same JSON line."""
    c = CONFIGS['cfg3']
    r = measure(c, c['precision'], max(5, args.steps//2), args.warmup,
                args.kernel_reps)
    return {'workload': r['config']['workload'], 'value': round(r['gdofs'], 4),
            'unit': 'GDOF*RHS/s', 'dtype': r['dtype'],
            'ms_per_step': round(r['ms_per_step'], 3),
            'config': r['config'], 'roofline': r['roofline'],
            'rhs_kernels': r['rhs_kernels'],
            'rhs_alg_bytes_per_dof': r['rhs_alg_bytes_per_dof'],
            'tuned_variants': r.get('tuned_variants')}


def measure(c, precision, steps, warmup, kernel_reps):
    """Time `steps` steps of the workload after `warmup` steps, then time
    each kernel of one RHS (+ stage update) with HIP events."""
    from pyfr_amd.backend import AMDBackend
    from pyfr_amd.mesh import BoxMesh
    from pyfr_amd.solvers import get_integrator

    cfg = make_cfg(c, precision)
    t0 = time.perf_counter()
    be = AMDBackend(cfg)
    integ = get_integrator(be, make_mesh(c), cfg)
    sysm = integ.system
    tsetup = time.perf_counter() - t0
    print(f'# {precision}: setup {tsetup:.1f} s, {sysm.gndofs} DOF',
          file=sys.stderr, flush=True)

    dt = c['dt']
    integ.run_steps(warmup, dt)
    be.wait()
    t0 = time.perf_counter()
    integ.run_steps(steps, dt)
    be.wait()
    t1 = time.perf_counter()
    nrhs = integ.nrhs_per_step*steps
    gdofs = sysm.gndofs*nrhs/(t1 - t0)/1e9
    ms_per_step = (t1 - t0)*1e3/steps
    print(f'# {precision}: {steps} steps in {t1 - t0:.3f} s: '
          f'{gdofs:.3f} GDOF*RHS/s', file=sys.stderr, flush=True)

    kerns = rhs_kernel_list(be, integ, c, precision)
    isz = np.dtype(be.fpdtype).itemsize
    times = time_kernels(be, kerns, kernel_reps)
    per = {}
    for k, ms in zip(kerns, times):
        nm = kernel_label(k)
        d = per.setdefault(nm, {'ms': 0.0, 'bytes': 0, 'flops': 0,
                                'launches': 0})
        d['ms'] += ms
        d['bytes'] += kernel_bytes(k, isz)
        d['flops'] += kernel_flops(k)
        d['launches'] += 1
    mpeak = MFMA_PEAK_TFS['f64' if isz == 8 else 'f32']
    for nm, d in per.items():
        d['GBs'] = d['bytes']/(d['ms']*1e-3)/1e9 if d['ms'] > 0 else 0.0
        d['frac'] = d['GBs']/HBM_PEAK_GBS
        msg = (f'#   {nm:22s} {d["ms"]:8.3f} ms/RHS  {d["GBs"]:8.1f} GB/s '
               f'({100*d["frac"]:5.1f}% of peak)')
        if d['flops']:
            d['TFs'] = d['flops']/(d['ms']*1e-3)/1e12
            d['mfma_frac'] = d['TFs']/mpeak
            msg += f'  {d["TFs"]:6.1f} TFLOP/s ({100*d["mfma_frac"]:4.1f}%)'
        print(msg, file=sys.stderr)
    dom = max(per, key=lambda n: per[n]['ms'])
    dd = per[dom]
    roof = {'bound': 'hbm', 'kernel': dom,
            'achieved': round(dd['GBs'], 1), 'peak': HBM_PEAK_GBS,
            'unit': 'GB/s', 'frac': round(dd['frac'], 4),
            'traffic': None, 'bytes_per_launch': dd['bytes']//dd['launches'],
            'avg_launch_ms': round(dd['ms']/dd['launches'], 4)}
    roof['traffic'] = measured_traffic(c, precision, dom)
    if dd.get('mfma_frac', 0) > dd['frac']:
        roof.update({'bound': 'mfma', 'achieved': round(dd['TFs'], 2),
                     'peak': mpeak, 'unit': 'TFLOP/s',
                     'frac': round(dd['mfma_frac'], 4),
                     'flops_per_launch': dd['flops']//dd['launches'],
                     'hbm_GBs': round(dd['GBs'], 1)})
    dtype = {'double': 'f64', 'single': 'f32',
             'mixed': 'f32 RHS / f64 state'}[precision]
    out = {
        'gdofs': gdofs, 'ms_per_step': ms_per_step, 'dtype': dtype,
        'config': {'workload': c['workload'].replace(
                       '-fp64', '-' + {'double': 'fp64', 'single': 'fp32',
                                       'mixed': 'mixed'}[precision]),
                   'kernels': 'fused' if FUSED else 'reference-decomposition',
                   'mesh': mesh_label(c), 'order': c['order'],
                   'dof': sysm.gndofs, 'scheme': c['scheme'], 'dt': dt,
                   'rhs_per_step': integ.nrhs_per_step},
        'roofline': roof,
        'rhs_kernels': {n: {k: (round(v, 4) if isinstance(v, float) else v)
                            for k, v in d.items()} for n, d in per.items()},
        'rhs_alg_bytes_per_dof': round(sum(d['bytes'] for d in per.values())
                                       / sysm.gndofs, 1),
        'rhs_kernel_ms_sum': round(sum(d['ms'] for d in per.values()), 3),
    }
    tuned = [t for p in be._providers if hasattr(p, 'tuning_report')
             for t in p.tuning_report()]
    if tuned:
        out['tuned_variants'] = tuned
        print('# tuned variants: ' + ', '.join(
            f'{t["kernel"]} {t["operator"]} -> {t["variant"]}'
            for t in tuned), file=sys.stderr)
    del integ, sysm, kerns
    be.wait()
    return out


if __name__ == '__main__':
    main()
