"""Configurations of the benchmark workloads: the analytic Taylor-Green
field on the periodic box, random-phase mode fields, and the ini
files (the reference's section and key names) for hex and tet/prism
runs."""
import numpy as np

from pyfr_amd.inifile import Inifile

TGV_ICS = {
    'rho': '1',
    'u': 'sin(x)*cos(y)*cos(z)',
    'v': '-cos(x)*sin(y)*cos(z)',
    'w': '0',
    'p': '100/1.4 + (cos(2*x) + cos(2*y))*(cos(2*z) + 2)/16',
}

NS_CONSTS = {'gamma': 1.4, 'mu': 1/1600, 'Pr': 0.71}


def make_cfg(system='euler', order=2, rsolver='rusanov', precision='double',
             scheme='rk4', dt=1e-3, beta=0.0, tau=0.0, visc_corr='none'):
    consts = {'gamma': 1.4} if system == 'euler' else dict(NS_CONSTS)
    d = {
        'backend': {'precision': precision},
        'constants': consts,
        'solver': {'system': system, 'order': order,
                   'viscosity-correction': visc_corr},
        'solver-time-integrator': {'formulation': 'explicit',
                                   'scheme': scheme, 'controller': 'none',
                                   'tstart': 0.0, 'dt': dt},
        'solver-interfaces': {'riemann-solver': rsolver, 'ldg-beta': beta,
                              'ldg-tau': tau},
        'solver-interfaces-quad': {'flux-pts': 'gauss-legendre'},
        'solver-elements-hex': {'soln-pts': 'gauss-legendre'},
        'soln-ics': TGV_ICS,
    }
    return Inifile.from_dict(d)


def mode_ics(seed=0, nmodes=8, amp=0.05, kmax=2):
    """Superposition of ``nmodes`` random-phase sinusoidal modes (integer
    wavevectors in [-kmax, kmax]^3, periodic on [0, 2 pi]^3) of amplitude
    ``amp`` in density and each velocity component, on rho = 1, p = 1/1.4
    (sound speed 1).  Returns initial-condition expressions."""
    rng = np.random.default_rng(seed)
    out = {}
    for var, base in (('rho', '1.0'), ('u', '0.0'), ('v', '0.0'),
                      ('w', '0.0')):
        terms = [base]
        for _ in range(nmodes):
            k = np.zeros(3, dtype=int)
            while not k.any():
                k = rng.integers(-kmax, kmax + 1, 3)
            ph = rng.uniform(0, 2*np.pi)
            terms.append(f'{amp!r}*sin({k[0]}*x + {k[1]}*y + {k[2]}*z + '
                         f'{ph!r})')
        out[var] = ' + '.join(terms)
    out['p'] = repr(1/1.4)
    return out


def make_simplex_cfg(order=2, precision='double', scheme='rk4', dt=1e-3,
                     ics=None, rsolver='rusanov', system='euler', beta=0.0,
                     tau=0.0):
    cfg = make_cfg(system, order, rsolver, precision=precision,
                   scheme=scheme, dt=dt, beta=beta, tau=tau)
    for k, v in (ics or mode_ics()).items():
        cfg.set('soln-ics', k, v)
    return cfg
