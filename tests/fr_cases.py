"""Shared configurations for the parity tests."""
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


def oracle_consts(system, beta=0.0, tau=0.0):
    if system == 'euler':
        return {'gamma': 1.4}
    c = dict(NS_CONSTS)
    c.update({'ldg-beta': beta, 'ldg-tau': tau})
    return c


def rel_l2(a, b):
    return float(np.linalg.norm((a - b).ravel())/np.linalg.norm(b.ravel()))
