"""Shared configurations for the parity tests (the case builders live in
pyfr_amd.cases, where bench.py also reads them)."""
import numpy as np

from pyfr_amd.cases import (NS_CONSTS, TGV_ICS, make_cfg,  # noqa: F401
                            make_simplex_cfg, mode_ics)


def oracle_consts(system, beta=0.0, tau=0.0):
    if system == 'euler':
        return {'gamma': 1.4}
    c = dict(NS_CONSTS)
    c.update({'ldg-beta': beta, 'ldg-tau': tau})
    return c


def rel_l2(a, b):
    return float(np.linalg.norm((a - b).ravel())/np.linalg.norm(b.ravel()))


def eval_ics(ics):
    """Callables of (x, y, z) for expressions of ``mode_ics``/``TGV_ICS``."""
    env = {'sin': np.sin, 'cos': np.cos, '__builtins__': {}}

    def mk(expr):
        return lambda x, y, z: eval(expr, env, {'x': x, 'y': y, 'z': z}) \
            + 0*x
    return {k: mk(v) for k, v in ics.items()}
