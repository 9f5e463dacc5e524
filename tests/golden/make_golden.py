"""Write the golden fixtures under tests/golden/ from data files held by the
reference (read as data, nothing imported or executed from it).

* hex_gleg_ord3.npz -- the arrays of ``pyfr/tests/hex-gleg-ord3.npz`` (M0..M3
  of a p=3 hex with Gauss-Legendre solution and flux points), the only golden
  data the reference's own tests hold for the hot path
  (``pyfr/tests/test_ele_mats.py:10-27``).
* gl_rules.npz -- Gauss-Legendre points and weights parsed from the
  reference's tabulated rule files ``pyfr/quadrules/{line,quad,hex}/
  gauss-legendre-n*.txt`` for 1..6 points per direction (these pin the point
  values and orderings that the oracle and the product generate in code).

Run in the development container:  python tests/golden/make_golden.py
"""
import os
import re

import numpy as np

REF = '/root/reference/pyfr'
OUT = os.path.dirname(os.path.abspath(__file__))


def parse_rule(path, ndim):
    pts, wts = [], []
    with open(path) as f:
        for line in f:
            vals = [float(v) for v in line.split()]
            if not vals:
                continue
            pts.append(vals[:ndim])
            wts.append(vals[ndim] if len(vals) > ndim else np.nan)
    return np.array(pts), np.array(wts)


def main():
    ref = np.load(f'{REF}/tests/hex-gleg-ord3.npz', allow_pickle=False)
    np.savez_compressed(f'{OUT}/hex_gleg_ord3.npz',
                        **{k: ref[k] for k in ref.files})

    rules = {}
    for kind, ndim, npow in (('line', 1, 1), ('quad', 2, 2), ('hex', 3, 3)):
        d = f'{REF}/quadrules/{kind}'
        for fn in sorted(os.listdir(d)):
            m = re.match(r'gauss-legendre-n(\d+)-', fn)
            if not m:
                continue
            npts = int(m[1])
            n1 = round(npts**(1/npow))
            if n1**npow != npts or n1 > 6:
                continue
            p, w = parse_rule(os.path.join(d, fn), ndim)
            rules[f'{kind}_n{n1}_pts'] = p
            rules[f'{kind}_n{n1}_wts'] = w
    np.savez_compressed(f'{OUT}/gl_rules.npz', **rules)
    print('wrote', sorted(rules))


if __name__ == '__main__':
    main()
