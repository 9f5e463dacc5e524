"""Regenerate pyfr_amd/quadtables.py from the published point sets held as
text tables by the reference (``pyfr/quadrules/{tet,tri,pri,pyr}/*.txt``):

* tetrahedron solution points: Shunn & Ham (2012), n = 4, 10, 20, 35;
* triangle flux points: Williams & Shunn (2014), n = 3, 6, 10, 15, and the
  36-point rule the reference integrates triangular faces with;
* prism solution points: the Williams-Shunn triangle x Gauss-Legendre line
  product, n = 6, 18, 40, 75;
* pyramid solution points: the reference's Gauss-Legendre pyramid sets,
  n = 5, 14, 30, 55 (points only; no weights);
* the other solution / flux point families the reference offers at
  p = 1..4: alpha-optimised tets and triangles, alpha-optimised and
  Williams-Shunn triangles x Gauss-Legendre-Lobatto lines for prisms, and
  Gauss-Legendre-Lobatto pyramids;
* element quadrature rules of flux anti-aliasing at p = 1..4: the rule the
  reference's default selection picks (``pyfr/shapes.py:146-176``, with
  ``pyfr/quadrules/__init__.py:52-79``: the fewest points among the rules of
  degree >= 2p + 1 for tets, >= 2(p + 2) for prisms).  Tets: Witherden &
  Vincent (2015), n = 8, 14, 35, 59.  Prisms: Witherden & Vincent n = 28
  (p = 1), Witherden n = 46, 82, 128 (p = 2..4).  At p = 2 two 46-point
  prism rules of degree 8 tie and the reference keeps the first its
  directory listing yields; this table takes the first in name order;
* triangle-face rules of surface-flux anti-aliasing at p = 1..4: the
  fewest points among the symmetric ('s') rules of degree >= 2(p + 1)
  (``pyfr/shapes.py:159-185``): Williams-Shunn n = 6, then Witherden &
  Vincent n = 12, 16, 25.

* CATALOG[shape]: (rule, npts, degree, flags) of every rule file of stated
  degree the reference holds for tets, prisms, pyramids, triangles, hexes
  and quads
  (from the file names, ``<rule>-n<npts>-d<degree>-<flags>.txt``), which
  ``quad-deg`` selection searches (``pyfr/quadrules/__init__.py:52-79``);
  the points of the rules that selection picks for degrees up to 10 (tets,
  prisms, pyramids) and 12 (triangles), over all families and within the
  solution points' families (the default of the initial conditions'
  projection), are tabulated too.

These are data (point coordinates and weights), needed bit for bit because
the FR operators are defined by them.  Run in the development container only:
    python -m pyfr_amd.make_quadtables
"""
import os
import re

SRC = '/root/reference/pyfr/quadrules'
RULES = {
    ('tet', 'shunn-ham', 4): 'tet/shunn-ham-n4-d2-spu.txt',
    ('tet', 'shunn-ham', 10): 'tet/shunn-ham-n10-d3-spu.txt',
    ('tet', 'shunn-ham', 20): 'tet/shunn-ham-n20-d5-spu.txt',
    ('tet', 'shunn-ham', 35): 'tet/shunn-ham-n35-d6-spu.txt',
    ('tri', 'williams-shunn', 3): 'tri/williams-shunn-n3-d2-spu.txt',
    ('tri', 'williams-shunn', 6): 'tri/williams-shunn-n6-d4-spu.txt',
    ('tri', 'williams-shunn', 10): 'tri/williams-shunn-n10-d5-spu.txt',
    ('tri', 'williams-shunn', 15): 'tri/williams-shunn-n15-d7-spu.txt',
    ('tri', 'williams-shunn', 36): 'tri/williams-shunn-n36-d12-spu.txt',
    ('pri', 'williams-shunn~gauss-legendre', 6):
        'pri/williams-shunn~gauss-legendre-n6-d2-pstu.txt',
    ('pri', 'williams-shunn~gauss-legendre', 18):
        'pri/williams-shunn~gauss-legendre-n18-d4-pstu.txt',
    ('pri', 'williams-shunn~gauss-legendre', 40):
        'pri/williams-shunn~gauss-legendre-n40-d5-pstu.txt',
    ('pri', 'williams-shunn~gauss-legendre', 75):
        'pri/williams-shunn~gauss-legendre-n75-d7-pstu.txt',
    ('tet', 'witherden-vincent', 8): 'tet/witherden-vincent-n8-d3-sp.txt',
    ('tet', 'witherden-vincent', 14): 'tet/witherden-vincent-n14-d5-sp.txt',
    ('tet', 'witherden-vincent', 35): 'tet/witherden-vincent-n35-d7-sp.txt',
    ('tet', 'witherden-vincent', 59): 'tet/witherden-vincent-n59-d9-sp.txt',
    ('pri', 'witherden-vincent', 28): 'pri/witherden-vincent-n28-d6-sp.txt',
    ('pri', 'witherden', 46): 'pri/witherden-n46-d8-sp.txt',
    ('pri', 'witherden', 82): 'pri/witherden-n82-d10-sp.txt',
    ('pri', 'witherden', 128): 'pri/witherden-n128-d12-sp.txt',
    ('tri', 'witherden-vincent', 12): 'tri/witherden-vincent-n12-d6-sp.txt',
    ('tri', 'witherden-vincent', 16): 'tri/witherden-vincent-n16-d8-sp.txt',
    ('tri', 'witherden-vincent', 25): 'tri/witherden-vincent-n25-d10-sp.txt',
    ('pyr', 'gauss-legendre', 5): 'pyr/gauss-legendre-n5-su.txt',
    ('pyr', 'gauss-legendre', 14): 'pyr/gauss-legendre-n14-su.txt',
    ('pyr', 'gauss-legendre', 30): 'pyr/gauss-legendre-n30-su.txt',
    ('pyr', 'gauss-legendre', 55): 'pyr/gauss-legendre-n55-su.txt',
}
# the other solution point families of [solver-elements-<type>] soln-pts
# and [solver-interfaces-tri] flux-pts at p = 1..4
for _n in (4, 10, 20, 35):
    RULES[('tet', 'alpha-opt', _n)] = f'tet/alpha-opt-n{_n}-su.txt'
for _n, _d in ((6, 1), (18, 3), (40, 5), (75, 7)):
    RULES[('pri', 'alpha-opt~gauss-legendre-lobatto', _n)] = \
        f'pri/alpha-opt~gauss-legendre-lobatto-n{_n}-stu.txt'
    RULES[('pri', 'williams-shunn~gauss-legendre-lobatto', _n)] = \
        f'pri/williams-shunn~gauss-legendre-lobatto-n{_n}-d{_d}-pstu.txt'
for _n in (5, 14, 30, 55):
    RULES[('pyr', 'gauss-legendre-lobatto', _n)] = \
        f'pyr/gauss-legendre-lobatto-n{_n}-su.txt'
for _n in (3, 6, 10, 15):
    RULES[('tri', 'alpha-opt', _n)] = f'tri/alpha-opt-n{_n}-su.txt'
NDIM = {'tet': 3, 'tri': 2, 'pri': 3, 'pyr': 3}
QDEG_MAX = {'tet': 10, 'pri': 10, 'tri': 12, 'pyr': 10}
# the solution points' families, which the projection of initial
# conditions searches by default (pyfr/solvers/base/elements.py:94-108)
IC_FAMILIES = {'tet': ('shunn-ham',),
               'pri': ('williams-shunn~gauss-legendre',)}


def catalog():
    out = {}
    for shape in ('tet', 'pri', 'pyr', 'tri', 'hex', 'quad'):
        rows = []
        for fn in sorted(os.listdir(os.path.join(SRC, shape))):
            m = re.match(r'([a-zA-Z0-9\-~+]+)-n(\d+)(?:-d(\d+))?'
                         r'(?:-([pstu]+))?\.txt$', fn)
            if m and m[3]:
                rows.append((m[1], int(m[2]), int(m[3]), m[4] or '', fn))
        out[shape] = rows
    return out


def select(rows, qdeg, name=None):
    """The fewest points among the rules of degree >= qdeg (ties: the
    first in name order)."""
    best = None
    for r in rows:
        if (name is None or r[0] == name) and r[2] >= qdeg:
            if best is None or r[1] < best[1]:
                best = r
    return best


def main():
    cat = catalog()
    rules = dict(RULES)
    for shape, qmax in QDEG_MAX.items():
        for q in range(1, qmax + 1):
            for fam in (None, *IC_FAMILIES.get(shape, ())):
                r = select(cat[shape], q, fam)
                if r:
                    rules.setdefault((shape, r[0], r[1]), f'{shape}/{r[4]}')
    out = ['"""Tabulated point sets (generated by make_quadtables.py; see its',
           'docstring for provenance).  TABLES[(shape, rule, npts)] =',
           '(points, weights) with points as tuples of floats; CATALOG[shape]',
           '= (rule, npts, degree, flags) of the rules of stated degree."""',
           '',
           'CATALOG = {']
    for shape, rows in cat.items():
        out.append(f'    {shape!r}: [')
        out += [f'        {r[:4]!r},' for r in rows]
        out.append('    ],')
    out += ['}', '', 'TABLES = {']
    for key, fn in rules.items():
        pts, wts = [], []
        with open(os.path.join(SRC, fn)) as f:
            for line in f:
                v = [float(x) for x in line.split()]
                if not v:
                    continue
                nd = NDIM[key[0]]
                pts.append(tuple(v[:nd]))
                wts.append(v[nd] if len(v) > nd else None)
        out.append(f'    {key!r}: (')
        out.append('        [' + ',\n         '.join(
            '(' + ', '.join(repr(c) for c in p) + ')' for p in pts) + '],')
        out.append('        [' + ', '.join(repr(w) for w in wts) + ']),')
    out.append('}')
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)),
                       'quadtables.py')
    with open(dst, 'w') as f:
        f.write('\n'.join(out) + '\n')
    print('wrote', dst)


if __name__ == '__main__':
    main()
