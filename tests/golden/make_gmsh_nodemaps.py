"""Write tests/golden/gmsh_nodemaps.json: a SHA-256 digest of each Gmsh ->
PyFR node map the reference tabulates (``_nodemaps`` in
``pyfr/readers/gmsh.py:58-150``) for the shapes this build reads (tets,
hexes, prisms, pyramids and their tri / quad faces; lines are left out).
The reference's file is parsed as text (``ast.literal_eval`` of the one
dict literal; nothing imported or executed), and only the digests are kept:
they pin the maps the reader derives from Gmsh's ordering rules without
holding the reference's tables.

Left out as well: ('tri', 21).  Its interior nodes follow another order
than every face interior of the volume maps, and the reference reader
cannot reach it (no Gmsh type of its ``_etype_map``, gmsh.py:34-42, has 21
nodes).

Run in the development container:  python tests/golden/make_gmsh_nodemaps.py
"""
import ast
import hashlib
import json
import os

import numpy as np

SRC = '/root/reference/pyfr/readers/gmsh.py'
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)),
                   'gmsh_nodemaps.json')
KINDS = {'tet', 'hex', 'pri', 'pyr', 'tri', 'quad'}
SKIP = {('tri', 21)}


def digest(m):
    return hashlib.sha256(np.asarray(m, dtype=np.int64).tobytes()).hexdigest()


def main():
    with open(SRC) as f:
        tree = ast.parse(f.read())
    maps = None
    for node in ast.walk(tree):
        if (isinstance(node, ast.Assign) and
                any(getattr(t, 'id', None) == '_nodemaps'
                    for t in node.targets)):
            maps = ast.literal_eval(node.value)
    if maps is None:
        raise SystemExit('no _nodemaps table found')
    out = {f'{k}-{n}': digest(m) for (k, n), m in sorted(maps.items())
           if k in KINDS and (k, n) not in SKIP}
    with open(OUT, 'w') as f:
        json.dump(out, f, indent=1, sort_keys=True)
        f.write('\n')
    print(f'wrote {len(out)} digests to {OUT}')


if __name__ == '__main__':
    main()
