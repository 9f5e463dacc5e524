"""Boundary faces of non-periodic boxes (CPU): every face is either paired
once or on exactly one named side, and the sides agree with the oracle's
geometric brute force."""
import numpy as np
import pytest

from oracle.geom import trilinear_ploc
from oracle.mesh import box_hex_verts, pair_faces_bruteforce
from oracle.basis import HexShapeOracle
from pyfr_amd.mesh import HEX_FACE_AXES, BoxMesh


@pytest.mark.parametrize('periodic', [(False, True, True),
                                      (False, False, True),
                                      (True, False, False),
                                      (False, False, False)])
def test_box_boundaries(periodic):
    n = 3
    m = BoxMesh(n, periodic=periodic)
    lhs, rhs = m.con
    faces = lhs + rhs + [f for v in m.bcon.values() for f in v]
    assert len(faces) == len(set(faces)) == 6*n**3
    want = {f'{"xyz"[a]}{s}' for a in range(3) if not periodic[a]
            for s in (0, 1)}
    assert set(m.bcon) == want
    for name, fs in m.bcon.items():
        ax, hi = 'xyz'.index(name[0]), name[1] == '1'
        assert len(fs) == n*n
        for _, e, f in fs:
            assert HEX_FACE_AXES[f] == (ax, 1 if hi else -1)
            ijk = (e % n, (e // n) % n, e // (n*n))
            assert ijk[ax] == (n - 1 if hi else 0)
    # the oracle's geometric pairing finds the same boundary faces
    sh = HexShapeOracle(1)
    plocf = trilinear_ploc(box_hex_verts(n), sh.fpts)
    bnd = {}
    pairs = pair_faces_bruteforce(plocf, 6, sh.nfacefpts, 2*np.pi,
                                  periodic, bnd)
    assert len(pairs) == len(lhs)
    for name, fs in bnd.items():
        assert sorted((e, f) for e, f, _ in fs) == \
            sorted((e, f) for _, e, f in m.bcon[name])


@pytest.mark.parametrize('n,periodic', [(3, (True, True, True)),
                                        (4, (False, True, True))])
def test_pair_orientation_matches_oracle(n, periodic):
    """Left/right sides as the reference picks them (smaller (face,
    element) key on the left): the product's pairs and the oracle's
    geometric ones agree side for side, which the asymmetric HLL flux
    needs."""
    m = BoxMesh(n, periodic=periodic)
    sh = HexShapeOracle(1)
    plocf = trilinear_ploc(box_hex_verts(n), sh.fpts)
    pairs = pair_faces_bruteforce(plocf, 6, sh.nfacefpts, 2*np.pi,
                                  periodic, {})
    orc = sorted((eL, fL, eR, fR) for eL, fL, _, eR, fR, _ in pairs)
    lhs, rhs = m.con
    prod = sorted((l[1], l[2], r[1], r[2]) for l, r in zip(lhs, rhs))
    assert orc == prod
    # and the rule itself: (face, element) of the left side is smaller
    assert all((fl, el) < (fr, er) for el, fl, er, fr in prod)
