"""The face-moment factorisation the 13-moment face kernels rely on, checked
on the oracle's own basis tables (CPU).

On a face normal to axis d, mode r of the basis is the product of its
Legendre factor in d, evaluated at -1 or +1 (a sign), and a polynomial in the
two tangential coordinates fixed by its tangential degrees.  So the lift of
a face flux onto mode r, sum_p w2_p phi_r(p) F_p, equals that sign times one
of NK face moments M_k = sum_p w2_p b_k(p) F_p, with k the mode's pair of
tangential degrees: 3 distinct moments for nbasis 4, 4 for nbasis 8, 6 for
nbasis 10.  fz::k_face13 writes the moments and lift13_mom applies the signs.
"""
import numpy as np
import pytest

from hermeshd_amd.config import HDConfig


def _deg(nb, r, ax):
    """Legendre degree of mode r along axis ax: 1, x, y, z, yz, zx, xy, then
    xyz (nbasis 8) or P2(x), P2(y), P2(z) (nbasis 10)."""
    if r == 0:
        return 0
    if r <= 3:
        return 1 if ax == r - 1 else 0
    if r <= 6:
        return 0 if ax == r - 4 else 1
    if nb == 8:
        return 1
    return 2 if ax == r - 7 else 0


# tangential axes of a face normal to d (the row and column of its points)
_TANG = {0: (2, 1), 1: (0, 2), 2: (1, 0)}


@pytest.mark.parametrize("nb,q,nk", [(4, 2, 3), (8, 2, 4), (10, 3, 6)])
def test_face_values_factor_through_tangential_pairs(oracle_mod, nb, q, nk):
    cfg = HDConfig(nx=4, ny=4, nz=4, nbasis=nb, iquad=q, nQ=5)
    t = oracle_mod.Oracle(cfg.as_ints()).tables()
    for d, axis in enumerate("xyz"):
        for side, key in ((0, f"b{axis}m"), (1, f"b{axis}p")):
            B = t[key]
            groups = {}
            for r in range(nb):
                sign = -1.0 if (_deg(nb, r, d) == 1 and side == 0) else 1.0
                pair = (_deg(nb, r, _TANG[d][0]), _deg(nb, r, _TANG[d][1]))
                col = sign * B[:, r]
                if pair in groups:
                    assert np.array_equal(col, groups[pair]), (axis, side, r)
                else:
                    groups[pair] = col
            assert len(groups) == nk
            # the moment vectors are independent (no two pairs collapse)
            assert np.linalg.matrix_rank(np.array(list(groups.values()))) == nk


def test_mutated_degree_breaks_the_factorisation(oracle_mod):
    """A wrong degree table (P2(x) taken as degree 0 in x) mixes modes with
    different face polynomials into one pair on the x faces, and the check
    above fails."""
    cfg = HDConfig(nx=4, ny=4, nz=4, nbasis=10, iquad=3, nQ=5)
    B = oracle_mod.Oracle(cfg.as_ints()).tables()["byp"]

    def bad_deg(r, ax):
        return 0 if (r == 7 and ax == 0) else _deg(10, r, ax)
    groups, clash = {}, False
    for r in range(10):
        pair = (bad_deg(r, _TANG[1][0]), bad_deg(r, _TANG[1][1]))
        if pair in groups and not np.array_equal(B[:, r], groups[pair]):
            clash = True
        groups.setdefault(pair, B[:, r])
    assert clash
