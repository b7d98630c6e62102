"""The Taylor-Green run's kinetic energy of the whole polynomial state
(bench.kinetic_energy_quadrature), against Legendre orthogonality (CPU).

At unit density the cell mean of 0.5 m_x^2 with m_x = sum_r a_r phi_r is
0.5 sum_r a_r^2 / cbasis_r, cbasis_r = prod over axes of (2 deg + 1): the
27-point Gauss rule is exact for these degree-4 integrands.
"""
import numpy as np
import pytest
import torch

import bench


def _cbasis(nb):
    degs = [(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1), (0, 1, 1), (1, 0, 1), (1, 1, 0)]
    degs += [(1, 1, 1)] if nb == 8 else [(2, 0, 0), (0, 2, 0), (0, 0, 2)]
    return np.array([(2 * a + 1) * (2 * b + 1) * (2 * c + 1) for a, b, c in degs[:nb]], dtype=float)


@pytest.mark.parametrize("nb", [8, 10])
def test_quadrature_kinetic_energy_matches_orthogonality(nb):
    rng = np.random.default_rng(4)
    ncell = (2, 3, 4)
    q = np.zeros((nb, 5) + ncell)
    q[0, 0] = 1.0                                  # unit density, no sub-cell density modes
    coef = rng.standard_normal((nb, 3) + ncell)    # modes of m_x, m_y, m_z
    q[:, 1:4] = coef
    q[0, 4] = 10.0
    want = float((0.5 * (coef ** 2 / _cbasis(nb)[:, None, None, None, None]).sum(axis=(0, 1))).mean())
    got = bench.kinetic_energy_quadrature(torch.from_numpy(q), nb)
    assert got == pytest.approx(want, rel=1e-13)


def test_cell_average_case():
    """Mode 1 only: the quadrature value is the cell averages' kinetic energy."""
    q = np.zeros((10, 5, 1, 1, 3))
    q[0, 0] = [1.0, 2.0, 0.5]
    q[0, 1] = [0.3, -0.2, 0.1]
    q[0, 2] = [0.0, 0.4, -0.6]
    got = bench.kinetic_energy_quadrature(torch.from_numpy(q), 10)
    want = float(np.mean(0.5 * (q[0, 1] ** 2 + q[0, 2] ** 2) / q[0, 0]))
    assert got == pytest.approx(want, rel=1e-14)
