"""The product's Laplacian points and coefficients (gpaw_amd.fd_operators.
guc_laplace_stencil) against the oracle's independent derivation
(oracle/stencil.py laplace_stencil, pinned by test_oracle_stencil.py) and the
closed form of an orthorhombic cell -- CPU only, no library load."""
from math import cos, pi, sin

import numpy as np
import pytest

from gpaw_amd.fd_operators import guc_laplace_stencil
from oracle.grid import Grid
from oracle.stencil import laplace_stencil

CELLS = [
    ('distorted hexagonal', 4,
     [(1, 0, 0), (1.02 * cos(pi / 3 - 0.02), 1.02 * sin(pi / 3 - 0.02), 0),
      (0, 0, 1.0)]),
    ('hexagonal', 4, [(1, 0, 0), (0.5, 3**0.5 / 2, 0), (0, 0, 1.1)]),
    ('fcc', 6, [(0, 1, 1), (1, 0, 1), (1, 1, 0)]),
    ('bcc', 4, [(-1, 1, 1), (1, -1, 1), (1, 1, -1)]),
    ('sc', 3, [1.1, 1.02, 1.03]),
    ('distorted sc', 6, [(1, 0, 0), (0.01, 1, 0), (0, 0.02, 1)]),
    ('MnO', 6, [(1, 0.5, 0.5), (0.5, 1, 0.5), (0.5, 0.5, 1)]),
]


@pytest.mark.parametrize('name,D,cell', CELLS)
@pytest.mark.parametrize('N_c', [(8, 8, 8), (8, 12, 16)])
def test_matches_oracle(name, D, cell, N_c):
    gd = Grid(N_c, cell)
    for n in range(1, 5):
        c0, o0 = laplace_stencil(gd.h_cv, -0.5, n)
        c1, o1, D1 = guc_laplace_stencil(gd.h_cv, -0.5, n)
        assert len(c1) == 2 * n * D1 + 1 == len(c0)
        if N_c == (8, 8, 8):
            assert D1 == D
        np.testing.assert_array_equal(o1, o0)
        np.testing.assert_allclose(c1, c0, rtol=0, atol=2e-15 * abs(c0).max())


def test_orthorhombic_closed_form():
    h = np.array([0.3, 0.35, 0.4])
    c, o, D = guc_laplace_stencil(np.diag(h), -0.5, 3)
    w = [-49 / 18, 3 / 2, -3 / 20, 1 / 90]
    assert D == 3
    assert c[0] == pytest.approx(-0.5 * w[0] * (1 / h**2).sum(), rel=1e-15)
    for p in range(1, len(c)):
        axis = int(np.flatnonzero(o[p])[0])
        j = abs(int(o[p][axis]))
        assert c[p] == pytest.approx(-0.5 * w[j] / h[axis]**2, rel=1e-15)
    # the z axis comes first, then y, then x; +1..+3 before -1..-3
    np.testing.assert_array_equal(o[1:7, 2], [1, 2, 3, -1, -2, -3])
    np.testing.assert_array_equal(o[13:19, 0], [1, 2, 3, -1, -2, -3])


def test_polynomial_exactness_general_cells():
    """Second derivatives of quadratics are exact on every cell."""
    for name, D, cell in CELLS:
        gd = Grid((8, 8, 8), cell)
        c, o, _ = guc_laplace_stencil(gd.h_cv, 1.0, 2)
        t = o @ gd.h_cv
        for A in (np.diag([1.0, 2.0, 3.0]),
                  np.array([[0, 1, 0.5], [1, 0, 0], [0.5, 0, 2.0]])):
            # f(r) = r^T A r: Laplacian 2 tr A, at the centre f(0) = 0
            val = sum(cp * tv @ A @ tv for cp, tv in zip(c, t))
            assert val == pytest.approx(2 * np.trace(A), abs=1e-10), name
