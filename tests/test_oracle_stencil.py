"""Pins the oracle stencil (oracle/stencil.py) -- CPU only.

Checks: the reference C build (bc_unpack + bmgs_fd), polynomial exactness as
in the reference's test_laplace.py:13-101, the plane-wave symbol of the
19-point operator with and without Bloch phases, and the analytic Laplacian of
a Gaussian (test_fd_laplace.py:33-57).
"""
from math import cos, pi, sin

import numpy as np
import pytest

from oracle.grid import Grid
from oracle.stencil import fd_apply, laplace_stencil

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
def test_polynomial_exactness(name, D, cell):
    for n in range(1, 5):
        N = 2 * n + 2
        gd = Grid((N, N, N), cell)
        coef_p, off_pc = laplace_stencil(gd.h_cv, 1.0, n)
        assert len(coef_p) == D * 2 * n + 1
        r_gv = gd.coordinates().transpose((1, 2, 3, 0))
        r_gv -= gd.cell_cv.sum(0) / 2
        for m in range(0, 2 * n + 1):
            for ix in range(m + 1):
                for iy in range(m - ix + 1):
                    iz = m - ix - iy
                    a_g = (r_gv ** (ix, iy, iz)).prod(3)
                    r = 2.0 if (m == 2 and max(ix, iy, iz) == 2) else 0.0
                    b_g = fd_apply(coef_p, off_pc, a_g, gd.pbc_c)
                    assert abs(b_g[n + 1, n + 1, n + 1] - r) < 2e-12


def test_coefficients_closed_form():
    """19-point, h = 0.2 A: centre and +-1..3 coefficients (scale -0.5)."""
    h = 0.2 / 0.52917721067
    gd = Grid((112, 112, 112), [112 * h] * 3)
    coef_p, off_pc = laplace_stencil(gd.h_cv, -0.5, 3)
    assert len(coef_p) == 19
    a = 1 / h**2
    np.testing.assert_allclose(coef_p[0], -0.5 * (-49 / 18) * 3 * a,
                               rtol=1e-14)
    for j, c in zip((1, 2, 3), (3 / 2, -3 / 20, 1 / 90)):
        sel = [p for p, o in enumerate(off_pc) if abs(o).max() == j]
        assert len(sel) == 6
        np.testing.assert_allclose(coef_p[sel], -0.5 * c * a, rtol=1e-14)
    np.testing.assert_allclose(coef_p[0], 28.58624, rtol=1e-6)


@pytest.mark.parametrize('k_c', [(0, 0, 0), (0.25, 0, 0), (0.1, -0.3, 0.45)])
def test_plane_wave_symbol(k_c):
    N_c = np.array([8, 10, 12])
    h_c = np.array([0.3, 0.35, 0.4])
    gd = Grid(N_c, N_c * h_c)
    coef_p, off_pc = laplace_stencil(gd.h_cv, -0.5, 3)
    m_c = np.array([1, -2, 3])
    q_c = 2 * pi * (m_c + np.array(k_c)) / (N_c * h_c)
    r_vg = gd.coordinates()
    psi = np.exp(1j * np.einsum('v,vxyz->xyz', q_c, r_vg))
    out = fd_apply(coef_p, off_pc, psi, gd.pbc_c, gd.phase_cd(k_c))
    w = [-49 / 18, 3 / 2, -3 / 20, 1 / 90]
    lam = -0.5 * sum((w[0] + 2 * sum(w[j] * cos(j * q * h)
                                      for j in (1, 2, 3))) / h**2
                     for q, h in zip(q_c, h_c))
    np.testing.assert_allclose(out, lam * psi, atol=1e-12 * abs(lam))


def test_gaussian_laplacian():
    lat = 8.0
    gd = Grid((32, 32, 32), (lat, lat, lat), False)
    x, y, z = gd.coordinates()
    sigma, mu = 1.5, lat / 2
    r2 = (x - mu)**2 + (y - mu)**2 + (z - mu)**2
    a = np.exp(-r2 / (2.0 * sigma))
    b_analytic = (r2 / sigma**2 - 3.0 / sigma) * a
    coef_p, off_pc = laplace_stencil(gd.h_cv, 1.0, 3)
    b = fd_apply(coef_p, off_pc, a, gd.pbc_c)
    assert abs(b - b_analytic)[2:-2, 2:-2, 2:-2].max() < 1e-2


CASES = [((8, 10, 12), 0.4, True, 3), ((8, 10, 12), 0.4, (True, False, True),
                                         3),
         ((6, 7, 9), 0.5, True, 1), ((12, 12, 12), 0.3, False, 2)]


@pytest.mark.parametrize('N_c,h,pbc,n', CASES)
@pytest.mark.parametrize('real', [True, False])
def test_against_reference_build(ref, rng, N_c, h, pbc, n, real):
    gd = Grid(N_c, np.array(N_c) * h, pbc)
    coef_p, off_pc = laplace_stencil(gd.h_cv, -0.5, n)
    a = rng.standard_normal((3,) + tuple(gd.n_c))
    ph = None
    if not real:
        a = a + 1j * rng.standard_normal(a.shape)
        ph = gd.phase_cd((0.25, -0.1, 0.3))
    out = np.zeros_like(a)
    ref.Operator(coef_p, off_pc, gd, real).apply(
        a, out, ph if ph is not None else None)
    mine = fd_apply(coef_p, off_pc, a, gd.pbc_c, ph)
    scale = abs(out).max()
    assert abs(mine - out).max() <= 1e-14 * scale


def test_against_reference_build_fcc(ref, rng):
    gd = Grid((8, 8, 8), [(0, 2, 2), (2, 0, 2), (2, 2, 0)])
    coef_p, off_pc = laplace_stencil(gd.h_cv, -0.5, 2)
    a = rng.standard_normal((2, 8, 8, 8))
    out = np.zeros_like(a)
    ref.Operator(coef_p, off_pc, gd, True).apply(a, out, None)
    np.testing.assert_allclose(fd_apply(coef_p, off_pc, a, gd.pbc_c), out,
                               rtol=0, atol=1e-13 * abs(out).max())
