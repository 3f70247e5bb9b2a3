"""Pins the oracle spline, spline_to_grid, LFC and V-cycle -- CPU only.

Checks: the reference C build (Spline, spline_to_grid, LFC integrate/add,
fd_precond); exact interpolation at knots; brute force over all periodic
images on the full grid; adjointness of add and integrate; the l=1
orthogonality of test_derivatives.py:17-27; closed forms of the V-cycle on
constant and checkerboard inputs.
"""
from math import pi

import numpy as np
import pytest

from oracle.grid import Grid
from oracle.lfc import LFC
from oracle.precond import Preconditioner
from oracle.spline import Y_CONST, Spline
from oracle.stencil import laplace_stencil


def test_spline_knots_and_continuity(rng):
    rcut = 2.5
    r = np.linspace(0, rcut, 51)
    f = np.exp(-r**2 / 0.7)
    s = Spline(0, rcut, f)
    f[-1] = 0.0
    np.testing.assert_allclose(s(r[:-1]), f[:-1], rtol=0, atol=1e-15)
    assert s(rcut) == 0.0 and s(3.0) == 0.0
    # C1 continuity at interior knots
    eps = 1e-7
    k = r[5:45]
    left = (s(k - eps) - s(k - 2 * eps)) / eps
    right = (s(k + 2 * eps) - s(k + eps)) / eps
    assert abs(left - right).max() < 1e-5


def test_spline_against_reference(ref, rng):
    for l in (0, 1):
        r = np.linspace(0, 2.0, 120)
        f = r**l * np.exp(-r**2 / 0.5)
        s = Spline(l, 2.0, f)
        sr = ref.Spline(l, 2.0, f)
        x = rng.uniform(0, 2.2, 300)
        assert max(abs(s(v) - sr(v)) for v in x) == 0.0


def _splines(rng, natoms):
    r = np.linspace(0, 2.0, 101)
    out = []
    for a in range(natoms):
        rc = rng.uniform(0.6, 1.0)
        g = np.exp(-r**2 / rc**2)
        out.append([Spline(0, 2.0, g), Spline(0, 2.0, g * 0.5),
                    Spline(1, 2.0, g), Spline(1, 2.0, g * 0.3)])
    return out


def _brute_dense(gd, spline_aj, spos_ac, k_c):
    """Functions summed over all periodic images, evaluated on every point."""
    r_vg = gd.coordinates().reshape(3, -1)
    N = r_vg.shape[1]
    rows = []
    for spos_c, spline_j in zip(spos_ac, spline_aj):
        for s in spline_j:
            nm = 2 * s.l + 1
            F = np.zeros((nm, N), complex)
            for i in range(-1, 2):
                for j in range(-1, 2):
                    for k in range(-1, 2):
                        n_c = np.array([i, j, k])
                        pos = np.dot(spos_c + n_c, gd.cell_cv)
                        d = r_vg - pos[:, None]
                        rr = np.sqrt((d**2).sum(0))
                        f = s(rr)
                        ph = np.exp(-2j * pi * np.dot(k_c, -n_c))
                        if s.l == 0:
                            Y = [f * Y_CONST[0]]
                        else:
                            c = Y_CONST[1]
                            Y = [f * c * d[1], f * c * d[2], f * c * d[0]]
                        F += np.array(Y) * ph
            rows.extend(F)
    return np.array(rows)


@pytest.mark.parametrize('k_c', [None, (0.25, 0.0, 0.1)])
def test_lfc_brute_force(rng, k_c):
    gd = Grid((14, 14, 16), np.array([14, 14, 16]) * 0.38)
    spline_aj = _splines(rng, 3)
    spos_ac = rng.uniform(0, 1, (3, 3))
    spos_ac[0] = (0.01, 0.98, 0.5)
    lfc = LFC(gd, spline_aj, k_c, complex if k_c else float)
    lfc.set_positions(spos_ac)
    B = _brute_dense(gd, spline_aj, spos_ac, np.zeros(3) if k_c is None
                     else np.array(k_c))
    a = rng.standard_normal((2,) + tuple(gd.n_c))
    if k_c:
        a = a + 1j * rng.standard_normal(a.shape)
    c = lfc.integrate(a)
    c_brute = gd.dv * a.reshape(2, -1) @ np.conj(B).T
    assert abs(c - c_brute).max() < 1e-13
    coef = rng.standard_normal(c.shape) + (1j * rng.standard_normal(c.shape)
                                           if k_c else 0)
    out = lfc.add(np.zeros_like(a), coef)
    out_brute = (coef @ B).reshape(a.shape)
    assert abs(out - out_brute).max() < 1e-13


def test_lfc_adjoint(rng):
    gd = Grid((12, 14, 16), np.array([12, 14, 16]) * 0.35)
    lfc = LFC(gd, _splines(rng, 4), (0.25, 0.1, 0.0), complex)
    lfc.set_positions(rng.uniform(0, 1, (4, 3)))
    psi = rng.standard_normal((3,) + tuple(gd.n_c)) * (1 + 0.5j)
    c = rng.standard_normal((3, lfc.Mmax)) + 1j * rng.standard_normal(
        (3, lfc.Mmax))
    lhs = np.vdot(lfc.add(np.zeros_like(psi), c), psi)
    rhs = np.vdot(c, lfc.integrate(psi)) / gd.dv
    assert abs(lhs - rhs) < 1e-11 * abs(lhs)


def test_lfc_l1_orthogonality():
    """test_derivatives.py:7-27: add(I) then integrate is diagonal for l=1."""
    a = 4.0
    gd = Grid((16, 20, 20), [a, a + 1, a + 2], (0, 1, 1))
    spos_ac = np.array([[0.25, 0.15, 0.35], [0.5, 0.5, 0.5]])
    s = Spline(0, 2.0, np.array([1, 0.9, 0.1, 0.0]))
    p = Spline(1, 2.0, np.array([1, 0.9, 0.1, 0.0]))
    lfc = LFC(gd, [[s], [s, p]], cut=True)
    lfc.set_positions(spos_ac)
    C = np.zeros((3, lfc.Mmax))
    C[:, 2:] = np.eye(3)
    psi = lfc.add(gd.zeros(3), C)
    C2 = lfc.integrate(psi)[:, 2:]
    d = C2.diagonal()
    assert np.ptp(d) < 4e-6
    C2 -= np.diag(d)
    assert abs(C2).max() < 5e-17


@pytest.mark.parametrize('k_c,dtype', [(None, float), ((0.25, 0, 0.1),
                                                       complex)])
def test_lfc_against_reference(ref, rng, k_c, dtype):
    gd = Grid((12, 14, 16), np.array([12, 14, 16]) * 0.35)
    spline_aj = _splines(rng, 3)
    spos_ac = rng.uniform(0, 1, (3, 3))
    spos_ac[0] = (0.02, 0.97, 0.5)
    lfc = LFC(gd, spline_aj, k_c, dtype)
    lfc.set_positions(spos_ac)
    R = ref.LFC(gd, spline_aj, spos_ac, k_c, dtype)
    assert R.Mmax == lfc.Mmax
    a = rng.standard_normal((2,) + tuple(gd.n_c))
    if dtype == complex:
        a = a + 1j * rng.standard_normal(a.shape)
    c = lfc.integrate(a)
    assert abs(c - R.integrate(a)).max() < 1e-14 * abs(c).max()
    coef = rng.standard_normal(c.shape) + (0 if dtype == float else
                                           1j * rng.standard_normal(c.shape))
    b1 = lfc.add(a.copy(), coef)
    b2 = R.add(a.copy(), coef)
    assert abs(b1 - b2).max() < 1e-14 * abs(b2).max()


@pytest.mark.parametrize('real', [True, False])
def test_precond_against_reference(ref, rng, real):
    gd0 = Grid((16, 16, 20), np.array([16, 16, 20]) * 0.3)
    c0, o0 = laplace_stencil(gd0.h_cv, -0.5, 3)
    R = rng.standard_normal((3,) + tuple(gd0.n_c))
    ph = None
    if not real:
        R = R + 1j * rng.standard_normal(R.shape)
        ph = gd0.phase_cd((0.25, 0, 0))
    d = Preconditioner(gd0, c0, o0)(R, ph)
    dr = ref.fd_precond(gd0, c0, o0, R, ph)
    assert abs(d - dr).max() <= 1e-14 * abs(dr).max()


def test_precond_closed_forms():
    """Constant input: restriction and interpolation preserve constants and
    the Laplacians annihilate them, so the 14 steps give d0 = -41 s c.
    Checkerboard (Nyquist) input: the nn=1 restriction removes it, so only the
    last smoothing step acts and d0 = -s R."""
    gd0 = Grid((16, 16, 16), [16 * 0.35] * 3)
    c0, o0 = laplace_stencil(gd0.h_cv, -0.5, 3)
    P = Preconditioner(gd0, c0, o0)
    s = P.step
    R = np.full((16, 16, 16), 0.7)
    np.testing.assert_allclose(P(R), -41 * s * R, rtol=1e-13)
    i = np.indices((16, 16, 16)).sum(0)
    R = (-1.0)**i
    np.testing.assert_allclose(P(R), -s * R, atol=1e-14)
    assert s == pytest.approx(0.66666666 / c0[0])
