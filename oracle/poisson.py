"""Jacobi relaxation and the multigrid FD Poisson solver -- oracle
restatement (test infrastructure).

* relax_jacobi: Operator.relax with relax_method 2 (c/operators.c:55-95,
  c/bmgs/relax.c Jacobi branch): n sweeps of
  f <- (1 - w) f + w (s - sum_{c != 0} coef_c f[shift_c]) / coef_0
  with the right-hand side built from the previous sweep's f (padded copy).
* relax_gauss_seidel: relax_method 1 (c/bmgs/relax.c:12-36): n sweeps
  over the points in C order, each solving its own equation in place,
  x = (s - sum_{c >= 1} coef_c a[p + offset_c]) / coef_0 with the
  coefficients in stencil order; `a` is a padded copy whose ghosts are
  filled from f before each sweep (periodic images or zero), so wrapped
  neighbours keep the sweep's starting values.
* FDPoissonSolver: gpaw/poisson.py:318-546 -- levels from repeated
  coarsening (at most 8), operators[0] = Laplace(gd, -1/(4 pi), nn) and
  7-point Laplacians on the coarse levels, nn=1 transformers, pre/post
  smoothing counts [2, 4, 4, ...] / [1, 4, 4, ...] (8 / 8 on the coarsest
  level), Jacobi weights [2/3, 1, 1, ...]; iterate2 is the recursive
  V-cycle; solve_neutral iterates until dv sum r^2 <= eps and removes the
  mean for fully periodic cells.

Pinned by tests/test_oracle_poisson.py: relax_jacobi and
relax_gauss_seidel against the reference C (Operator.relax of oracle/_ref),
relax_jacobi against a direct NumPy sweep, relax_gauss_seidel by its
defining property (each point's equation holds with its earlier
neighbours' new values); the
solver's converged potential against the exact solution of the same
discrete equation from FFT (the stencil is diagonal in plane waves).
"""
import numpy as np

from .stencil import laplace_stencil, fd_apply
from .transformers import Transformer


def relax_jacobi(coefs, offsets, f, s, n, w, pbc_c):
    """n Jacobi sweeps in place on f (one grid)."""
    offsets = np.asarray(offsets)
    c0 = coefs[np.flatnonzero((offsets == 0).all(1))[0]]
    for _ in range(n):
        Lf = fd_apply(coefs, offsets, f[None], pbc_c)[0]
        x = Lf - c0 * f                       # off-centre part
        f[:] = (1.0 - w) * f + w * (s - x) / c0
    return f


def relax_gauss_seidel(coefs, offsets, f, s, n, pbc_c):
    """n Gauss-Seidel sweeps in place on f (one grid; small grids only:
    a Python loop over the points)."""
    offsets = np.asarray(offsets, int)
    coefs = np.asarray(coefs, float)
    i0 = np.flatnonzero((offsets == 0).all(1))[0]
    others = [c for c in range(len(coefs)) if c != i0]
    r = np.abs(offsets).max(0)
    n_c = np.array(f.shape)
    inv = 1.0 / coefs[i0]
    for _ in range(n):
        # padded copy, ghosts from the sweep's starting values
        a = np.zeros(tuple(n_c + 2 * r))
        a[r[0]:r[0] + n_c[0], r[1]:r[1] + n_c[1], r[2]:r[2] + n_c[2]] = f
        for ax in range(3):
            if pbc_c[ax] and r[ax] > 0:
                sl = [slice(None)] * 3
                src = [slice(None)] * 3
                sl[ax] = slice(0, r[ax])
                src[ax] = slice(n_c[ax], n_c[ax] + r[ax])
                a[tuple(sl)] = a[tuple(src)]
                sl[ax] = slice(n_c[ax] + r[ax], n_c[ax] + 2 * r[ax])
                src[ax] = slice(r[ax], 2 * r[ax])
                a[tuple(sl)] = a[tuple(src)]
        for p0 in range(n_c[0]):
            for p1 in range(n_c[1]):
                for p2 in range(n_c[2]):
                    q = (p0 + r[0], p1 + r[1], p2 + r[2])
                    x = 0.0
                    for c in others:
                        o = offsets[c]
                        x += a[q[0] + o[0], q[1] + o[1], q[2] + o[2]] * coefs[c]
                    x = (s[p0, p1, p2] - x) * inv
                    a[q] = x
                    f[p0, p1, p2] = x
    return f


class FDPoissonSolver:
    def __init__(self, nn=3, eps=2e-10, maxiter=1000, relax='J'):
        self.relax = relax
        self.nn = nn
        self.eps = eps
        self.maxiter = maxiter

    def set_grid_descriptor(self, gd):
        self.gd = gd
        scale = -0.25 / np.pi
        self.gds = [gd]
        self.operators = [laplace_stencil(gd.h_cv, scale, self.nn)]
        self.interpolators = []
        self.restrictors = []
        self.presmooths = [2]
        self.postsmooths = [1]
        self.weights = [2.0 / 3.0]
        level = 0
        while level < 8:
            if (gd.N_c % 2).any():
                break
            gd2 = gd.coarsen()
            self.gds.append(gd2)
            self.operators.append(laplace_stencil(gd2.h_cv, scale, 1))
            self.interpolators.append(Transformer(gd2, gd, 1))
            self.restrictors.append(Transformer(gd, gd2, 1))
            self.presmooths.append(4)
            self.postsmooths.append(4)
            self.weights.append(1.0)
            level += 1
            gd = gd2
        self.levels = level
        self.presmooths[level] = 8
        self.postsmooths[level] = 8
        self.phis = [None] + [g.zeros() for g in self.gds[1:]]
        self.rhos = [g.zeros() for g in self.gds]

    def _apply(self, level, a):
        c, o = self.operators[level]
        return fd_apply(c, o, a[None], self.gds[level].pbc_c)[0]

    def _relax(self, level, n):
        c, o = self.operators[level]
        if self.relax == 'GS':
            relax_gauss_seidel(c, o, self.phis[level], self.rhos[level], n,
                               self.gds[level].pbc_c)
        else:
            relax_jacobi(c, o, self.phis[level], self.rhos[level], n,
                         self.weights[level], self.gds[level].pbc_c)

    def iterate2(self, level=0):
        if level < self.levels:
            self._relax(level, self.presmooths[level])
            residual = self._apply(level, self.phis[level]) - self.rhos[level]
            self.rhos[level + 1][:] = self.restrictors[level].apply(
                residual[None])[0]
            self.phis[level + 1][:] = 0.0
            self.iterate2(level + 1)
            self.phis[level] -= self.interpolators[level].apply(
                self.phis[level + 1][None])[0]
        self._relax(level, self.postsmooths[level])
        if level == 0:
            residual = self._apply(0, self.phis[0]) - self.rhos[0]
            return float(np.dot(residual.ravel(), residual.ravel())) * \
                self.gds[0].dv

    def solve_neutral(self, phi, rho):
        self.phis[0] = phi
        self.rhos[0][:] = rho
        niter = 1
        while self.iterate2() > self.eps and niter < self.maxiter:
            niter += 1
        if niter == self.maxiter:
            raise RuntimeError('Poisson solver did not converge in %d '
                               'iterations!' % self.maxiter)
        if self.gd.pbc_c.all():
            phi -= phi.mean()
        return niter


def fft_solution(gd, coefs, offsets, rho):
    """Exact solution of L phi = rho (zero mean) for a periodic grid and a
    stencil L, from its plane-wave symbol."""
    n_c = gd.N_c
    k = np.meshgrid(*[2 * np.pi * np.fft.fftfreq(n) for n in n_c],
                    indexing='ij')
    sym = np.zeros(tuple(n_c))
    for c, o in zip(coefs, np.asarray(offsets)):
        sym = sym + c * np.cos(k[0] * o[0] + k[1] * o[1] + k[2] * o[2])
    rk = np.fft.fftn(rho)
    sym[0, 0, 0] = 1.0
    phik = rk / sym
    phik[0, 0, 0] = 0.0
    return np.fft.ifftn(phik).real
