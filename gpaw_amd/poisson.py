"""Multigrid finite-difference Poisson solver on the GPU (drop-in for
gpaw.poisson.FDPoissonSolver with relax='J').

Solves L phi = rho with L = -1/(4 pi) nabla^2 (phi in Hartree for rho in
electrons / Bohr^3), the reference's construction (gpaw/poisson.py:318-546):
levels by repeated coarsening while the grid is even (at most 8), the nn
stencil on the fine level and 7-point Laplacians below, nn=1 restriction /
interpolation, weighted Jacobi smoothing (FDOperator.relax) with pre/post
sweep counts [2, 4, ...] / [1, 4, ...] (8 / 8 on the coarsest level) and
weights [2/3, 1, ...].  iterate2 is the recursive V-cycle; solve_neutral
repeats it until dv sum r^2 <= eps and removes the mean of phi for fully
periodic cells.  Gauss-Seidel (relax='GS', a sequential sweep) is not
provided.
"""
import numpy as np
import torch

from .fd_operators import Laplace
from .transformers import Transformer


class PoissonConvergenceError(RuntimeError):
    pass


class FDPoissonSolver:
    def __init__(self, nn=3, relax='J', eps=2e-10, maxiter=1000,
                 remove_moment=None, use_charge_center=False, **kwargs):
        if relax != 'J':
            raise NotImplementedError(
                'Relaxation method %s: only Jacobi (J) is provided' % relax)
        if remove_moment:
            raise NotImplementedError('remove_moment')
        self.nn = nn
        self.relax = relax
        self.relax_method = 2
        self.eps = eps
        self.maxiter = maxiter
        self._initialized = False

    def todict(self):
        return {'name': 'fd', 'nn': self.nn, 'relax': self.relax,
                'eps': self.eps}

    def get_stencil(self):
        return self.nn

    def create_laplace(self, gd, scale=1.0, n=1, dtype=float):
        return Laplace(gd, scale, n, dtype)

    def set_grid_descriptor(self, gd):
        self.gd = gd
        scale = -0.25 / np.pi
        self.operators = [self.create_laplace(gd, scale, self.nn)]
        self.interpolators = []
        self.restrictors = []
        self.presmooths = [2]
        self.postsmooths = [1]
        self.weights = [2.0 / 3.0]
        level = 0
        while level < 8:
            try:
                gd2 = gd.coarsen()
            except ValueError:
                break
            self.operators.append(self.create_laplace(gd2, scale, 1))
            self.interpolators.append(Transformer(gd2, gd))
            self.restrictors.append(Transformer(gd, gd2))
            self.presmooths.append(4)
            self.postsmooths.append(4)
            self.weights.append(1.0)
            level += 1
            gd = gd2
        self.levels = level
        self._initialized = False

    def _init(self):
        if self._initialized:
            return
        gds = [op.gd for op in self.operators]
        self.rhos = [g.empty() for g in gds]
        self.phis = [None] + [g.empty() for g in gds[1:]]
        self.residuals = [g.empty() for g in gds]
        self.step = 0.66666666 / self.operators[0].get_diagonal_element()
        self.presmooths[self.levels] = 8
        self.postsmooths[self.levels] = 8
        self._initialized = True

    def solve_neutral(self, phi, rho, timer=None):
        """phi (in/out, initial guess) and rho: device arrays of the grid.
        Returns the number of V-cycles."""
        self._init()
        self.phis[0] = phi
        self.rhos[0].copy_(torch.as_tensor(rho, device='cuda'))
        niter = 1
        while self.iterate2(self.step) > self.eps and niter < self.maxiter:
            niter += 1
        if niter == self.maxiter:
            raise PoissonConvergenceError(
                'Poisson solver did not converge in %d iterations!'
                % self.maxiter)
        if self.gd.pbc_c.all():
            phi -= phi.mean()
        return niter

    def iterate2(self, step, level=0):
        """One V-cycle from `level` down; returns dv sum r^2 at level 0."""
        self._init()
        op = self.operators[level]
        residual = self.residuals[level]
        if level < self.levels:
            op.relax(self.relax_method, self.phis[level], self.rhos[level],
                     self.presmooths[level], self.weights[level])
            # residual = L phi - rho in one stencil pass
            op.apply(self.phis[level][None], residual[None],
                     src_xg=self.rhos[level][None], beta=-1.0)
            self.restrictors[level].apply(residual, self.rhos[level + 1])
            self.phis[level + 1].zero_()
            self.iterate2(4.0 * step, level + 1)
            self.interpolators[level].apply(self.phis[level + 1], residual)
            self.phis[level] -= residual
        op.relax(self.relax_method, self.phis[level], self.rhos[level],
                 self.postsmooths[level], self.weights[level])
        if level == 0:
            op.apply(self.phis[0][None], residual[None],
                     src_xg=self.rhos[0][None], beta=-1.0)
            return float(torch.dot(residual.ravel(), residual.ravel())) * \
                self.gd.dv
