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

With graph=True (the default) solve_neutral runs the first V-cycle
eagerly, then records one V-cycle (every launch from the fine-level
smoothing to the final residual, about 60 kernels on a 224^3 grid) in a
HIP graph and replays it; only the residual norm is read back per cycle.
The cycle works on the solver's own phi buffer, so the one recording
serves every later solve on this grid.
"""
import numpy as np
import torch

from .fd_operators import Laplace
from .transformers import Transformer


class PoissonConvergenceError(RuntimeError):
    pass


class FDPoissonSolver:
    def __init__(self, nn=3, relax='J', eps=2e-10, maxiter=1000,
                 remove_moment=None, use_charge_center=False, graph=True,
                 **kwargs):
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
        self.use_graph = graph
        self._graph = None
        self.errors = []
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
        self._graph = None
        self._initialized = False

    def _init(self):
        if self._initialized:
            return
        gds = [op.gd for op in self.operators]
        self.rhos = [g.empty() for g in gds]
        self.phis = [None] + [g.empty() for g in gds[1:]]
        self.residuals = [g.empty() for g in gds]
        self.phi0 = gds[0].empty()
        self.step = 0.66666666 / self.operators[0].get_diagonal_element()
        self.presmooths[self.levels] = 8
        self.postsmooths[self.levels] = 8
        self._initialized = True

    def solve_neutral(self, phi, rho, timer=None):
        """phi (in/out, initial guess) and rho: device arrays of the grid.
        Returns the number of V-cycles; self.errors holds dv sum r^2 after
        each."""
        self._init()
        self.rhos[0].copy_(torch.as_tensor(rho, device='cuda'))
        self.errors = []
        if not self.use_graph:
            self.phis[0] = phi
            niter = 1
            while self._error(self.iterate2(self.step)) > self.eps and \
                    niter < self.maxiter:
                niter += 1
        else:
            self.phis[0] = self.phi0
            self.phi0.copy_(torch.as_tensor(phi, device='cuda'))
            niter = 1
            err = self._error(self.iterate2(self.step))
            if err > self.eps and niter < self.maxiter:
                graph = self._recorded_cycle()
                while err > self.eps and niter < self.maxiter:
                    graph.replay()
                    niter += 1
                    err = self._error(self._residual_norm())
            phi.copy_(self.phi0)
        if niter == self.maxiter:
            raise PoissonConvergenceError(
                'Poisson solver did not converge in %d iterations!'
                % self.maxiter)
        if self.gd.pbc_c.all():
            phi -= phi.mean()
        return niter

    def _error(self, err):
        self.errors.append(err)
        return err

    def _residual_norm(self):
        r = self.residuals[0]
        return float(torch.dot(r.ravel(), r.ravel())) * self.gd.dv

    def _recorded_cycle(self):
        """One V-cycle (through the level-0 residual) recorded in a HIP
        graph; the eager cycle before it has initialised every kernel."""
        if self._graph is None:
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g):
                self._cycle(self.step)
            self._graph = g
        return self._graph

    def iterate2(self, step, level=0):
        """One V-cycle from `level` down; returns dv sum r^2 at level 0."""
        self._cycle(step, level)
        if level == 0:
            return self._residual_norm()

    def _cycle(self, step, level=0):
        """The launches of one V-cycle from `level` down (at level 0
        ending with the residual L phi - rho)."""
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
            self._cycle(4.0 * step, level + 1)
            self.interpolators[level].apply(self.phis[level + 1], residual)
            self.phis[level] -= residual
        op.relax(self.relax_method, self.phis[level], self.rhos[level],
                 self.postsmooths[level], self.weights[level])
        if level == 0:
            op.apply(self.phis[0][None], residual[None],
                     src_xg=self.rhos[0][None], beta=-1.0)
