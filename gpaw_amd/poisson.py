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
periodic cells; solve handles a net charge and removed multipoles
(uniform background, Gaussian monopole, remove_moment).  relax='GS'
smooths with Gauss-Seidel sweeps in the reference's C order
(FDOperator.relax method 1, one launch per hyperplane of the grid).

With graph=True (the default) solve_neutral runs the first V-cycle
eagerly, then records one V-cycle (every launch from the fine-level
smoothing to the final residual, about 60 kernels on a 224^3 grid) in a
HIP graph and replays it; only the residual norm is read back per cycle.
The cycle works on the solver's own phi buffer, so the one recording
serves every later solve on this grid.
"""
from math import factorial, pi, sqrt

import numpy as np
import torch
from scipy.special import erf

from .fd_operators import Laplace
from .transformers import Transformer


# Real solid harmonics r^l Y_lm for L = l^2 + m (l <= 2), in the m order of
# the reference (c/bmgs/spherical_harmonics.h): a published constant table.
def solid_harmonic(L, x, y, z, r2):
    if L == 0:
        return 0.28209479177387814 + 0.0 * r2
    if 1 <= L <= 3:
        return 0.48860251190291992 * (y, z, x)[L - 1]
    return {4: 1.0925484305920792 * x * y,
            5: 1.0925484305920792 * y * z,
            6: 0.31539156525252005 * (3 * z * z - r2),
            7: 1.0925484305920792 * x * z,
            8: 0.54627421529603959 * (x * x - y * y)}[L]


def _lmax_check(L):
    if not 0 <= L <= 8:
        raise ValueError('generalized Gaussians for l <= 2 (L < 9) only')
    return int(sqrt(L))


def gaussian_density(a, L, x, y, z, r2):
    """Generalized Gaussian g_L = k_l exp(-a r^2) r^l Y_L with
    k_l = (4 pi)^-1/2 l! / (2l+1)! (4a)^(l+3/2), normalised so that
    int g_L r^l Y_L d^3r = 1 (gpaw/utilities/gauss.py:95-159)."""
    l = _lmax_check(L)
    k = (4 * pi) ** -0.5 * factorial(l) / factorial(2 * l + 1) * \
        (4 * a) ** (l + 1.5)
    return k * np.exp(-a * r2) * solid_harmonic(L, x, y, z, r2)


def gaussian_potential(a, L, x, y, z, r2):
    """Potential of gaussian_density(a, L) (nabla^2 v = -4 pi g): for a
    density f(r) r^l Y_L, v = 4 pi / (2l+1) Y_L [r^-(l+1) int_0^r f s^(2l+2)
    ds + r^l int_r^inf f s ds].  With f = k exp(-a s^2) the integrals are
    I_n(r) = int_0^r s^n exp(-a s^2) ds from I_0 = sqrt(pi / a) erf(sqrt(a) r)
    / 2 and I_n = ((n - 1) I_(n-2) - r^(n-1) exp(-a r^2)) / (2a), and
    exp(-a r^2) / (2a)."""
    l = _lmax_check(L)
    k = (4 * pi) ** -0.5 * factorial(l) / factorial(2 * l + 1) * \
        (4 * a) ** (l + 1.5)
    r = np.sqrt(r2)
    e = np.exp(-a * r2)
    I = sqrt(pi / a) * erf(sqrt(a) * r) / 2
    for n in range(2, 2 * l + 3, 2):
        I = ((n - 1) * I - r ** (n - 1) * e) / (2 * a)
    radial = I / r ** (2 * l + 1) + e / (2 * a)
    return 4 * pi / (2 * l + 1) * k * radial * solid_harmonic(L, x, y, z, r2)


class Gaussian:
    """Generalized Gaussians centred in a non-periodic cell (the cell's
    centre or `center`, in Bohr) with the reference's width: exponent
    a = 4 * 19 * max_c |icell_c|^2 at the centre, 19 / r_min^2 at a given
    centre (r_min: its distance to the nearest face); r^2 is kept >= 1e-12
    at the centre point (gpaw/utilities/gauss.py:108-123,
    gpaw/utilities/tools.py:96-115)."""

    def __init__(self, gd, a=19.0, center=None):
        self.gd = gd
        cell_center = gd.cell_cv.sum(0) / 2
        if center is None:
            self.a = 4 * a * (gd.icell_cv ** 2).sum(1).max()
            r0 = cell_center
        else:
            r0 = np.asarray(center, float)
            r_min = (cell_center - np.abs(r0 - cell_center)).min()
            self.a = a / r_min ** 2
        r_vG = gd.get_grid_point_coordinates()          # (3, n0, n1, n2)
        self.xyz = [r_vG[c] - r0[c] for c in range(3)]
        r2 = sum(x * x for x in self.xyz)
        self.r2 = np.where(r2 < 1e-12, 1e-12, r2)

    def get_gauss(self, L):
        return gaussian_density(self.a, L, *self.xyz, self.r2)

    def get_gauss_pot(self, L):
        return gaussian_potential(self.a, L, *self.xyz, self.r2)

    def get_moment(self, n, L):
        """dv sum n r^l Y_L (n: a host array of the grid)."""
        return float((n * solid_harmonic(L, *self.xyz, self.r2)).sum()) * \
            self.gd.dv

    def remove_moment(self, n, L, q=None):
        """n -= q g_L (q: the L moment of n); returns q v_L, or 0 when the
        moment is below 1e-7."""
        if q is None:
            q = self.get_moment(n, L)
        if abs(q) < 1e-7:
            return 0.0
        n -= q * self.get_gauss(L)
        return q * self.get_gauss_pot(L)


class PoissonConvergenceError(RuntimeError):
    pass


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, float)).cuda()


class FDPoissonSolver:
    def __init__(self, nn=3, relax='J', eps=2e-10, maxiter=1000,
                 remove_moment=None, use_charge_center=False,
                 metallic_electrodes=False,
                 use_charged_periodic_corrections=False, graph=True,
                 **kwargs):
        if relax not in ('J', 'GS'):
            raise ValueError('Relaxation method %s: J (Jacobi) or GS '
                             '(Gauss-Seidel)' % relax)
        if metallic_electrodes:
            raise NotImplementedError('metallic_electrodes')
        if use_charged_periodic_corrections:
            raise NotImplementedError(
                'use_charged_periodic_corrections (Madelung term)')
        self.remove_moment = remove_moment
        self.use_charge_center = use_charge_center
        self._gauss = None
        self.nn = nn
        self.relax = relax
        self.relax_method = 1 if relax == 'GS' else 2
        self.eps = eps
        self.maxiter = maxiter
        self.use_graph = graph
        self._graph = None
        self.errors = []
        self._initialized = False

    def todict(self):
        d = {'name': 'fd', 'nn': self.nn, 'relax': self.relax,
             'eps': self.eps}
        if self.remove_moment:
            d['remove_moment'] = self.remove_moment
        if self.use_charge_center:
            d['use_charge_center'] = self.use_charge_center
        return d

    # ---------------------------------------------------------- solve --
    def solve(self, phi, rho, charge=None, maxcharge=1e-6,
              zero_initial_phi=False, timer=None):
        """phi (in/out) for rho of any total charge, as
        BasePoissonSolver.solve (gpaw/poisson.py:178-303): with
        remove_moment (non-periodic cells) the multipoles L < remove_moment
        are carried by Gaussians whose potentials are added back; a neutral
        system is solved with the uniform background of its (tiny) charge
        removed; a charged periodic one with its uniform background removed;
        a charged non-periodic one with its monopole carried by a Gaussian
        at the cell centre (or at the centre of the majority charge,
        use_charge_center).  Mixed boundary conditions with a net charge
        raise NotImplementedError.  Returns the number of V-cycles."""
        self._init()
        gd = self.gd
        rho = torch.as_tensor(rho, device='cuda')
        assert tuple(phi.shape) == tuple(gd.n_c) == tuple(rho.shape)
        actual_charge = float(rho.sum()) * gd.dv
        background = actual_charge / gd.dv / float(np.prod(gd.n_c))
        if self.remove_moment:
            assert not gd.pbc_c.any()
            if self._gauss is None or self._gauss[0] is not None:
                self._gauss = (None, Gaussian(gd))
            gauss = self._gauss[1]
            rho_neutral = rho.clone()
            phi_cor = torch.zeros_like(phi)
            for L in range(self.remove_moment):
                Y = _dev(solid_harmonic(L, *gauss.xyz, gauss.r2))
                q = float((rho_neutral * Y).sum()) * gd.dv
                if abs(q) < 1e-7:
                    continue
                rho_neutral -= q * _dev(gauss.get_gauss(L))
                phi_cor += q * _dev(gauss.get_gauss_pot(L))
            phi -= phi_cor
            niter = self.solve_neutral(phi, rho_neutral, timer=timer)
            phi += phi_cor
            return niter
        if charge is None:
            charge = actual_charge
        if abs(charge) <= maxcharge:
            return self.solve_neutral(phi, rho - background, timer=timer)
        if gd.pbc_c.all():
            if zero_initial_phi:
                phi.zero_()
            return self.solve_neutral(phi, rho - background, timer=timer)
        if gd.pbc_c.any():
            raise NotImplementedError('Charged systems with mixed periodic/'
                                      'zero boundary conditions')
        center = None
        if self.use_charge_center:
            sign = 1.0 if actual_charge > 0 else -1.0
            rho_sign = torch.clamp(rho * sign, min=0.0).cpu().numpy()
            r_vG = gd.get_grid_point_coordinates()
            absolute = rho_sign.sum() * gd.dv
            center = np.array([(r_vG[c] * rho_sign).sum() * gd.dv / absolute
                               for c in range(3)])
            border = np.inner(gd.h_cv, np.array((7, 7, 7)))
            far = np.inner(gd.h_cv, gd.N_c) - border
            if np.any(center > far) or np.any(center < border):
                raise RuntimeError('Poisson solver: center of charge outside '
                                   'borders - please increase box')
        key = None if center is None else tuple(center)
        if self._gauss is None or self._gauss[0] != ('mono', key):
            g = Gaussian(gd, center=center)
            self._gauss = (('mono', key), g, _dev(g.get_gauss(0)),
                           _dev(g.get_gauss_pot(0)))
        rho_gauss, phi_gauss = self._gauss[2], self._gauss[3]
        q = actual_charge / sqrt(4 * pi)       # monopole moment
        rho_neutral = rho - q * rho_gauss
        if zero_initial_phi:
            phi.zero_()
        else:
            phi -= q * phi_gauss
        niter = self.solve_neutral(phi, rho_neutral, timer=timer)
        phi += q * phi_gauss
        return niter

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
