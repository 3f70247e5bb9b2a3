"""Blocked RMM-DIIS on the GPU (gpaw/eigensolvers/rmmdiis.py counterpart).

FDWaveFunctions holds one k-point's wave functions psit (nbands, n0, n1, n2)
on the device with eigenvalues eps and projections P = <p|psit>.  Rotations
write a new buffer and re-bind wfs.psit (the reference copies back into
psit); read wfs.psit after an iteration rather than keeping the tensor.
RMMDIIS.iterate restates Eigensolver.iterate + RMMDIIS.iterate_one_k_point
(eigensolver.py:121-147, rmmdiis.py:63-254):

  1. subspace diagonalisation (eigensolver.py:175-236): H_nm =
     dv <psi_n|T+v|psi_m> + sum_a P_n^* dH_a P_m, eigh, rotate psit, Htpsit
     and P by the eigenvectors;
  2. residuals R = (T+v) psi - eps psi + sum_a p_a (P dH_a - eps P dS_a)
     (eigensolver.py:153-172);
  3. per block of B bands: error += w_n int |R_n|^2; dpsi = M R
     (preconditioner); dR = (H - eps S) dpsi; lambda = -<R|dR>/<dR|dR>;
     psi += lambda dpsi, R += lambda dR; niter - 1 DIIS steps (minimise
     |sum_i alpha_i R_i| with sum_i alpha_i = 1 over the stored iterates);
     final trial step psi += 0.1 M R;
  4. orthonormalisation (fdpw.py:394-424): S = dv psi psi^H + P dS P^H =
     L L^H, psi <- L^-1 psi, P <- L^-1 P.

Grid work (stencils, projectors, V-cycle, band-wise dots and axpys) and the
subspace matrices S and H (gh_gram: lower triangles on the matrix cores,
FP64 sums) run in libgpawhip; eigh and Cholesky run in FP64 on the GPU
through torch (rocSOLVER) and the rotations psi <- U psi through torch
(rocBLAS) in the precision of the wave functions.  With float32 /
complex64 wave functions (the mixed-precision form) everything on the grid
is single precision while S, H, eigh, Cholesky and every reduction (errors,
DIIS matrices) stay FP64.
"""
import torch

from .. import blas
from .eigensolver import Eigensolver, _Kpt, band_weights


class FDWaveFunctions:
    """One k-point's wave functions on the device (fdpw.py counterpart)."""

    def __init__(self, gd, psit, pt, preconditioner, phase_cd=None):
        self.gd = gd
        self.psit = psit
        self.pt = pt
        self.preconditioner = preconditioner
        self.phase_cd = phase_cd
        self.dtype = psit.dtype
        self.nbands = psit.shape[0]
        self.dv = gd.dv
        self.eps = torch.zeros(self.nbands, dtype=torch.float64,
                               device='cuda')
        self.P = None
        self.orthonormalized = False

    @staticmethod
    def flat(a):
        return a.reshape(a.shape[0], -1)

    def projections(self):
        if self.P is None and self.pt is not None:
            self.P = self.pt.integrate_M(self.psit)
        return self.P

    def overlap(self, ham):
        """S' = dv psi psi^H + P dS P^H (the complex conjugate of
        S_nm = <psi_n|S|psi_m>; equal to it for real wave functions)."""
        psi = self.flat(self.psit)
        S = blas.gram(psi, alpha=self.dv, hermitian=True)
        P = self.projections()
        if P is not None and ham is not None and ham.dS_flat is not None:
            S = S + (ham.dS(P) @ P.conj().T).to(S.dtype)
            S = 0.5 * (S + S.conj().T)
        return S

    def add_to_density_from_k_point_with_occupation(self, nt_G, f_n):
        """nt_G += sum_n f_n |psit_n|^2 (wavefunctions/fd.py:95-100)."""
        from ..density import add_to_density_bands
        return add_to_density_bands(f_n, self.psit, nt_G)

    def orthonormalize(self, ham=None):
        """psi <- L^-1 psi, P <- L^-1 P with L L^H = S' (fdpw.py:394-424)."""
        S = self.overlap(ham)
        L = torch.linalg.cholesky(S)
        Linv = torch.linalg.solve_triangular(
            L, torch.eye(len(S), dtype=S.dtype, device=S.device), upper=False)
        psi = self.flat(self.psit)
        Linv = Linv.to(psi.dtype)   # the rotation in the data precision
        # the rotated wave functions go to a new buffer (no copy back):
        # psit is re-bound, as is every rotation below
        self.psit = (Linv @ psi).reshape(self.psit.shape)
        if self.P is not None:
            self.P = Linv @ self.P
        self.orthonormalized = True


class RMMDIIS(Eigensolver):
    """RMM-DIIS eigensolver (rmmdiis.py:10-50 parameters).

    blocksize: bands updated together.  Every step of the iteration is per
    band, so the result does not depend on it (the reference's own
    blocked-vs-unblocked test, test_blocked_rmm_diis.py); None (default)
    takes all bands at once, which fills the GPU (the reference's FD default
    is 10, sized for CPU caches)."""

    def __init__(self, keep_htpsit=True, blocksize=None, niter=3, rtol=1e-16,
                 limit_lambda=False, trial_step=0.1):
        if not keep_htpsit:
            raise NotImplementedError('keep_htpsit=False')
        self.keep_htpsit = keep_htpsit
        self.blocksize = blocksize
        self.niter = niter
        self.rtol = rtol
        self.limit_lambda = limit_lambda
        self.trial_step = trial_step
        self.orthonormalization_required = True
        self.error = None

    # ----------------------------------------------------------- pieces --
    # ------------------------------------------------------------ driver --
    def iterate(self, ham, wfs, weights=None):
        """One sweep over all bands (Eigensolver.iterate for one k-point).
        Returns the error sum_n w_n int |R_n|^2 dv (weights w_n default 1)."""
        if not wfs.orthonormalized:
            wfs.orthonormalize(ham)
        error = self.iterate_one_k_point(ham, wfs, weights)
        if self.orthonormalization_required:
            wfs.P = None
            wfs.orthonormalize(ham)
            wfs.projections()
        self.error = error
        return error

    def iterate_one_k_point(self, ham, wfs, weights=None):
        nb = wfs.nbands
        weights = band_weights(weights, nb)
        R = self.subspace_diagonalize(ham, wfs)
        psit = wfs.psit
        self.calculate_residuals(ham, wfs, psit, wfs.P, wfs.eps, R)
        kpt = _Kpt(wfs.phase_cd)
        B = self.blocksize or nb
        errors = []
        for n1 in range(0, nb, B):
            n2 = min(n1 + B, nb)
            psib = psit[n1:n2]
            Rb = R[n1:n2]
            eps = wfs.eps[n1:n2].contiguous()
            errors.append(weights[n1:n2] * self._update_block(
                ham, wfs, kpt, psib, Rb, eps))
        return float(torch.cat(errors).sum())

    def _step(self, x, lam, dx, out):
        """out = x + lam dx per band, one pass (out may be dx)."""
        one = torch.ones_like(lam)
        return blas.multi_lincomb([x, dx], torch.stack([one, lam], 1), out)

    def _update_block(self, ham, wfs, kpt, psib, Rb, eps):
        """The DIIS iterations of one block (rmmdiis.py:124-254); psib is
        overwritten with the result, Rb is used as workspace.  Returns the
        per-band errors dv <R|R> of the residuals on entry.

        The reference copies each iterate into its DIIS history and updates
        the block in place; here each iterate is a buffer of its own that the
        history keeps, and every update x + lam dx or sum_i alpha_i x_i is one
        pass that writes a new iterate (no copies)."""
        # RdR = dv <dR|R>, dRdR = dv <dR|dR> and the error dv <R|R> from one
        # pass over dR and R
        dpsit = self.precondition(wfs, Rb, kpt)
        dR = torch.empty_like(dpsit)
        self.residual_of(ham, wfs, dpsit, eps, dR)
        G = blas.multi_gram([dR, Rb], scale=wfs.dv)
        if G.is_complex():
            G = G.real
        judge = getattr(wfs.preconditioner, 'judge', None)
        if judge is not None:
            # fraction of |R|^2 left after the line search along dpsit: a
            # preconditioner with a fallback may reject its step per band;
            # those bands are redone with the fallback
            rho = 1.0 - G[:, 0, 1] ** 2 / (G[:, 0, 0] * G[:, 1, 1])
            bad = judge(rho)
            if bad is not None and bool(bad.any()):
                idx = bad.nonzero()[:, 0]
                Rbad = Rb[idx].contiguous()
                d_fb = wfs.preconditioner.fallback(Rbad, kpt)
                dR_fb = torch.empty_like(d_fb)
                self.residual_of(ham, wfs, d_fb, eps[idx].contiguous(),
                                 dR_fb)
                dpsit[idx] = d_fb
                dR[idx] = dR_fb
                G = blas.multi_gram([dR, Rb], scale=wfs.dv)
                if G.is_complex():
                    G = G.real
        error = G[:, 1, 1]
        lam = -G[:, 0, 1] / G[:, 0, 0]
        if self.limit_lambda:
            lam = self._limit(lam)
        psi_hist = [psib, self._step(psib, lam, dpsit, dpsit)]
        R_hist = [Rb, self._step(Rb, lam, dR, dR)]
        psi_c, R_c = psi_hist[-1], R_hist[-1]
        for nit in range(1, self.niter):
            if nit > 1 or self.limit_lambda:
                psi_c, R_c = self._diis_combine(wfs, psi_hist, R_hist, nit)
            if nit < self.niter - 1:
                dpsit = self.precondition(wfs, R_c, kpt)
                psi_n = self._step(psi_c, lam, dpsit, dpsit)
                R_n = torch.empty_like(psi_n)
                self.residual_of(ham, wfs, psi_n, eps, R_n)
                psi_hist.append(psi_n)
                R_hist.append(R_n)
                psi_c, R_c = psi_n, R_n
        dpsit = self.precondition(wfs, R_c, kpt)
        if self.trial_step is not None:
            lam = torch.full_like(lam, self.trial_step)
        self._step(psi_c, lam, dpsit, psib)
        return error

    def _limit(self, lam):
        lower = self.limit_lambda['lower']
        upper = self.limit_lambda['upper']
        if self.limit_lambda.get('absolute', False):
            lam = torch.where(lam.abs() < lower, lower * lam.sign(), lam)
            lam = torch.where(lam.abs() > upper, upper * lam.sign(), lam)
        else:
            lam = lam.clamp(lower, upper)
        return lam

    def _diis_combine(self, wfs, psi_hist, R_hist, nit):
        """Per band: alpha minimising |sum_i alpha_i R_i| with
        sum_i alpha_i = 1 over the nit + 1 stored iterates; returns
        (sum_i alpha_i psi_i, sum_i alpha_i R_i) as new arrays, the terms
        added in the reference's order: alpha_nit x_nit first, then
        i = 0 .. nit - 1 (rmmdiis.py:200-239)."""
        k = nit + 1
        # R_nn[b, i, j] = dv <R_i|R_j> (one pass over the k residuals)
        R_nn = blas.multi_gram(R_hist[:k], scale=wfs.dv)
        nbk = R_nn.shape[0]
        A = -torch.ones((nbk, k + 1, k + 1), dtype=R_nn.dtype, device='cuda')
        A[:, :k, :k] = R_nn
        A[:, -1, -1] = 0.0
        x = torch.zeros((nbk, k + 1, 1), dtype=R_nn.dtype, device='cuda')
        x[:, -1, 0] = -1.0
        alpha = torch.linalg.solve(A, x)[:, :k, 0]
        order = [nit] + list(range(nit))
        coef = alpha[:, order].contiguous()
        return tuple(blas.multi_lincomb([hist[i] for i in order], coef)
                     for hist in (psi_hist, R_hist))
