"""Davidson on the GPU (gpaw/eigensolvers/davidson.py counterpart).

Davidson.iterate restates Eigensolver.iterate + Davidson.iterate_one_k_point
(eigensolver.py:121-147, davidson.py:80-222) for one k-point, all bands in
one block:

  1. subspace diagonalisation and residuals R = (H - eps S) psi, as in
     RMM-DIIS;
  2. niter times: psi2 = M R (the V-cycle), P2 = <p|psi2>; the 2B x 2B pair
     in the basis (psi, psi2) -- lower triangles, as the reference hands
     them to scipy.linalg.eigh(lower=True):
        H = [[diag(eps), .], [<psi2|H|psi>, <psi2|H|psi2>]]
        S = [[1,         .], [<psi2|S|psi>, <psi2|S|psi2>]]
     with <psi2|H|psi> = <H psi2|psi> (H Hermitian) and the non-local parts
     P2^* dH P^T, P2^* dS P^T; the generalized problem solved by Cholesky
     reduction on the GPU (FP64); psi <- C[:B,:B]^T psi + C[B:,:B]^T psi2,
     the same for P, eps <- e[:B]; R recomputed except after the last
     expansion;
  3. error = sum_n w_n dv <R_n|R_n> of the residuals entering the last
     expansion (davidson.py:126-127).  No orthonormalisation afterwards: the
     eigenvectors of the pair are S-orthonormal.

The four 256 x 1.4 M products per expansion are gh_gram calls (matrix
cores, FP64 sums; the diagonal blocks from lower triangles) and the rotation
is a GEMM through torch (rocBLAS) in the data precision; grid work
(stencil, projectors, V-cycle) runs in libgpawhip.
"""
import torch

from .. import blas
from .eigensolver import Eigensolver, _Kpt, band_weights


def _hermitian_from_lower(A):
    """The Hermitian matrix whose lower triangle is that of A."""
    L = torch.tril(A, -1)
    D = torch.diagonal(A).real
    return L + L.conj().T + torch.diag(D).to(A.dtype)


def eigh_pair_lower(H, S):
    """Generalized eigenproblem H x = e S x from the lower triangles of
    Hermitian H and positive definite S (the reduction LAPACK's
    sygv/hegv performs: S = L L^H, eigh(L^-1 H L^-H), x = L^-H y).
    Returns (e ascending, X with the eigenvectors as columns)."""
    H = _hermitian_from_lower(H)
    S = _hermitian_from_lower(S)
    L = torch.linalg.cholesky(S)
    eye = torch.eye(len(S), dtype=S.dtype, device=S.device)
    Linv = torch.linalg.solve_triangular(L, eye, upper=False)
    A = Linv @ H @ Linv.conj().T
    A = 0.5 * (A + A.conj().T)
    e, Y = torch.linalg.eigh(A)
    return e, Linv.conj().T @ Y


class Davidson(Eigensolver):
    """Davidson eigensolver (davidson.py:18-50 parameters): niter expansions
    of the subspace by the preconditioned residuals per iterate."""

    def __init__(self, niter=2, keep_htpsit=True):
        if not keep_htpsit:
            raise NotImplementedError('keep_htpsit=False')
        self.niter = niter
        self.keep_htpsit = keep_htpsit
        self.orthonormalization_required = False
        self.error = None

    def __repr__(self):
        return 'Davidson(niter=%d)' % self.niter

    def todict(self):
        return {'name': 'dav', 'niter': self.niter}

    def iterate(self, ham, wfs, weights=None):
        """One iterate for one k-point (Eigensolver.iterate).  Returns the
        error sum_n w_n int |R_n|^2 dv (weights w_n default 1)."""
        if not wfs.orthonormalized:
            wfs.orthonormalize(ham)
        error = self.iterate_one_k_point(ham, wfs, weights)
        wfs.orthonormalized = True
        self.error = error
        return error

    def _pair(self, ham, wfs, psit, P, psit2, P2, Hpsit2):
        """Lower blocks of the 2B x 2B (H, S) pair."""
        dv = wfs.dv
        nb = psit.shape[0]
        psi, psi2, h2 = wfs.flat(psit), wfs.flat(psit2), wfs.flat(Hpsit2)
        mdt = torch.promote_types(psi.dtype, torch.float64)
        H = torch.zeros((2 * nb, 2 * nb), dtype=mdt, device='cuda')
        S = torch.zeros_like(H)
        H[:nb, :nb] = torch.diag(wfs.eps).to(mdt)
        S[:nb, :nb] = torch.eye(nb, dtype=mdt, device='cuda')
        # X_xy = dv sum conj(a_x) b_y on the matrix cores (FP64 sums); the
        # diagonal blocks from their lower triangles
        H22 = blas.gram(psi2, h2, alpha=dv, hermitian=True, conj=True)
        H21 = blas.gram(h2, psi, alpha=dv, conj=True)
        S22 = blas.gram(psi2, alpha=dv, hermitian=True, conj=True)
        S21 = blas.gram(psi2, psi, alpha=dv, conj=True)
        if P2 is not None:
            H22 = H22 + (P2.conj() @ ham.dH(P2).T).to(mdt)
            H21 = H21 + (P2.conj() @ ham.dH(P).T).to(mdt)
            if ham.dS_flat is not None:
                S22 = S22 + (P2.conj() @ ham.dS(P2).T).to(mdt)
                S21 = S21 + (P2.conj() @ ham.dS(P).T).to(mdt)
        H[nb:, nb:] = H22
        H[nb:, :nb] = H21
        S[nb:, nb:] = S22
        S[nb:, :nb] = S21
        return H, S

    def iterate_one_k_point(self, ham, wfs, weights=None):
        nb = wfs.nbands
        weights = band_weights(weights, nb)
        R = self.subspace_diagonalize(ham, wfs)
        self.calculate_residuals(ham, wfs, wfs.psit, wfs.P, wfs.eps, R)
        kpt = _Kpt(wfs.phase_cd)
        error = None
        for nit in range(self.niter):
            psit, P = wfs.psit, wfs.P
            if nit == self.niter - 1:
                error = float((weights * blas.multi_dotc(
                    R, R, scale=wfs.dv)).sum())
            psit2 = self.precondition(wfs, R, kpt)
            P2 = ham.pt.integrate_M(psit2) if ham.pt is not None else None
            # R is free now: it takes (T + v) psit2
            self.apply_h_local(ham, wfs, psit2, R)
            H, S = self._pair(ham, wfs, psit, P, psit2, P2, R)
            e, C = eigh_pair_lower(H, S)
            C1 = C[:nb, :nb].T.contiguous().to(psit.dtype)
            C2 = C[nb:, :nb].T.contiguous().to(psit.dtype)
            new = C1 @ wfs.flat(psit)
            new.addmm_(C2, wfs.flat(psit2))
            wfs.psit = new.reshape(psit.shape)
            if P is not None:
                wfs.P = torch.addmm(C1 @ P, C2, P2)
            wfs.eps = e[:nb].to(torch.float64).contiguous()
            del psit2, new
            if nit < self.niter - 1:
                self.residual_of(ham, wfs, wfs.psit, wfs.eps, R, P=wfs.P)
        return error
