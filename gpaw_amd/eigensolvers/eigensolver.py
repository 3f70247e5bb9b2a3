"""Pieces shared by the eigensolver drivers (gpaw/eigensolvers/
eigensolver.py counterpart): the local Hamiltonian on device tensors,
residuals (eigensolver.py:153-172), subspace diagonalisation
(eigensolver.py:175-236), band-wise integrals and the preconditioner call.
"""
import torch

from .. import blas


class _Kpt:
    def __init__(self, phase_cd):
        self.phase_cd = phase_cd


def band_weights(weights, nb):
    """Per-band error weights on the device (default 1)."""
    if weights is None:
        return torch.ones(nb, dtype=torch.float64, device='cuda')
    return torch.as_tensor(weights, dtype=torch.float64, device='cuda')


class Eigensolver:
    """Base of RMMDIIS and Davidson (one k-point, device tensors)."""

    def apply_h_local(self, ham, wfs, a, out):
        ham.apply_local(a, out, wfs.phase_cd)

    def calculate_residuals(self, ham, wfs, psit, P, eps, R):
        """R <- R - eps psit + sum_a p_a (P dH_a - eps P dS_a), where R holds
        (T + v) psit on entry (eigensolver.py:153-172)."""
        blas.multi_axpy(-eps, psit, R)
        self.add_nonlocal_residual(ham, P, eps, R)

    def add_nonlocal_residual(self, ham, P, eps, R):
        """R += sum_a p_a (P dH_a - eps P dS_a)."""
        if ham.pt is not None:
            ham.pt.add_M(R, ham.dH(P, eps_x=eps))

    def residual_of(self, ham, wfs, x, eps, out, P=None):
        """out = (H - eps S) x: the local part (T + v - eps) x in one
        stencil pass, then the projector terms (with P = <p|x> unless
        given)."""
        ham.apply_local_shifted(x, out, eps, wfs.phase_cd)
        if ham.pt is not None:
            if P is None:
                P = ham.pt.integrate_M(x)
            self.add_nonlocal_residual(ham, P, eps, out)
        return out

    def subspace_diagonalize(self, ham, wfs):
        """Rotate psit, P to the eigenvectors of the subspace Hamiltonian;
        returns (T + v) psit of the rotated wave functions."""
        psit = wfs.psit
        Htpsit = torch.empty_like(psit)
        self.apply_h_local(ham, wfs, psit, Htpsit)
        psi = wfs.flat(psit)
        hpsi = wfs.flat(Htpsit)
        # H_xy = dv <psi_x|T + v|psi_y>: lower triangle on the matrix cores
        # (FP64 sums whatever the data precision), the upper one mirrored
        H = blas.gram(psi, hpsi, alpha=wfs.dv, hermitian=True, conj=True)
        P = wfs.projections()
        if P is not None:
            H = H + (P.conj() @ ham.dH(P).T).to(H.dtype)
            H = 0.5 * (H + H.conj().T)
        eps, C = torch.linalg.eigh(H)
        U = C.T.contiguous().to(psi.dtype)   # rotations in the data precision
        wfs.psit = (U @ psi).reshape(psit.shape)
        Htpsit = (U @ hpsi).reshape(psit.shape)
        if P is not None:
            wfs.P = U @ P
        wfs.eps = eps.to(torch.float64).contiguous()
        return Htpsit

    def integrate(self, wfs, a, b):
        """Re <a_x|b_x> per band."""
        return blas.multi_dotc(a, b, scale=wfs.dv)

    def precondition(self, wfs, R, kpt):
        return wfs.preconditioner(R, kpt)
