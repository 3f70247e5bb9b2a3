"""Davidson eigensolver -- oracle restatement (test infrastructure).

NumPy/SciPy restatement of one Eigensolver.iterate with the reference's
Davidson (eigensolvers/davidson.py:80-222, diagonalizerbackend.py:10-42) for
one k-point, all bands in one block:

  subspace diagonalisation; R = (H - eps S) psi;
  niter times: psi2 = precond(R), P2 = <p|psi2>;
    the 2B x 2B pair in the basis (psi, psi2), lower triangles only
      H = [[diag(eps), .], [<psi2|H|psi>, <psi2|H|psi2>]]
      S = [[1,         .], [<psi2|S|psi>, <psi2|S|psi2>]]
    (non-local parts P2^* dH P^T and P2^* dS P^T; <psi2|H|psi> as
    <H psi2|psi>, H Hermitian), scipy.linalg.eigh(H, S, lower=True);
    psi <- C[:B, :B]^T psi + C[B:, :B]^T psi2 (same for P), eps <- e[:B];
    R recomputed except after the last expansion.
  The error is sum_n w_n dv <R_n|R_n> of the residuals entering the last
  expansion (davidson.py:126-127).  No orthonormalisation afterwards: the
  eigenvectors of the pair are S-orthonormal.

Pinned by tests/test_oracle_davidson.py against what the mathematics fixes:
after an iterate psi is S-orthonormal and diagonalises H (eps are its
Rayleigh quotients); each eps can only fall relative to the subspace
eigenvalues of the start (Rayleigh-Ritz in a larger space, Cauchy
interlacing); repeated iterates converge to the lowest generalized
eigenvalues of the dense (H, S) pair, real and with Bloch phases.
"""
import numpy as np
import scipy.linalg as sl

from .rmmdiis import flat, subspace_diagonalize


def _nonlocal(ham, X_aii, Pa, Pb):
    """sum_a Pa^* X_a Pb^T (rows: bands of Pa, columns: bands of Pb)."""
    if Pa is None or X_aii is None:
        return 0.0
    return Pa.conj() @ ham._blocks(X_aii, Pb).T


def pair_matrices(ham, psi, P, eps, psi2, P2):
    """Lower triangles of the 2B x 2B (H, S) pair in the basis (psi, psi2)."""
    dv = ham.gd.dv
    nb = len(psi)
    A, A2 = flat(psi), flat(psi2)
    Hp2 = flat(ham.apply_local(psi2))
    dtype = np.result_type(A.dtype, A2.dtype)
    H = np.zeros((2 * nb, 2 * nb), dtype)
    S = np.zeros((2 * nb, 2 * nb), dtype)
    H[:nb, :nb] = np.diag(eps)
    S[:nb, :nb] = np.eye(nb)
    H[nb:, nb:] = dv * A2.conj() @ Hp2.T + _nonlocal(ham, ham.dH_aii, P2, P2)
    H[nb:, :nb] = dv * Hp2.conj() @ A.T + _nonlocal(ham, ham.dH_aii, P2, P)
    S[nb:, nb:] = dv * A2.conj() @ A2.T + _nonlocal(ham, ham.dS_aii, P2, P2)
    S[nb:, :nb] = dv * A2.conj() @ A.T + _nonlocal(ham, ham.dS_aii, P2, P)
    return H, S


def davidson_iterate(ham, psi, precond, niter=2, weights=None):
    """One Eigensolver.iterate (Davidson) on an S-orthonormal psi.
    Returns (error, eps, psi)."""
    dv = ham.gd.dv
    nb = len(psi)
    if weights is None:
        weights = np.ones(nb)
    eps, psi, R, P = subspace_diagonalize(ham, psi)
    R = R - eps[:, None, None, None] * psi
    if P is not None:
        ham.lfc.add(R, ham.dH(P, eps))
    error = None
    for nit in range(niter):
        if nit == niter - 1:
            error = float(np.dot(weights, dv * np.real(
                np.sum(flat(R).conj() * flat(R), axis=1))))
        psi2 = precond(R, ham.phase_cd)
        P2 = ham.projections(psi2)
        H, S = pair_matrices(ham, psi, P, eps, psi2, P2)
        e, C = sl.eigh(H, S, lower=True)
        eps = e[:nb]
        C1, C2 = C[:nb, :nb].T, C[nb:, :nb].T
        psi = (C1 @ flat(psi) + C2 @ flat(psi2)).reshape(psi.shape)
        if P is not None:
            P = C1 @ P + C2 @ P2
        if nit < niter - 1:
            R = ham.apply_local(psi) - eps[:, None, None, None] * psi
            if P is not None:
                ham.lfc.add(R, ham.dH(P, eps))
    return error, eps, psi
