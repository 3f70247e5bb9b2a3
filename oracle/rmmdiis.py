"""Subspace algebra and blocked RMM-DIIS -- oracle restatement (test
infrastructure).

NumPy restatement of
  * the PAW pseudo-Hamiltonian on wave functions (hamiltonian.py:23-44,
    437-479; wavefunctions/fd.py:82-88): H psi = T psi + v psi +
    sum_a p_a dH_a <p_a|psi>;
  * subspace diagonalisation (eigensolvers/eigensolver.py:175-236):
    H_nm = <psi_n|T + v|psi_m> + sum_a P_n^* dH_a P_m, eigh, rotation of
    psi, (T + v) psi and P by the eigenvectors;
  * residuals (eigensolver.py:153-172): R = (T + v) psi - eps psi +
    sum_a p_a (P dH_a - eps P dS_a);
  * orthonormalisation (wavefunctions/fdpw.py:394-424, matrix.py:229-262):
    S = <psi|psi> + P dS P^H = L L^H, psi <- L^-1 psi, P <- L^-1 P;
  * one RMM-DIIS sweep (eigensolvers/rmmdiis.py:63-254).

Pinned by tests/test_oracle_rmmdiis.py against what the mathematics fixes:
the dense matrix of H is Hermitian (so add and integrate are adjoint with
the right phases); the rotated subspace is diagonal with the eigenvalues of
the projected matrix; orthonormalisation gives <psi'|S|psi'> = 1 in
Gram-Schmidt order; repeated sweeps converge to the lowest eigenvalues of
the dense matrix (numpy.linalg.eigh) and the residual norm falls; the
line-search lambda is the minimiser of |R + lambda dR|; the DIIS weights sum
to one and minimise the combined residual over that affine set.
"""
import numpy as np

from .stencil import fd_apply


class Hamiltonian:
    def __init__(self, gd, kin, vt_G, lfc=None, dH_aii=None, dS_aii=None,
                 phase_cd=None):
        self.gd = gd
        self.kin = kin                 # (coefs, offsets)
        self.vt_G = vt_G
        self.lfc = lfc
        self.dH_aii = dH_aii
        self.dS_aii = dS_aii
        self.phase_cd = phase_cd

    def apply_local(self, psi):
        """(T + v) psi  (hamiltonian.py:437-479 without the non-local part)."""
        return (fd_apply(self.kin[0], self.kin[1], psi, self.gd.pbc_c,
                         self.phase_cd) + self.vt_G * psi)

    def _blocks(self, X_aii, P):
        out = np.zeros_like(P)
        for a, X in enumerate(X_aii):
            M = self.lfc.M_a[a]
            n = self.lfc.Mmax_a[a]
            out[:, M:M + n] = P[:, M:M + n] @ X
        return out

    def dH(self, P, eps=None):
        """C = P dH (- eps P dS) per atom (hamiltonian.py:23-44,
        eigensolver.py:162-166)."""
        C = self._blocks(self.dH_aii, P)
        if eps is not None and self.dS_aii is not None:
            C -= eps[:, None] * self._blocks(self.dS_aii, P)
        return C

    def projections(self, psi):
        return None if self.lfc is None else self.lfc.integrate(psi)

    def apply(self, psi):
        out = self.apply_local(psi)
        if self.lfc is not None:
            self.lfc.add(out, self.dH(self.projections(psi)))
        return out

    def residual(self, psi, eps):
        """(H - eps S) psi."""
        R = self.apply_local(psi) - eps[:, None, None, None] * psi
        if self.lfc is not None:
            self.lfc.add(R, self.dH(self.projections(psi), eps))
        return R


def flat(a):
    return a.reshape(len(a), -1)


def subspace_diagonalize(ham, psi):
    """Returns eps, rotated psi, rotated (T + v) psi, rotated P."""
    dv = ham.gd.dv
    Ht = ham.apply_local(psi)
    H = dv * flat(psi).conj() @ flat(Ht).T
    P = ham.projections(psi)
    if P is not None:
        H += P.conj() @ ham.dH(P).T
    H = 0.5 * (H + H.conj().T)
    eps, C = np.linalg.eigh(H)
    U = C.T
    shape = psi.shape
    psi = (U @ flat(psi)).reshape(shape)
    Ht = (U @ flat(Ht)).reshape(shape)
    P = None if P is None else U @ P
    return eps, psi, Ht, P


def overlap(ham, psi, P):
    """S' = dv psi psi^H + P dS P^H (complex conjugate of <psi_n|S|psi_m>)."""
    S = ham.gd.dv * flat(psi) @ flat(psi).conj().T
    if P is not None and ham.dS_aii is not None:
        S = S + ham._blocks(ham.dS_aii, P) @ P.conj().T
    return S


def orthonormalize(ham, psi, P=None):
    if P is None:
        P = ham.projections(psi)
    L = np.linalg.cholesky(overlap(ham, psi, P))
    Linv = np.linalg.inv(L)
    psi = (Linv @ flat(psi)).reshape(psi.shape)
    return psi, (None if P is None else Linv @ P)


def line_search(R, dR, dv):
    """lambda = -Re<R|dR> / <dR|dR> per band (rmmdiis.py:157-164)."""
    RdR = dv * np.real(np.sum(flat(dR).conj() * flat(R), axis=1))
    dRdR = dv * np.real(np.sum(flat(dR).conj() * flat(dR), axis=1))
    return -RdR / dRdR


def diis_weights(R_i, dv):
    """alpha minimising |sum_i alpha_i R_i| with sum alpha_i = 1
    (rmmdiis.py:207-224): one band, R_i of shape (k, G)."""
    k = len(R_i)
    A = -np.ones((k + 1, k + 1), R_i.dtype)
    A[:k, :k] = dv * R_i.conj() @ R_i.T
    A[-1, -1] = 0.0
    x = np.zeros(k + 1, R_i.dtype)
    x[-1] = -1.0
    return np.linalg.solve(A, x)[:-1]


def rmmdiis_sweep(ham, psi, precond, blocksize=10, niter=3, trial_step=0.1,
                  weights=None):
    """One Eigensolver.iterate on an orthonormal psi: subspace
    diagonalisation, blocked RMM-DIIS, orthonormalisation.  Returns
    (error, eps, psi)."""
    dv = ham.gd.dv
    nb = len(psi)
    if weights is None:
        weights = np.ones(nb)
    eps, psi, R, P = subspace_diagonalize(ham, psi)
    R = R - eps[:, None, None, None] * psi
    if P is not None:
        ham.lfc.add(R, ham.dH(P, eps))
    error = 0.0
    for n1 in range(0, nb, blocksize):
        n2 = min(n1 + blocksize, nb)
        e = eps[n1:n2]
        psib = psi[n1:n2]
        Rb = R[n1:n2]
        error += np.sum(weights[n1:n2] * dv *
                        np.real(np.sum(flat(Rb).conj() * flat(Rb), 1)))
        psi_hist = [psib.copy()]
        R_hist = [Rb.copy()]
        dpsi = precond(Rb, ham.phase_cd)
        dR = ham.residual(dpsi, e)
        lam = line_search(Rb, dR, dv)
        psib += lam[:, None, None, None] * dpsi
        Rb += lam[:, None, None, None] * dR
        for nit in range(1, niter):
            psi_hist.append(psib.copy())
            R_hist.append(Rb.copy())
            if nit > 1:
                for b in range(n2 - n1):
                    alpha = diis_weights(
                        np.array([flat(r)[b] for r in R_hist]), dv)
                    psib[b] = sum(al * p[b] for al, p in zip(alpha, psi_hist))
                    Rb[b] = sum(al * r[b] for al, r in zip(alpha, R_hist))
            if nit < niter - 1:
                dpsi = precond(Rb, ham.phase_cd)
                psib += lam[:, None, None, None] * dpsi
                Rb[:] = ham.residual(psib, e)
        dpsi = precond(Rb, ham.phase_cd)
        if trial_step is not None:
            lam = np.full(n2 - n1, trial_step)
        psib += lam[:, None, None, None] * dpsi
    psi, P = orthonormalize(ham, psi)
    return error, eps, psi


def dense_matrix(ham, dtype=float):
    """The grid operator H as a dense (N, N) matrix (small grids only)."""
    n_c = tuple(ham.gd.n_c)
    N = int(np.prod(n_c))
    E = np.eye(N, dtype=dtype).reshape((N,) + n_c)
    return flat(ham.apply(E)).T
