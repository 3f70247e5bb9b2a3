"""Pin the oracle's Davidson iterate (oracle/davidson.py) against what the
mathematics fixes: S-orthonormal, H-diagonal output whose eps are the
Rayleigh quotients; eigenvalues that can only fall relative to the start's
subspace eigenvalues (Rayleigh-Ritz over a larger space); convergence to the
lowest generalized eigenvalues of the dense (H, S) pair."""
import numpy as np
import pytest
import scipy.linalg as sl

from oracle import davidson as odv
from oracle import rmmdiis as orm
from test_oracle_rmmdiis import dense_pair, small_system, start_guess

CASES = [None, (0.25, -0.1, 0.3)]


@pytest.mark.parametrize('k_c', CASES)
def test_davidson_output_orthonormal_and_diagonal(k_c):
    ham, pc, rng = small_system(N=8, h=0.45, k_c=k_c, seed=6)
    dtype = float if k_c is None else complex
    Hm, Sm = dense_pair(ham)
    psi, _ = orm.orthonormalize(
        ham, rng.standard_normal((5, 8, 8, 8)).astype(dtype))
    err, eps, psi2 = odv.davidson_iterate(ham, psi, pc)
    X = orm.flat(psi2)
    dv = ham.gd.dv
    G = dv * X.conj() @ Sm @ X.T
    D = dv * X.conj() @ Hm @ X.T
    assert abs(G - np.eye(5)).max() < 1e-10
    assert abs(D - np.diag(eps)).max() < 1e-10 * abs(eps).max()
    assert err > 0


@pytest.mark.parametrize('k_c', CASES)
def test_davidson_eigenvalues_fall(k_c):
    """Rayleigh-Ritz over span(psi, psi2) contains span(psi): every Ritz
    value is at most the matching subspace eigenvalue of the start."""
    ham, pc, rng = small_system(N=8, h=0.45, k_c=k_c, seed=7)
    dtype = float if k_c is None else complex
    psi, _ = orm.orthonormalize(
        ham, rng.standard_normal((5, 8, 8, 8)).astype(dtype))
    eps0 = orm.subspace_diagonalize(ham, psi)[0]
    _, eps1, _ = odv.davidson_iterate(ham, psi, pc, niter=1)
    assert (eps1 <= eps0 + 1e-12).all()
    assert (eps1 < eps0 - 1e-6).any()


@pytest.mark.parametrize('k_c', CASES)
def test_davidson_converges_to_dense(k_c):
    """Seven bands, the lowest five converged (weights 1, 1, 1, 1, 1, 0, 0):
    the reference's Davidson keeps no history, so unconverged buffer bands
    above the target set are what make it converge quickly."""
    ham, pc, rng = small_system(N=8, h=0.45, k_c=k_c, seed=8)
    Hm, Sm = dense_pair(ham)
    nb, nconv = 7, 5
    w, W = sl.eigh(Hm, Sm, subset_by_index=[0, nb - 1])
    psi, _ = orm.orthonormalize(ham, start_guess(ham, rng, nb, W, 0.3))
    weights = np.array([1.0] * nconv + [0.0] * (nb - nconv))
    errs = []
    for _ in range(24):
        err, eps, psi = odv.davidson_iterate(ham, psi, pc, weights=weights)
        errs.append(err)
    assert errs[-1] < 1e-8 * errs[0]
    assert abs(eps[:nconv] - w[:nconv]).max() < 1e-9
