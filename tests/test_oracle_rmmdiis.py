"""Pin the oracle's subspace algebra and RMM-DIIS sweep (oracle/rmmdiis.py)
against what the mathematics fixes: Hermiticity of the dense operator,
diagonal rotated subspace, S-orthonormality in Gram-Schmidt order, the
line-search and DIIS optimality conditions, and convergence of repeated
sweeps to the lowest generalized eigenvalues of the dense (H, S) pair."""
import numpy as np
import pytest
import scipy.linalg as sl

from oracle import grid as og
from oracle import lfc as olfc
from oracle import precond as opc
from oracle import rmmdiis as orm
from oracle import spline as osp
from oracle import stencil as ost


def small_system(N=12, h=0.4, natoms=2, k_c=None, seed=0, dS=True):
    """Two atoms with 8 projectors (l = 0, 0, 1, 1), a Gaussian well per
    atom, random symmetric dH (and dS) blocks."""
    rng = np.random.default_rng(seed)
    gd = og.Grid((N,) * 3, (N * h,) * 3)
    r = np.linspace(0, 2.0, 101)
    spl, raw_spl = [], []
    for a in range(natoms):
        rc = rng.uniform(0.6, 1.0, 4)
        fs = [(l, np.exp(-r**2 / rc[j]**2)) for j, l in enumerate((0, 0, 1, 1))]
        raw_spl.append(fs)
        spl.append([osp.Spline(l, 2.0, f) for l, f in fs])
    spos = rng.uniform(0, 1, (natoms, 3))
    dtype = float if k_c is None else complex
    lfc = olfc.LFC(gd, spl, k_c, dtype)
    lfc.set_positions(spos)
    dH, dS_ = [], []
    for a in range(natoms):
        X = rng.normal(0, 0.1, (8, 8))
        dH.append(X + X.T)
        Y = rng.normal(0, 0.02, (8, 8))
        dS_.append(Y + Y.T)
    L = N * h
    g = np.indices((N,) * 3).transpose(1, 2, 3, 0) * h
    vt = np.zeros((N,) * 3)
    for p in spos @ gd.cell_cv:
        d = g - p
        d -= np.round(d / L) * L
        vt -= 1.5 * np.exp(-(d**2).sum(-1) / 1.5**2)
    c, o = ost.laplace_stencil(gd.h_cv, -0.5, 3)
    ph = None if k_c is None else gd.phase_cd(k_c)
    ham = orm.Hamiltonian(gd, (c, o), vt, lfc, dH, dS_ if dS else None, ph)
    pc = opc.Preconditioner(gd, c, o)
    # raw inputs, for building the same system on the GPU side
    ham.raw = dict(N=N, h=h, k_c=k_c, splines=raw_spl, rcut=2.0, spos=spos,
                   dH=dH, dS=dS_ if dS else None, vt=vt)
    return ham, pc, rng


def dense_pair(ham):
    """Dense H (from the operator applied to unit vectors) and S."""
    gd, lfc = ham.gd, ham.lfc
    dtype = float if ham.phase_cd is None else complex
    Hm = orm.dense_matrix(ham, dtype)
    N = Hm.shape[0]
    Sm = np.eye(N, dtype=dtype)
    if ham.dS_aii is not None:
        B = lfc.dense_matrix()
        for a, dS in enumerate(ham.dS_aii):
            M = lfc.M_a[a]
            Bm = B[M:M + len(dS)]
            Sm = Sm + gd.dv * Bm.T @ dS @ Bm.conj()
    return Hm, Sm


def start_guess(ham, rng, nb, W, noise):
    N = tuple(ham.gd.n_c)
    psi = W.T.reshape((nb,) + N) / ham.gd.dv**0.5
    pert = rng.standard_normal(psi.shape)
    if np.iscomplexobj(W):
        pert = pert + 1j * rng.standard_normal(psi.shape)
    return psi + noise * pert / np.abs(pert).max() * np.abs(psi).max()


@pytest.mark.parametrize('k_c', [None, (0.25, -0.1, 0.3)])
def test_dense_operator_hermitian(k_c):
    ham, _, _ = small_system(N=8, h=0.45, k_c=k_c)
    Hm, Sm = dense_pair(ham)
    assert abs(Hm - Hm.conj().T).max() < 1e-12 * abs(Hm).max()
    assert abs(Sm - Sm.conj().T).max() < 1e-14
    assert np.linalg.eigvalsh(Sm).min() > 0.5


def test_orthonormalize_gram_schmidt_order():
    ham, _, rng = small_system(N=8, h=0.45)
    psi = rng.standard_normal((5, 8, 8, 8))
    psi2, P2 = orm.orthonormalize(ham, psi)
    Hm, Sm = dense_pair(ham)
    X = orm.flat(psi2)
    # <psi'_n|S|psi'_m> = delta_nm with the grid inner product dv a^H b
    G = ham.gd.dv * X.conj() @ Sm @ X.T
    assert abs(G - np.eye(5)).max() < 1e-12
    # Gram-Schmidt: psi'_0 is psi_0 normalized; span{psi'_0..k} unchanged
    n0 = (ham.gd.dv * orm.flat(psi)[0] @ Sm @ orm.flat(psi)[0])**0.5
    assert abs(X[0] - orm.flat(psi)[0] / n0).max() < 1e-12
    Q = np.linalg.lstsq(orm.flat(psi)[:3].T, X[2], rcond=None)
    assert np.linalg.norm(orm.flat(psi)[:3].T @ Q[0] - X[2]) < 1e-10
    # P rotates with psi
    assert abs(P2 - ham.projections(psi2)).max() < 1e-12


def test_subspace_diagonal():
    ham, _, rng = small_system(N=8, h=0.45)
    psi, _ = orm.orthonormalize(ham, rng.standard_normal((4, 8, 8, 8)))
    Hm, Sm = dense_pair(ham)
    A = orm.flat(psi)
    Hsub = ham.gd.dv * A.conj() @ Hm @ A.T
    eps, psi2, Ht2, P2 = orm.subspace_diagonalize(ham, psi)
    assert abs(eps - np.linalg.eigvalsh(0.5 * (Hsub + Hsub.T))).max() < 1e-12
    X = orm.flat(psi2)
    D = ham.gd.dv * X.conj() @ Hm @ X.T
    assert abs(D - np.diag(eps)).max() < 1e-11
    assert abs(orm.flat(Ht2) - orm.flat(ham.apply_local(psi2))).max() < 1e-11
    assert abs(P2 - ham.projections(psi2)).max() < 1e-12


def test_line_search_and_diis_optimal():
    rng = np.random.default_rng(3)
    R = rng.standard_normal((3, 4, 4, 4))
    dR = rng.standard_normal((3, 4, 4, 4))
    lam = orm.line_search(R, dR, 0.7)
    for b in range(3):
        f = lambda t: np.sum((R[b] + t * dR[b])**2)  # noqa: E731
        assert f(lam[b]) <= min(f(lam[b] + 1e-4), f(lam[b] - 1e-4))
    Ri = rng.standard_normal((3, 50)) + 1j * rng.standard_normal((3, 50))
    al = orm.diis_weights(Ri, 0.7)
    assert abs(al.sum() - 1) < 1e-12
    base = np.linalg.norm(al @ Ri)
    for _ in range(20):
        d = rng.standard_normal(3) + 1j * rng.standard_normal(3)
        d -= d.mean()           # stay on sum(alpha) = 1
        assert np.linalg.norm((al + 1e-3 * d) @ Ri) >= base - 1e-12


@pytest.mark.parametrize('k_c', [None, (0.25, -0.1, 0.3)])
def test_rmmdiis_converges_to_dense(k_c):
    ham, pc, rng = small_system(N=12, h=0.4, k_c=k_c)
    Hm, Sm = dense_pair(ham)
    nb = 6
    w, W = sl.eigh(Hm, Sm, subset_by_index=[0, nb - 1])
    psi = start_guess(ham, rng, nb, W, 0.3)
    psi, _ = orm.orthonormalize(ham, psi)
    errors = []
    for it in range(30):
        err, eps, psi = orm.rmmdiis_sweep(ham, psi, pc, blocksize=4)
        errors.append(err)
    assert errors[-1] < 1e-6 * errors[0]
    assert abs(eps[:4] - w[:4]).max() < 1e-10
