"""Pin the oracle's Jacobi relaxation and multigrid Poisson solver
(oracle/poisson.py): relaxation against the reference C (Operator.relax of
the reference build) and a direct sweep; the solver against the exact FFT
solution of the same discrete equation."""
import numpy as np
import pytest

from oracle import grid as og
from oracle import poisson as opo
from oracle import refbuild
from oracle.stencil import laplace_stencil


def charge(gd, rng):
    """Neutral pair of Gaussian charges plus a little noise (zero mean)."""
    n_c = gd.N_c
    h = np.diag(gd.h_cv)
    r = np.indices(n_c).transpose(1, 2, 3, 0) * h
    L = np.diag(gd.cell_cv)
    rho = np.zeros(tuple(n_c))
    for q, p in ((1.0, 0.3 * L), (-1.0, 0.65 * L)):
        d = r - p
        d -= np.round(d / L) * L
        rho += q * np.exp(-(d**2).sum(-1) / 1.5**2)
    rho += 0.01 * rng.standard_normal(rho.shape)
    return rho - rho.mean()


@pytest.mark.skipif(not refbuild.available(), reason='reference C not built')
@pytest.mark.parametrize('nn,pbc', [(3, True), (1, True), (2, (True, False,
                                                                 True))])
def test_relax_jacobi_vs_reference(nn, pbc):
    rng = np.random.default_rng(1)
    N_c = np.array([12, 14, 16])
    gd = og.Grid(N_c, N_c * 0.3, pbc)
    c, o = laplace_stencil(gd.h_cv, -0.25 / np.pi, nn)
    f = rng.standard_normal(tuple(gd.n_c))
    s = rng.standard_normal(tuple(gd.n_c))
    op = refbuild.Operator(c, o, gd, True)
    ref = f.copy()
    op.relax(2, ref, s, 3, 2.0 / 3.0)
    out = opo.relax_jacobi(c, o, f.copy(), s, 3, 2.0 / 3.0, gd.pbc_c)
    assert abs(out - ref).max() < 1e-13 * abs(ref).max()


def test_relax_jacobi_single_sweep_formula():
    """One sweep with w = 1 solves each point's equation with the
    neighbours of the old iterate: c0 f' + sum_c coef_c f[shift] = s."""
    rng = np.random.default_rng(2)
    N_c = np.array([8, 8, 8])
    gd = og.Grid(N_c, N_c * 0.4)
    c, o = laplace_stencil(gd.h_cv, 1.0, 2)
    f = rng.standard_normal(tuple(N_c))
    s = rng.standard_normal(tuple(N_c))
    f1 = opo.relax_jacobi(c, o, f.copy(), s, 1, 1.0, gd.pbc_c)
    c0 = c[0]
    off = sum(cc * np.roll(f, tuple(-oo), axis=(0, 1, 2))
              for cc, oo in zip(c[1:], o[1:]))
    assert abs(c0 * f1 + off - s).max() < 1e-12


@pytest.mark.parametrize('N', [32, 24])
def test_poisson_solver_vs_fft(N):
    rng = np.random.default_rng(3)
    gd = og.Grid((N,) * 3, (N * 0.35,) * 3)
    rho = charge(gd, rng)
    ps = opo.FDPoissonSolver(nn=3, eps=1e-20)
    ps.set_grid_descriptor(gd)
    phi = gd.zeros()
    niter = ps.solve_neutral(phi, rho)
    c, o = ps.operators[0]
    exact = opo.fft_solution(gd, c, o, rho)
    assert niter < 200
    assert abs(phi - exact).max() < 1e-8 * abs(exact).max()
    # levels: coarsen while N is even (32 -> ... -> 1, 24 -> ... -> 3)
    assert ps.levels == {32: 5, 24: 3}[N]


@pytest.mark.skipif(not refbuild.available(), reason='reference C not built')
@pytest.mark.parametrize('nn,pbc', [(3, True), (1, True), (2, (True, False,
                                                                 True))])
def test_relax_gauss_seidel_vs_reference(nn, pbc):
    rng = np.random.default_rng(3)
    N_c = np.array([8, 10, 12])
    gd = og.Grid(N_c, N_c * 0.3, pbc)
    c, o = laplace_stencil(gd.h_cv, -0.25 / np.pi, nn)
    f = rng.standard_normal(tuple(gd.n_c))
    s = rng.standard_normal(tuple(gd.n_c))
    op = refbuild.Operator(c, o, gd, True)
    ref = f.copy()
    op.relax(1, ref, s, 2, 1.0)
    out = opo.relax_gauss_seidel(c, o, f.copy(), s, 2, gd.pbc_c)
    assert abs(out - ref).max() < 1e-13 * abs(ref).max()


def test_relax_gauss_seidel_defining_property():
    """After one sweep every point's equation holds with the new values of
    the neighbours before it in C order and the starting values of those
    after it (and of every wrapped neighbour)."""
    rng = np.random.default_rng(4)
    N_c = np.array([6, 7, 8])
    gd = og.Grid(N_c, N_c * 0.4)
    c, o = laplace_stencil(gd.h_cv, 1.0, 2)
    f0 = rng.standard_normal(tuple(N_c))
    s = rng.standard_normal(tuple(N_c))
    f1 = opo.relax_gauss_seidel(c, o, f0.copy(), s, 1, gd.pbc_c)
    worst = 0.0
    for p in np.ndindex(*N_c):
        lhs = c[0] * f1[p]
        for cc, oo in zip(c[1:], o[1:]):
            q = np.array(p) + oo
            wrapped = ((q < 0) | (q >= N_c)).any()
            qm = tuple(q % N_c)
            earlier = not wrapped and tuple(q) < p
            lhs += cc * (f1[qm] if earlier else f0[qm])
        worst = max(worst, abs(lhs - s[p]))
    assert worst < 1e-12
