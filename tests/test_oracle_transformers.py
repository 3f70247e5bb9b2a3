"""Pins the oracle transformers (oracle/transformers.py) -- CPU only.

Checks: the reference C build (cgpaw.Transformer), integral conservation
(test_transformations.py:10-51, 1e-14), and an independent brute-force
construction of the separable filters as dense 1-D Bloch-periodic matrices.
"""
import numpy as np
import pytest

from oracle.grid import Grid
from oracle.transformers import WEIGHTS, Transformer


def test_weights_closed_form():
    """Lagrange interpolation weights at the midpoint: sum = 1/2 and exact
    for polynomials of degree < K (coefs() in transformers.py:141-154)."""
    for K, w in WEIGHTS.items():
        nn = K // 2
        x = np.arange(-nn + 1, nn + 1)  # nodes around the midpoint 0.5
        for p in range(K):
            val = sum(wt * ((0 - t)**p + (1 + t)**p) for t, wt in
                      enumerate(w))
            assert abs(val - 0.5**p) < 1e-15
        assert len(x) == K


@pytest.mark.parametrize('nn', [1, 2, 3, 4])
@pytest.mark.parametrize('real', [True, False])
def test_integral_conservation(nn, real):
    rng = np.random.RandomState(8)
    gd = Grid((20, 20, 20), (1, 1, 1))
    a = rng.random_sample((20, 20, 20))
    ph = None
    if not real:
        a = a + 1j * rng.random_sample(a.shape)
        ph = np.ones((3, 2), complex)
    for gd2 in (gd.refine(), gd.coarsen()):
        b = Transformer(gd, gd2, nn).apply(a, ph)
        assert abs(gd.integrate(a) - gd2.integrate(b)) < 1e-14


def _dense_1d(n, nn, phase, restrict):
    """Dense 1-D Bloch-periodic filter matrix built from the weights."""
    K = 2 * nn
    w = WEIGHTS[K]

    def wrap(i):
        f = 1.0
        if i < 0:
            i += n
            f = np.conj(phase)
        elif i >= n:
            i -= n
            f = phase
        return i, f
    if restrict:
        T = np.zeros((n // 2, n), complex)
        for o in range(n // 2):
            taps = [(0, 1.0)] + [(s * (2 * j + 1), wj) for j, wj in
                                 enumerate(w) for s in (1, -1)]
            for off, wt in taps:
                i, f = wrap(2 * o + off)
                T[o, i] += 0.5 * wt * f
    else:
        T = np.zeros((2 * n, n), complex)
        for i in range(n):
            T[2 * i, i] += 1.0
            for t, wt in enumerate(w):
                for idx in (i - t, i + 1 + t):
                    j, f = wrap(idx)
                    T[2 * i + 1, j] += wt * f
    return T


@pytest.mark.parametrize('nn', [1, 3])
@pytest.mark.parametrize('restrict', [True, False])
def test_against_dense_filters(rng, nn, restrict):
    N_c = (8, 12, 16) if restrict else (4, 6, 8)
    gd = Grid(N_c, np.array(N_c) * 0.3)
    k_c = (0.25, -0.1, 0.3)
    ph = gd.phase_cd(k_c)
    a = rng.standard_normal(N_c) + 1j * rng.standard_normal(N_c)
    gd2 = gd.coarsen() if restrict else gd.refine()
    b = Transformer(gd, gd2, nn).apply(a, ph)
    T = [_dense_1d(N_c[c], nn, ph[c, 0], restrict) for c in range(3)]
    b2 = np.einsum('ai,bj,ck,ijk->abc', T[0], T[1], T[2], a)
    assert abs(b - b2).max() < 1e-13


@pytest.mark.parametrize('nn', [1, 2, 3, 4])
@pytest.mark.parametrize('real', [True, False])
@pytest.mark.parametrize('pbc', [True, (True, False, True), False])
def test_against_reference_build(ref, rng, nn, real, pbc):
    gd0 = Grid((16, 20, 24), np.array([16, 20, 24]) * 0.3, pbc)
    gd1 = gd0.coarsen()
    ph = None if real else gd0.phase_cd((0.25, -0.1, 0.3))
    phr = np.ones((3, 2), complex) if ph is None else ph
    for gin, gout in ((gd0, gd1), (gd1, gd0)):
        a = rng.standard_normal((2,) + tuple(gin.n_c))
        if not real:
            a = a + 1j * rng.standard_normal(a.shape)
        mine = Transformer(gin, gout, nn).apply(a, ph)
        out = np.zeros((2,) + tuple(gout.n_c), a.dtype)
        ref.Transformer(gin, gout, nn, real).apply(a, out, phr)
        assert abs(mine - out).max() <= 1e-15 * max(1, abs(out).max())
