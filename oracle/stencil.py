"""Finite-difference stencils -- oracle restatement (test infrastructure only).

laplace_stencil  follows fd_operators.py:178-232 (GUCLaplace: the 13 shortest
                 lattice directions, exact second-moment condition, Chelikowsky
                 table :21-28, offset order centre, +1..+n, -1..-n per
                 direction).
pad              follows bc.c:146-258 with one domain: zero everything, paste
                 the interior, then per axis i=0,1,2 copy the low interior slab
                 into the high ghost times phase_cd[i,0] and the high interior
                 slab into the low ghost times phase_cd[i,1]; non-periodic
                 ghosts stay zero.  Later axes copy whole padded planes of
                 earlier axes (cfd=0 of bc_init), so corners carry products of
                 phases.
fd_apply         follows bmgs/fd.c:9-33: out[j] = sum_c coef_c * a[i + off_c].

Pinned by tests/test_oracle_stencil.py against the reference C build
(oracle/_ref: bc_unpack1/2 + bmgs_fd via cgpaw.Operator), polynomial exactness
(test_laplace.py:82-93), the plane-wave symbol of the stencil and the analytic
Laplacian of a Gaussian.
"""
import numpy as np

CHELIKOWSKY = [[0],
               [-2, 1],
               [-5 / 2, 4 / 3, -1 / 12],
               [-49 / 18, 3 / 2, -3 / 20, 1 / 90],
               [-205 / 72, 8 / 5, -1 / 5, 8 / 315, -1 / 560],
               [-5269 / 1800, 5 / 3, -5 / 21, 5 / 126, -5 / 1008, 1 / 3150],
               [-5369 / 1800, 12 / 7, -15 / 56, 10 / 189, -1 / 112, 2 / 1925,
                -1 / 16632]]


def laplace_stencil(h_cv, scale=1.0, n=1):
    """Return (coef_p, offset_pc) of GUCLaplace (fd_operators.py:178-232).

    Candidate directions: the 13 points of the 3x3x3 cube after the centre in
    C order.  They are ranked by |sum_c m_c u_c|^2 with u_c the grid axes
    scaled to unit length with numpy's default argsort (the reference's
    ranking, including its order among ties); the first D = 3, 4, 5, 6 are
    tried until sum_d a_d t_d t_d^T = 1 (t_d = m_d . h_cv) has an exact
    solution, found here from the normal equations.  Points: the centre,
    then per direction +1..+n and -1..-n, weights from the Chelikowsky
    table."""
    h_cv = np.asarray(h_cv, float)
    cands = [(i // 9 - 1, (i // 3) % 3 - 1, i % 3 - 1) for i in range(14, 27)]
    u_cv = h_cv / np.sqrt((h_cv ** 2).sum(1))[:, None]
    key = np.sum((np.array(cands, float) @ u_cv) ** 2, axis=1)
    ranked = list(np.argsort(key))
    pairs = [(0, 0), (1, 1), (2, 2), (1, 2), (0, 2), (0, 1)]
    rhs = np.array([1.0, 1.0, 1.0, 0.0, 0.0, 0.0])
    for D in range(3, 7):
        t = [np.dot(cands[i], h_cv) for i in ranked[:D]]
        B = np.array([[tv[p] * tv[q] for tv in t] for p, q in pairs])
        BtB = B.T @ B
        if np.linalg.matrix_rank(BtB) < D:
            if D == 6:
                raise AssertionError('weird unit cell')
            continue
        a_d = np.linalg.solve(BtB, B.T @ rhs)
        if np.sum((B @ a_d - rhs) ** 2) < 1e-14:
            break
    a_d = a_d * scale
    w = CHELIKOWSKY[n]
    offsets = [(0, 0, 0)]
    coefs = [w[0] * a_d.sum()]
    for d in range(D):
        m = np.array(cands[ranked[d]])
        for sign in (1, -1):
            for j in range(1, n + 1):
                offsets.append(tuple(sign * j * m))
                coefs.append(a_d[d] * w[j])
    return np.array(coefs, float), np.array(offsets, int)


def pad(a_xg, pad_cd, pbc_c, phase_cd=None):
    """Padded copy of a_xg (..., n0, n1, n2) with periodic/Bloch/zero halos.

    pad_cd[c] = (low, high) ghost widths; a periodic ghost wraps at most once.
    """
    pad_cd = np.asarray(pad_cd, int)
    xshape = a_xg.shape[:-3]
    n_c = np.array(a_xg.shape[-3:])
    complex_ = np.iscomplexobj(a_xg)
    b = np.zeros(xshape + tuple(n_c + pad_cd.sum(1)), a_xg.dtype)
    nx = len(xshape)
    sl = [slice(None)] * nx + [slice(p, p + n) for p, n in zip(pad_cd[:, 0],
                                                                n_c)]
    b[tuple(sl)] = a_xg
    for i in range(3):
        lo, hi = pad_cd[i]
        n = n_c[i]
        if not pbc_c[i]:
            continue
        assert lo <= n and hi <= n, 'halo wider than domain'
        ax = nx + i

        def take(start, stop):
            idx = [slice(None)] * b.ndim
            idx[ax] = slice(start, stop)
            return tuple(idx)
        # low interior slab -> high ghost, times phase_cd[i, 0]
        if hi > 0:
            src = b[take(lo, lo + hi)].copy()
            if complex_:
                src = src * phase_cd[i, 0]
            b[take(lo + n, lo + n + hi)] = src
        # high interior slab -> low ghost, times phase_cd[i, 1]
        if lo > 0:
            src = b[take(lo + n - lo, lo + n)].copy()
            if complex_:
                src = src * phase_cd[i, 1]
            b[take(0, lo)] = src
    return b


def fd_apply(coef_p, offset_pc, a_xg, pbc_c, phase_cd=None):
    """out = sum_p coef_p * a[g + offset_p] (bmgs/fd.c:9-33 over a padded copy)."""
    offset_pc = np.asarray(offset_pc, int)
    mp = int(np.abs(offset_pc).max())
    n_c = np.array(a_xg.shape[-3:])
    assert (mp <= n_c).all(), 'Stencil longer than domain'
    b = pad(a_xg, [(mp, mp)] * 3, pbc_c, phase_cd)
    out = np.zeros_like(a_xg)
    nx = a_xg.ndim - 3
    for coef, off in zip(coef_p, offset_pc):
        idx = [slice(None)] * nx + [slice(mp + o, mp + o + n)
                                    for o, n in zip(off, n_c)]
        out += coef * b[tuple(idx)]
    return out
