"""Finite-difference operators (drop-in for gpaw.fd_operators).

FDOperator / Laplace / GUCLaplace keep the reference's constructor and
method signatures (gpaw/fd_operators.py:36-141,171-232).  The stencil runs
in libgpawhip (gh_stencil_*); the Laplacian's points and coefficients are
derived on the host (guc_laplace_stencil) and agree with the reference's.
"""
import ctypes
import itertools

import numpy as np
import torch

from . import _lib
from ._arrays import (Staged, dtype_code, is_complex, real_dtype,
                      require_gpu, stream_ptr, torch_dtype)

# Expansion coefficients of the finite-difference Laplacian, J. R.
# Chelikowsky et al., Phys. Rev. B 50, 11355 (1994) -- the table of
# fd_operators.py:21-28 (a published table; kept identical).
laplace = [[0],
           [-2, 1],
           [-5 / 2, 4 / 3, -1 / 12],
           [-49 / 18, 3 / 2, -3 / 20, 1 / 90],
           [-205 / 72, 8 / 5, -1 / 5, 8 / 315, -1 / 560],
           [-5269 / 1800, 5 / 3, -5 / 21, 5 / 126, -5 / 1008, 1 / 3150],
           [-5369 / 1800, 12 / 7, -15 / 56, 10 / 189, -1 / 112, 2 / 1925,
            -1 / 16632]]


class FDOperator:
    """Stencil operator out = sum_p coef_p in[g + offset_p] with periodic
    (Bloch) or zero boundaries (fd_operators.py:36-141)."""

    def __init__(self, coef_p, offset_pc, gd, dtype=float, description=None,
                 xp=None):
        offset_pc = np.asarray(offset_pc, int).reshape(-1, 3)
        coef_p = np.ascontiguousarray(coef_p, float)
        assert coef_p.ndim == 1 and len(coef_p) == len(offset_pc)
        mp = np.abs(offset_pc).max()
        assert (mp <= gd.n_c).all(), 'Stencil longer than domain'
        self.cfd = all(sum(o != 0 for o in off) < 2 for off in offset_pc)
        self.mp = mp
        self.gd = gd
        self.dtype = dtype
        self.tdtype = torch_dtype(dtype)
        self.shape = tuple(int(n) for n in gd.n_c)
        self.npoints = len(coef_p)
        self.coef_p = coef_p
        self.offset_pc = offset_pc
        M_c = gd.n_c + 2 * mp
        self.offset_p = np.dot(offset_pc, [M_c[1] * M_c[2], M_c[2], 1])
        self.xp = xp
        L = _lib.lib()
        h = ctypes.c_void_p()
        self._keep = [_lib.f64(coef_p), _lib.i64(offset_pc.ravel()),
                      _lib.i64(gd.n_c), _lib.i32(gd.pbc_c.astype(np.int32))]
        _lib.check(L.gh_stencil_create(
            ctypes.byref(h), len(coef_p), self._keep[0][1], self._keep[1][1],
            self._keep[2][1], self._keep[3][1], dtype_code(dtype)))
        self.handle = h
        self._lib = L
        self._work = None
        if description is None:
            description = '%d point finite-difference stencil' % self.npoints
        self.description = description

    def __del__(self):
        h = getattr(self, 'handle', None)
        if h is not None and h.value:
            self._lib.gh_stencil_destroy(h)
            self.handle = None

    def __str__(self):
        return '<' + self.description + '>'

    def apply(self, in_xg, out_xg, phase_cd=None, vt_g=None, src_xg=None,
              beta=1.0):
        """out = L in  [+ vt * in] [+ beta * src]  (fd_operators.py:114).

        vt_g (optional) fuses the local potential of
        Hamiltonian.apply_local_potential (hamiltonian.py:437-443)."""
        assert tuple(in_xg.shape) == tuple(out_xg.shape)
        assert tuple(in_xg.shape[-3:]) == self.shape
        nb = int(np.prod(in_xg.shape[:-3], dtype=np.int64))
        ph, php = _lib.phase_arg(phase_cd)
        with Staged(in_xg, dtype=self.dtype) as a, \
                Staged(out_xg, writeback=True, dtype=self.dtype) as b:
            vt = src = None
            if vt_g is not None:
                vt = Staged(vt_g, dtype=real_dtype(self.dtype)).t
                assert tuple(vt.shape) == self.shape
            if src_xg is not None:
                src = Staged(src_xg, dtype=self.dtype).t
                assert src.shape == a.shape
            _lib.check(self._lib.gh_stencil_apply(
                self.handle, a.data_ptr(), b.data_ptr(), nb, php,
                vt.data_ptr() if vt is not None else None,
                src.data_ptr() if src is not None else None, float(beta),
                stream_ptr()))

    def apply_shifted(self, in_xg, out_xg, eps_x, phase_cd=None, vt_g=None):
        """out[x] = (L + vt - eps[x]) in[x] for device arrays: the local part
        of the residual (H - eps) psi (eigensolver.py:153-172) in one pass;
        eps_x: per-band FP64 values.  Non-orthogonal stencils run the plain
        apply followed by a band-wise axpy."""
        assert tuple(in_xg.shape) == tuple(out_xg.shape)
        assert tuple(in_xg.shape[-3:]) == self.shape
        assert in_xg.is_cuda and out_xg.is_cuda
        assert in_xg.dtype == out_xg.dtype == torch_dtype(self.dtype)
        assert in_xg.is_contiguous() and out_xg.is_contiguous()
        nb = int(np.prod(in_xg.shape[:-3], dtype=np.int64))
        eps = torch.as_tensor(eps_x, dtype=torch.float64,
                              device='cuda').reshape(-1).contiguous()
        assert eps.numel() == nb
        if not self.cfd:
            from . import blas
            self.apply(in_xg, out_xg, phase_cd, vt_g=vt_g)
            gsize = int(np.prod(self.shape))
            blas.multi_axpy(-eps, in_xg.reshape(nb, gsize),
                            out_xg.reshape(nb, gsize))
            return
        ph, php = _lib.phase_arg(phase_cd)
        vt = None
        if vt_g is not None:
            vt = Staged(vt_g, dtype=real_dtype(self.dtype)).t
            assert tuple(vt.shape) == self.shape
        _lib.check(self._lib.gh_stencil_apply_shifted(
            self.handle, in_xg.data_ptr(), out_xg.data_ptr(), nb, php,
            vt.data_ptr() if vt is not None else None, eps.data_ptr(),
            stream_ptr()))

    def __call__(self, in_xR, out_xR=None):
        if out_xR is None:
            out_xR = in_xR.new_empty(in_xR.shape) if hasattr(
                in_xR, 'new_empty') else np.empty_like(in_xR)
        self.apply(in_xR, out_xR)
        return out_xR

    def relax(self, relax_method, f_g, s_g, n, w=None):
        """n relaxation sweeps of L f = s in place on f (fd_operators.py:126,
        c/operators.c:55-95, c/bmgs/relax.c).  relax_method 2: Jacobi with
        weight w (default 1); relax_method 1: Gauss-Seidel in the
        reference's C-order (w unused), run as one launch per hyperplane
        i0 + i1 + i2 = const (stencils along the grid axes only)."""
        if relax_method not in (1, 2):
            raise ValueError('relax_method %r: 1 (Gauss-Seidel) or 2 '
                             '(Jacobi)' % (relax_method,))
        if relax_method == 1 and not self.cfd:
            raise NotImplementedError(
                'Gauss-Seidel relaxation needs a stencil along the grid axes')
        assert tuple(f_g.shape) == self.shape == tuple(s_g.shape)
        assert not is_complex(self.dtype)
        w = 1.0 if w is None else float(w)
        with Staged(f_g, writeback=True, dtype=self.dtype) as f, \
                Staged(s_g, dtype=self.dtype) as s:
            if self._work is None or self._work.shape != f.shape:
                self._work = torch.empty_like(f)
            _lib.check(self._lib.gh_stencil_relax(
                self.handle, int(relax_method), f.data_ptr(), s.data_ptr(),
                self._work.data_ptr(), int(n), w, stream_ptr()))

    def get_diagonal_element(self):
        return self._lib.gh_stencil_diagonal(self.handle)

    def set_tile(self, rows=0, planes=0, depth=0):
        """Tile shape of the stencil kernel: rows per block, x-planes per
        block, prefetch depth (0 = automatic)."""
        _lib.check(self._lib.gh_stencil_set_tile(self.handle, int(rows),
                                                 int(planes), int(depth)))


def Laplace(gd, scale=1.0, n=1, dtype=float, xp=None):
    if n == 9:
        raise NotImplementedError('FFT Laplacian (n=9) is not provided')
    return GUCLaplace(gd, scale, n, dtype, xp=xp)


def _neighbour_directions():
    """The 13 directions d of the 3 x 3 x 3 grid-point cube with d > 0
    lexicographically (one of each +-d pair), in lexicographic order."""
    return np.array([d for d in itertools.product((-1, 0, 1), repeat=3)
                     if d > (0, 0, 0)])


def _second_moments(t_dv):
    """Row d: the six independent entries (xx, yy, zz, yz, xz, xy) of the
    outer product t_d t_d^T of the Cartesian step vector t_d."""
    x, y, z = t_dv.T
    return np.stack([x * x, y * y, z * z, y * z, x * z, x * y], axis=1)


def guc_laplace_stencil(h_cv, scale=1.0, n=1):
    """Coefficients and offsets of the order-2n Laplacian on a grid with
    step vectors h_cv (the stencil of fd_operators.py:171-232).

    Along a grid direction d with Cartesian step t_d, the 1-D expansion
    w_0 f(r) + sum_j w_j (f(r + j t_d) + f(r - j t_d)) (Chelikowsky weights)
    approximates (t_d . grad)^2 f.  Summing directions with weights a_d gives
    the Laplacian when sum_d a_d t_d t_d^T = 1 (3 x 3).  Directions are taken
    shortest first, the length measured with every grid axis scaled to unit
    length and ranked by NumPy's default argsort, as the reference ranks them
    (so equal lengths come in that sort's order, which is not always the
    lexicographic one: the body-centred cubic cell); the first D = 3..6
    of them for which the six second-moment equations have an exact
    solution are used.  An orthorhombic cell needs only the three axes, with
    a_d = 1 / h_d^2.

    Returns (coefs, offsets, D): the centre first, then per direction the
    points +1..+n and -1..-n."""
    h_cv = np.asarray(h_cv, float)
    dirs = _neighbour_directions()
    unit_cv = h_cv / np.linalg.norm(h_cv, axis=1)[:, None]
    # equal lengths must compare equal, so the key is formed as a plain sum
    # of squared components (x^2 + y^2) + z^2
    rank_key = np.sum((dirs @ unit_cv) ** 2, axis=1)
    order = np.argsort(rank_key)
    if np.count_nonzero(h_cv - np.diag(np.diag(h_cv))) == 0:
        ndir = 3
        used = dirs[order[:3]]
        a_d = 1.0 / np.einsum('dv,dv->d', used @ h_cv, used @ h_cv)
    else:
        target = np.array([1.0, 1.0, 1.0, 0.0, 0.0, 0.0])
        for ndir in range(3, 7):
            used = dirs[order[:ndir]]
            B = _second_moments(used @ h_cv).T          # (6, ndir)
            a_d = np.linalg.lstsq(B, target, rcond=None)[0]
            if np.sum((B @ a_d - target) ** 2) < 1e-14:
                if np.linalg.matrix_rank(B) < ndir:
                    raise ValueError(
                        'grid directions are degenerate for this cell; '
                        'build the Laplacian on a reduced (Niggli) cell')
                break
        else:
            raise ValueError('no finite-difference Laplacian for this cell')
    w = np.asarray(laplace[n], float)
    a_d = a_d * scale
    steps = np.arange(1, n + 1)[:, None]
    coefs = [w[0] * a_d.sum()]
    offsets = [np.zeros((1, 3), int)]
    for a, d in zip(a_d, used):
        coefs.extend(np.concatenate([a * w[1:], a * w[1:]]))
        offsets.extend([steps * d, -steps * d])
    return np.array(coefs), np.concatenate(offsets), ndir


class GUCLaplace(FDOperator):
    def __init__(self, gd, scale=1.0, n=1, dtype=float, xp=None):
        """Laplacian for general non orthorhombic grid; scale=-0.5 gives the
        kinetic energy operator, n the range (error O(h^2n))."""
        coefs, offsets, D = guc_laplace_stencil(gd.h_cv, scale, n)
        FDOperator.__init__(self, coefs, offsets, gd, dtype, xp=xp)
        self.ndirections = D
        self.description = (
            'finite-difference Laplacian: %d points (%d directions x 2 x %d '
            '+ centre), error O(h^%d)' % (self.npoints, D, n, 2 * n))
