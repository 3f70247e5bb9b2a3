"""Serial grid descriptor: the geometry the operators are built from.

Mirrors the attributes of the reference's GridDescriptor for one domain
(gpaw/grid_descriptor.py:77-160): N_c, cell_cv, pbc_c, n_c, beg_c, end_c,
h_cv, dv, neighbor_cd and sdisp_cd (gpaw/domain.py:110-139), plus coarsen /
refine (:328-341), get_boxes (:344-411) and array constructors that return
device tensors.  Lengths are in Bohr.  On a non-periodic axis the first grid
point is excluded (n_c = N_c - 1, beg_c = 1) as in the reference.
"""
import itertools

import numpy as np
import torch

from ._arrays import torch_dtype


class GridDescriptor:
    ndim = 3

    def __init__(self, N_c, cell_cv=(1, 1, 1), pbc_c=True, comm=None):
        if comm is not None and getattr(comm, 'size', 1) != 1:
            raise NotImplementedError('domain decomposition is not built: '
                                      'one grid per GPU')
        self.N_c = np.array(N_c, int)
        if (self.N_c != np.asarray(N_c)).any():
            raise ValueError('Non-int number of grid points %s' % (N_c,))
        cell_cv = np.array(cell_cv, float)
        if cell_cv.ndim == 1:
            cell_cv = np.diag(cell_cv)
        self.cell_cv = cell_cv
        if isinstance(pbc_c, (bool, int, np.bool_)):
            pbc_c = (bool(pbc_c),) * 3
        self.pbc_c = np.array(pbc_c, bool)
        self.beg_c = np.where(self.pbc_c, 0, 1)
        self.end_c = self.N_c.copy()
        self.n_c = self.end_c - self.beg_c
        self.h_cv = self.cell_cv / self.N_c[:, None]
        self.volume = abs(np.linalg.det(self.cell_cv))
        self.dv = self.volume / self.N_c.prod()
        self.icell_cv = np.linalg.inv(self.cell_cv).T
        self.orthogonal = not (self.cell_cv -
                               np.diag(self.cell_cv.diagonal())).any()
        self.neighbor_cd = np.zeros((3, 2), int)
        self.sdisp_cd = np.zeros((3, 2), int)
        for c in range(3):
            if self.pbc_c[c]:
                self.sdisp_cd[c] = (1, -1)
            else:
                self.neighbor_cd[c] = -1

    def __repr__(self):
        return 'GridDescriptor(%s, pbc=%s)' % (self.N_c.tolist(),
                                               self.pbc_c.tolist())

    def new_descriptor(self, N_c=None, cell_cv=None, pbc_c=None):
        return GridDescriptor(self.N_c if N_c is None else N_c,
                              self.cell_cv if cell_cv is None else cell_cv,
                              self.pbc_c if pbc_c is None else pbc_c)

    def coarsen(self):
        if (self.N_c % 2).any():
            raise ValueError('Grid %s not divisible by 2!' % self.N_c)
        return self.new_descriptor(self.N_c // 2)

    def refine(self):
        return self.new_descriptor(self.N_c * 2)

    def phase_cd(self, k_c):
        """Bloch phases exp(2 pi i sdisp_cd k_c) (kpt_descriptor.py:283)."""
        k_c = np.asarray(k_c, float)
        return np.exp(2j * np.pi * self.sdisp_cd * k_c[:, None])

    # ------------------------------------------------------------ arrays --
    def zeros(self, n=(), dtype=float, device='cuda'):
        if isinstance(n, int):
            n = (n,)
        return torch.zeros(tuple(n) + tuple(self.n_c),
                           dtype=torch_dtype(dtype), device=device)

    def empty(self, n=(), dtype=float, device='cuda'):
        if isinstance(n, int):
            n = (n,)
        return torch.empty(tuple(n) + tuple(self.n_c),
                           dtype=torch_dtype(dtype), device=device)

    def get_grid_point_coordinates(self):
        return np.dot(np.indices(self.n_c, float).T + self.beg_c,
                      self.h_cv).T.copy()

    def integrate(self, a_xg, b_yg=None, global_integral=True,
                  hermitian=False):
        """Integrals over the grid (grid_descriptor.py:275-326).

        One array: a_x = dv sum_g a_xg.  Two: r[x, y] = dv sum_g
        conj(a_xg) b_yg -- the rk form when b_yg is a_xg, the r2k form
        (symmetrised) when hermitian, otherwise the mmm form -- on the
        matrix cores with FP64 sums (gh_gram).  Returns a device tensor, or
        a Python scalar for a 0-d result."""
        from .utilities.blas import r2k, rk, mmm
        a = a_xg if isinstance(a_xg, torch.Tensor) else torch.from_numpy(
            np.ascontiguousarray(a_xg)).cuda()
        xshape = tuple(a.shape[:-3])
        gsize = int(np.prod(a.shape[-3:]))
        if b_yg is None:
            result = a.reshape(xshape + (gsize,)).sum(dim=-1) * self.dv
            return result.item() if result.ndim == 0 else result
        same = b_yg is a_xg
        b = a if same else (b_yg if isinstance(b_yg, torch.Tensor) else
                            torch.from_numpy(np.ascontiguousarray(b_yg)).cuda())
        A = a.reshape((-1, gsize)).contiguous()
        B = b.reshape((-1, gsize)).contiguous()
        mdt = torch.complex128 if A.is_complex() else torch.float64
        if A.dtype != mdt:
            # single-precision data: FP32 products, FP64 sums (gh_gram)
            from .blas import gram
            if same:
                res = gram(A, alpha=self.dv, hermitian=True)
            elif hermitian:
                G = gram(B, A)
                res = 0.5 * self.dv * (G + G.conj().T)
            else:
                res = gram(B, A, alpha=self.dv)
        else:
            res = torch.zeros((len(B), len(A)), dtype=mdt, device='cuda')
            if same:
                rk(self.dv, A, 0.0, res)
            elif hermitian:
                r2k(0.5 * self.dv, A, B, 0.0, res)
            else:
                mmm(self.dv, B, 'N', A, 'C', 0.0, res)
        yshape = tuple(b.shape[:-3])
        result = res.T.reshape(xshape + yshape)
        return result.item() if result.ndim == 0 else result

    def get_boxes(self, spos_c, rcut, cut=True):
        """Boxes of grid points that may lie within rcut of the scaled
        position spos_c, one box per periodic image the sphere reaches
        (the contract of grid_descriptor.py:344-411).

        Each box is (beg_c, end_c, disp_c): a half-open index range of this
        domain and the image (in cells, along each axis) that the sphere
        centre must be shifted by to cover it.  Along axis c the sphere spans
        the unwrapped index range [ceil(x - r), ceil(x + r)) with x = spos_c
        N_c and r = rcut / (distance between lattice planes) grid steps; the
        range is cut at multiples of N_c into one piece per image, and each
        piece is clipped to the domain.  Boxes come in the order of the
        images, axis 0 slowest."""
        spos_c = np.asarray(spos_c, float)
        steps_c = rcut * np.sqrt((self.icell_cv ** 2).sum(axis=1)) * self.N_c
        lo_c = np.ceil(spos_c * self.N_c - steps_c).astype(int)
        hi_c = np.ceil(spos_c * self.N_c + steps_c).astype(int)
        for c in np.flatnonzero(~self.pbc_c):
            if not cut and (lo_c[c] < 0 or hi_c[c] > self.N_c[c]):
                raise GridBoundsError(
                    'sphere at scaled position %s reaches beyond the '
                    'non-periodic boundary of axis %d' % (tuple(spos_c), c))
            lo_c[c] = max(lo_c[c], 0)
            hi_c[c] = min(hi_c[c], self.N_c[c])
        pieces_c = [self._image_pieces(lo_c[c], hi_c[c], c) for c in range(3)]
        return [(np.array(beg), np.array(end), np.array(disp))
                for beg, end, disp in
                (zip(*combo) for combo in itertools.product(*pieces_c))]

    def _image_pieces(self, lo, hi, c):
        """[(beg, end, image)] of the unwrapped range [lo, hi) along axis c,
        one piece per periodic image, clipped to [beg_c, end_c)."""
        n = int(self.N_c[c])
        out = []
        for image in range(lo // n, -(-hi // n)):
            beg = max(lo - image * n, int(self.beg_c[c]))
            end = min(hi - image * n, int(self.end_c[c]))
            if beg < end:
                out.append((beg, end, image))
        return out


class GridBoundsError(ValueError):
    pass
