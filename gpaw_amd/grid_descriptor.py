"""Serial grid descriptor: the geometry the operators are built from.

Mirrors the attributes of the reference's GridDescriptor for one domain
(gpaw/grid_descriptor.py:77-160): N_c, cell_cv, pbc_c, n_c, beg_c, end_c,
h_cv, dv, neighbor_cd and sdisp_cd (gpaw/domain.py:110-139), plus coarsen /
refine (:328-341), get_boxes (:344-411) and array constructors that return
device tensors.  Lengths are in Bohr.  On a non-periodic axis the first grid
point is excluded (n_c = N_c - 1, beg_c = 1) as in the reference.
"""
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

    def get_boxes(self, spos_c, rcut, cut=True):
        """Boxes of grid points enclosing a sphere, one per periodic image
        (grid_descriptor.py:344-411)."""
        N_c = self.N_c
        ncut = rcut * (self.icell_cv ** 2).sum(axis=1) ** 0.5 * N_c
        npos_c = np.asarray(spos_c) * N_c
        beg_c = np.ceil(npos_c - ncut).astype(int)
        end_c = np.ceil(npos_c + ncut).astype(int)
        for c in range(3):
            if not self.pbc_c[c]:
                if cut:
                    beg_c[c] = max(beg_c[c], 0)
                    end_c[c] = min(end_c[c], N_c[c])
                elif beg_c[c] < 0 or end_c[c] > N_c[c]:
                    raise GridBoundsError(
                        'Box at %s crosses boundary' % (tuple(spos_c),))
        range_c = ([], [], [])
        for c in range(3):
            b = beg_c[c]
            e = b
            while e < end_c[c]:
                b0 = b % N_c[c]
                e = min(end_c[c], b + N_c[c] - b0)
                b1 = b + self.beg_c[c] - b0 if b0 < self.beg_c[c] else b
                e0 = b0 - b + e
                e1 = e - (e0 - self.end_c[c]) if e0 > self.end_c[c] else e
                if e1 > b1:
                    range_c[c].append((b1, e1))
                b = e
        boxes = []
        for b0, e0 in range_c[0]:
            for b1, e1 in range_c[1]:
                for b2, e2 in range_c[2]:
                    b = np.array((b0, b1, b2))
                    e = np.array((e0, e1, e2))
                    bb = b % N_c
                    ee = bb + e - b
                    disp = (b - bb) // N_c
                    bb = np.maximum(bb, self.beg_c)
                    ee = np.minimum(ee, self.end_c)
                    if (bb < ee).all():
                        boxes.append((bb, ee, disp))
        return boxes


class GridBoundsError(ValueError):
    pass
