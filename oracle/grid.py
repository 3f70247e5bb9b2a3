"""Serial grid geometry -- oracle restatement (test infrastructure only).

Follows gpaw/grid_descriptor.py:77-160 (N_c, n_c, beg_c, end_c, h_cv, dv),
:328-341 (coarsen/refine), :344-411 (get_boxes) and gpaw/domain.py:110-139
(neighbor_cd, sdisp_cd) for a single domain.

Pinned by: tests/test_grid_boxes.py (boxes cover exactly the
grid points inside a sphere for brute-force enumeration over periodic images).
"""
import numpy as np

BOHR_ANGSTROM = 0.52917721067  # ase.units.Bohr (CODATA 2014), Angstrom


class Grid:
    """Single-domain grid descriptor.

    For a non-periodic axis the first grid point (index 0) is excluded, as in
    grid_descriptor.py:127-131 (n_p[0] = 1).
    """

    def __init__(self, N_c, cell_cv, pbc_c=True):
        self.N_c = np.array(N_c, int)
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
        # domain.py:110-139 with one domain: neighbour is self (rank 0) on a
        # periodic axis, -1 otherwise; sdisp = +1 for d=0, -1 for d=1.
        self.neighbor_cd = np.zeros((3, 2), int)
        self.sdisp_cd = np.zeros((3, 2), int)
        for c in range(3):
            if self.pbc_c[c]:
                self.sdisp_cd[c] = (1, -1)
            else:
                self.neighbor_cd[c] = -1

    def coarsen(self):
        assert not (self.N_c % 2).any()
        return Grid(self.N_c // 2, self.cell_cv, self.pbc_c)

    def refine(self):
        return Grid(self.N_c * 2, self.cell_cv, self.pbc_c)

    def zeros(self, n=(), dtype=float):
        if isinstance(n, int):
            n = (n,)
        return np.zeros(tuple(n) + tuple(self.n_c), dtype)

    def coordinates(self):
        """r_vg of the grid points (grid_descriptor.py:708-711)."""
        return np.dot(np.indices(self.n_c, float).T + self.beg_c,
                      self.h_cv).T.copy()

    def integrate(self, a_xg):
        return a_xg.reshape(a_xg.shape[:-3] + (-1,)).sum(-1) * self.dv

    def phase_cd(self, k_c):
        """kpt_descriptor.py:283-284: exp(2 pi i sdisp_cd k_c)."""
        k_c = np.asarray(k_c, float)
        return np.exp(2j * np.pi * self.sdisp_cd * k_c[:, None])

    def get_boxes(self, spos_c, rcut, cut=True):
        """grid_descriptor.py:344-411 (serial domain)."""
        N_c = self.N_c
        ncut = rcut * (self.icell_cv ** 2).sum(axis=1) ** 0.5 * N_c
        npos_c = spos_c * N_c
        beg_c = np.ceil(npos_c - ncut).astype(int)
        end_c = np.ceil(npos_c + ncut).astype(int)
        if cut:
            for c in range(3):
                if not self.pbc_c[c]:
                    beg_c[c] = max(beg_c[c], 0)
                    end_c[c] = min(end_c[c], N_c[c])
        else:
            for c in range(3):
                if not self.pbc_c[c] and (beg_c[c] < 0 or end_c[c] > N_c[c]):
                    raise ValueError('box crosses non-periodic boundary')
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
