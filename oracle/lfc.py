"""Localized functions (projectors) -- oracle restatement (test infrastructure).

spline_to_grid follows lfc.c:887-977: for every point of a box, x = sum_c
g_c h_cv[c] - pos_v, r = |x|; points with r < rcut get spline(r) * Y_lm in C
(z-run) order, and each (g0, g1) row contributes one [begin, end) run of
flattened grid indices G measured from the grid corner beg_c.
LFC follows Sphere.set_position (lfc.py:77-119), LFC._update (lfc.py:273-374:
global M numbering atom-major, spline-major, m-minor; phase_qW =
exp(2 pi i k . sdisp_W)), integrate (lfc.py:552-631, lfc.c:432-505:
c_xM = dv * sum_G a_G A_GM [* phase_W]) and add (lfc.py:392-472,
lfc.c:818-885: a_G += sum_M A_GM c_M [* conj(phase_W)]).

The sums here are dense scatters (np.add.at), not the reference's event-list
sweep, so agreement with the reference is to rounding, not bitwise.

Pinned by tests/test_oracle_lfc.py against the reference C build
(cgpaw.spline_to_grid and cgpaw.LFC integrate/add), brute force over all
periodic images on the full grid, the adjointness
<add(c), psi> = <c, integrate(psi)> / dv, and the l=1 orthogonality of
test_derivatives.py:17-27.
"""
import numpy as np

from .spline import spherical_harmonics


def spline_to_grid(spline, beg_c, end_c, pos_v, h_cv, n_c, gdcorner_c):
    beg_c = np.asarray(beg_c, int)
    end_c = np.asarray(end_c, int)
    g = np.indices(end_c - beg_c).reshape(3, -1).T + beg_c  # C order
    x = h_cv[0, 0] * g[:, 0] + h_cv[1, 0] * g[:, 1] + h_cv[2, 0] * g[:, 2] \
        - pos_v[0]
    y = h_cv[0, 1] * g[:, 0] + h_cv[1, 1] * g[:, 1] + h_cv[2, 1] * g[:, 2] \
        - pos_v[1]
    z = h_cv[0, 2] * g[:, 0] + h_cv[1, 2] * g[:, 1] + h_cv[2, 2] * g[:, 2] \
        - pos_v[2]
    r2 = x * x + y * y + z * z
    r = np.sqrt(r2)
    inside = r < spline.get_cutoff()
    f = spline(r[inside])
    A_gm = spherical_harmonics(spline.l, f, x[inside], y[inside], z[inside])
    gi = g[inside] - np.asarray(gdcorner_c, int)
    G = (gi[:, 0] * n_c[1] + gi[:, 1]) * n_c[2] + gi[:, 2]
    # z-runs: consecutive G with the same (g0, g1)
    G_B = []
    if len(G):
        row = gi[:, 0] * n_c[1] + gi[:, 1]
        starts = np.flatnonzero(np.r_[True, row[1:] != row[:-1]])
        ends = np.r_[starts[1:], len(G)]
        for s, e in zip(starts, ends):
            G_B.extend([G[s], G[e - 1] + 1])
    return A_gm, np.array(G_B, np.intc), G


class LFC:
    """Serial LocalizedFunctionsCollection (no normalisation)."""

    def __init__(self, gd, spline_aj, k_c=None, dtype=float, cut=False):
        self.gd = gd
        self.cut = cut
        self.spline_aj = spline_aj
        self.k_c = None if k_c is None else np.asarray(k_c, float)
        self.dtype = dtype

    def set_positions(self, spos_ac):
        gd = self.gd
        spos_ac = np.asarray(spos_ac, float)
        self.volumes = []  # (atom, M_global, A_gm, G_g, sdisp_c)
        self.M_a = []
        self.Mmax_a = []
        M_atom = 0
        for a, (spos_c, spline_j) in enumerate(zip(spos_ac, self.spline_aj)):
            self.M_a.append(M_atom)
            M = 0
            for spline in spline_j:
                rcut = spline.get_cutoff()
                for beg_c, end_c, sdisp_c in gd.get_boxes(spos_c, rcut,
                                                          cut=self.cut):
                    pos_v = np.dot(spos_c - sdisp_c, gd.cell_cv)
                    A_gm, G_b, G_g = spline_to_grid(spline, beg_c, end_c,
                                                    pos_v, gd.h_cv, gd.n_c,
                                                    gd.beg_c)
                    if len(G_b):
                        self.volumes.append((a, M_atom + M, A_gm, G_g,
                                             np.array(sdisp_c)))
                M += 2 * spline.l + 1
            self.Mmax_a.append(M)
            M_atom += M
        self.Mmax = M_atom

    def _phase(self, sdisp_c):
        if self.k_c is None:
            return 1.0
        return np.exp(2j * np.pi * np.dot(self.k_c, sdisp_c))

    def integrate(self, a_xG):
        """Return c_xM (flattened M index) = dv * sum a(G) A(G) [phase]."""
        xshape = a_xG.shape[:-3]
        a = a_xG.reshape((-1,) + (int(np.prod(a_xG.shape[-3:])),))
        c = np.zeros((len(a), self.Mmax), a.dtype)
        for atom, M, A_gm, G_g, sdisp in self.volumes:
            ph = self._phase(sdisp)
            c[:, M:M + A_gm.shape[1]] += (a[:, G_g] * (ph * self.gd.dv)) @ A_gm
        return c.reshape(xshape + (self.Mmax,))

    def add(self, a_xG, c_xM):
        """a_xG += sum_M A(G) c_M [conj phase] (in place)."""
        shape = a_xG.shape
        a = a_xG.reshape((-1, int(np.prod(shape[-3:]))))
        c = c_xM.reshape((-1, self.Mmax))
        for atom, M, A_gm, G_g, sdisp in self.volumes:
            ph = np.conj(self._phase(sdisp))
            contrib = (c[:, M:M + A_gm.shape[1]] @ A_gm.T) * ph
            for x in range(len(a)):
                np.add.at(a[x], G_g, contrib[x])
        return a_xG

    def dict_from_M(self, c_xM):
        return {a: c_xM[..., M:M + n] for a, (M, n) in
                enumerate(zip(self.M_a, self.Mmax_a))}

    def dense_matrix(self):
        """Dense (Mmax, N) matrix of the functions (without phases); for
        small test grids only."""
        N = int(np.prod(self.gd.n_c))
        dt = complex if self.k_c is not None else float
        B = np.zeros((self.Mmax, N), dt)
        for atom, M, A_gm, G_g, sdisp in self.volumes:
            ph = np.conj(self._phase(sdisp))
            for m in range(A_gm.shape[1]):
                np.add.at(B[M + m], G_g, A_gm[:, m] * ph)
        return B
