"""Localized functions collection (drop-in for gpaw.lfc.LFC).

LFC(gd, spline_aj, kd, cut, dtype) / set_positions / dict / integrate / add
keep the signatures and the coefficient layout of the reference
(gpaw/lfc.py:219-631, c/lfc.c): c_axi[a] has shape xshape + (ni,), functions
numbered atom-major, spline-major, m-minor (lfc.py:277-331).

Set-up (per geometry): for every atom, spline and periodic image box
(gd.get_boxes, lfc.py:77-119) the function values are evaluated by
gh_spline_to_grid; the atom's values on all its (grid point, image) pairs are
merged into one dense table of nm_a columns and handed to gh_lfc_create.
integrate / add then run entirely on the GPU.
"""
import ctypes
from math import pi

import numpy as np
import torch

from . import _lib
from ._arrays import Staged, dtype_code, is_complex, stream_ptr, torch_dtype
from .grid_descriptor import GridBoundsError


class LocalizedFunctionsCollection:
    def __init__(self, gd, spline_aj, kd=None, cut=False, dtype=float,
                 integral=None, forces=None, xp=None):
        # integral: normalise on the grid (lfc.py:125-179) -- a number or
        # one per atom; the first function of each atom must be l = 0
        self.integral = integral
        self.gd = gd
        self.spline_aj = [list(s) for s in spline_aj]
        self.cut = cut
        self.dtype = dtype
        self.xp = xp
        if kd is None:
            self.ibzk_qc = np.zeros((1, 3))
            self.gamma = True
        else:
            self.ibzk_qc = np.atleast_2d(np.asarray(
                getattr(kd, 'ibzk_qc', kd), float))
            self.gamma = bool(getattr(kd, 'gamma',
                                      not self.ibzk_qc.any()))
        self.my_atom_indices = None
        self.handle = None
        self._lib = _lib.lib()
        self.spos_ac = None

    def __del__(self):
        self._free()

    def _free(self):
        h = getattr(self, 'handle', None)
        if h is not None and h.value:
            self._lib.gh_lfc_destroy(h)
        self.handle = None

    # ------------------------------------------------------------ set-up --
    def set_positions(self, spos_ac, atom_partition=None):
        spos_ac = np.asarray(spos_ac, float)
        assert len(spos_ac) == len(self.spline_aj)
        if self.spos_ac is not None and self.spos_ac.shape == spos_ac.shape \
                and not (self.spos_ac - spos_ac).any():
            return
        self._update(spos_ac)
        self.spos_ac = spos_ac.copy()

    def _update(self, spos_ac):
        gd = self.gd
        L = self._lib
        ngpts = int(np.prod(gd.n_c))
        h_cv = np.ascontiguousarray(gd.h_cv, float)
        sdisp_list = []
        sdisp_index = {}
        G_parts, w_parts, A_parts = [], [], []
        entry_ptr = [0]
        self.M_a = []
        self.Mmax_a = []
        M = 0
        for a, (spos_c, spline_j) in enumerate(zip(spos_ac, self.spline_aj)):
            self.M_a.append(M)
            nm_a = sum(2 * s.l + 1 for s in spline_j)
            keys, cols, vals = [], [], []
            Mj = 0
            for spline in spline_j:
                nm = 2 * spline.l + 1
                try:
                    boxes = gd.get_boxes(spos_c, spline.get_cutoff(),
                                         self.cut)
                except GridBoundsError as e:
                    e.args = ['Atom %d too close to edge: %s' % (a, e)]
                    raise
                for beg_c, end_c, sdisp_c in boxes:
                    key = tuple(int(s) for s in sdisp_c)
                    if key not in sdisp_index:
                        sdisp_index[key] = len(sdisp_list)
                        sdisp_list.append(key)
                    w = sdisp_index[key]
                    npts = int(np.prod(end_c - beg_c))
                    A = np.empty((npts, nm))
                    G = np.empty(npts, np.int32)
                    pos_v = np.dot(spos_c - sdisp_c, gd.cell_cv)
                    args = (_lib.i64(beg_c), _lib.i64(end_c),
                            _lib.f64(pos_v), _lib.i64(gd.n_c),
                            _lib.i64(gd.beg_c))  # arrays alive for the call
                    n = L.gh_spline_to_grid(
                        spline.l, spline.nbins, spline.dr,
                        spline.coefs.ctypes.data_as(_lib.PD),
                        args[0][1], args[1][1], args[2][1],
                        h_cv.ctypes.data_as(_lib.PD), args[3][1],
                        args[4][1], A.ctypes.data_as(_lib.PD),
                        G.ctypes.data_as(_lib.PI32))
                    _lib.check(0 if n >= 0 else int(n))
                    if n == 0:
                        continue
                    keys.append(w * np.int64(ngpts) + G[:n].astype(np.int64))
                    cols.append((Mj, A[:n]))
                Mj += nm
            if keys:
                allkeys = np.concatenate(keys)
                ukeys, inv = np.unique(allkeys, return_inverse=True)
                table = np.zeros((len(ukeys), nm_a))
                off = 0
                for k, (Mj0, A) in zip(keys, cols):
                    table[inv[off:off + len(k)], Mj0:Mj0 + A.shape[1]] = A
                    off += len(k)
                if self.integral is not None:
                    self._normalize(table, a, spline_j)
                G_parts.append((ukeys % ngpts).astype(np.int32))
                w_parts.append((ukeys // ngpts).astype(np.int32))
                A_parts.append(table.ravel())
                entry_ptr.append(entry_ptr[-1] + len(ukeys))
            else:
                entry_ptr.append(entry_ptr[-1])
            self.Mmax_a.append(nm_a)
            M += nm_a
        self.Mmax = M
        natoms = len(spos_ac)
        self.my_atom_indices = list(range(natoms))
        self.atom_indices = [a for a in range(natoms)
                             if entry_ptr[a + 1] > entry_ptr[a]]
        M_ptr = np.array(self.M_a + [M], np.int32)
        G_e = (np.concatenate(G_parts) if G_parts
               else np.zeros(0, np.int32))
        w_e = (np.concatenate(w_parts) if w_parts
               else np.zeros(0, np.int32))
        A = np.concatenate(A_parts) if A_parts else np.zeros(0)
        sdisp_Wc = np.array(sdisp_list, float).reshape(-1, 3)
        self.sdisp_wc = sdisp_Wc
        self.nentries = len(G_e)
        self.npoints_union = len(np.unique(G_e)) if len(G_e) else 0
        complex_ = is_complex(self.dtype)
        if complex_:
            # one table per k-point (phase_qW, lfc.py:339-340)
            phase_qw = np.exp(2j * pi * self.ibzk_qc.dot(sdisp_Wc.T))
        else:
            if not self.gamma and self.ibzk_qc.any():
                raise ValueError('real LFC at a non-Gamma k-point')
            phase_qw = None
        self._free()
        h = ctypes.c_void_p()
        keep = [_lib.i64(entry_ptr), _lib.i32(M_ptr), _lib.i32(G_e),
                _lib.i32(w_e), _lib.f64(A)]
        ph = (np.ascontiguousarray(phase_qw, complex)
              if phase_qw is not None else None)
        _lib.check(self._lib.gh_lfc_create(
            ctypes.byref(h), ngpts, natoms, keep[0][1], keep[1][1],
            keep[2][1], keep[3][1], keep[4][1], len(self.ibzk_qc),
            len(sdisp_list),
            ph.ctypes.data_as(_lib.PD) if ph is not None else None,
            float(self.gd.dv), dtype_code(self.dtype)))
        self.handle = h
        self.M_ptr = M_ptr

    def _normalize(self, table, a, spline_j):
        """Grid normalisation of one atom's functions (lfc.py:125-179): with
        I_M = dv sum over the atom's points of A_M, the l = 0 function is
        scaled to integrate to `integral` and every other function has that
        scaled function times I_M / integral subtracted, so it integrates
        to zero on the grid."""
        integral = self.integral
        if not np.isscalar(integral):
            integral = integral[a]
        if integral < 1e-15:
            return
        if spline_j[0].l != 0:
            raise ValueError('normalisation needs an l = 0 first function')
        I_m = self.gd.dv * table.sum(axis=0)
        table[:, 0] *= integral / I_m[0]
        if table.shape[1] > 1:
            table[:, 1:] -= np.outer(table[:, 0], I_m[1:] / integral)

    # ---------------------------------------------------------- helpers --
    def get_function_count(self, a):
        return self.Mmax_a[a]

    def dict(self, shape=(), derivative=False, zero=False):
        if derivative:
            raise NotImplementedError('derivatives are not provided')
        if isinstance(shape, int):
            shape = (shape,)
        td = torch_dtype(self.dtype)
        make = torch.zeros if zero else torch.empty
        return {a: make(tuple(shape) + (self.Mmax_a[a],), dtype=td,
                        device='cuda') for a in self.my_atom_indices}

    def M_to_ai(self, src_xM, dst_axi):
        for a in self.my_atom_indices:
            M1 = self.M_a[a]
            dst_axi[a] = src_xM[..., M1:M1 + self.Mmax_a[a]].clone()

    def ai_to_M(self, src_axi, dst_xM):
        for a in self.my_atom_indices:
            M1 = self.M_a[a]
            dst_xM[..., M1:M1 + self.Mmax_a[a]] = src_axi[a]

    # ------------------------------------------------------ fast paths --
    def _kpoint(self, q):
        """Phase-table index of k-point q (q = -1: the only k-point)."""
        nq = len(self.ibzk_qc)
        if q == -1:
            if nq > 1 and is_complex(self.dtype):
                raise ValueError('q must be given: the LFC has %d k-points'
                                 % nq)
            return 0
        if not 0 <= q < nq:
            raise ValueError('k-point index %d out of range (%d k-points)'
                             % (q, nq))
        return 0 if not is_complex(self.dtype) else int(q)

    def _check_grid(self, a_xG):
        if tuple(a_xG.shape[-3:]) != tuple(int(n) for n in self.gd.n_c):
            raise ValueError('array of shape %s does not live on the grid %s'
                             % (tuple(a_xG.shape), tuple(self.gd.n_c)))

    def integrate_M(self, a_xG, c_xM=None, q=-1):
        """c_xM[x, M] = dv * sum_G a(G) A_M(G) on the GPU."""
        self._check_grid(a_xG)
        xshape = tuple(a_xG.shape[:-3])
        nb = int(np.prod(xshape, dtype=np.int64))
        if c_xM is None:
            c_xM = torch.empty(xshape + (self.Mmax,),
                               dtype=torch_dtype(self.dtype), device='cuda')
        elif tuple(c_xM.shape) != xshape + (self.Mmax,):
            raise ValueError('coefficients of shape %s, expected %s'
                             % (tuple(c_xM.shape), xshape + (self.Mmax,)))
        kq = self._kpoint(q)
        with Staged(a_xG, dtype=self.dtype) as a, \
                Staged(c_xM, writeback=True, dtype=self.dtype) as c:
            _lib.check(self._lib.gh_lfc_integrate(
                self.handle, a.data_ptr(), c.data_ptr(), nb, kq,
                stream_ptr()))
        return c_xM

    def add_M(self, a_xG, c_xM, q=-1):
        """a(G) += sum_M c_xM[x, M] A_M(G) on the GPU."""
        self._check_grid(a_xG)
        xshape = tuple(a_xG.shape[:-3])
        if tuple(c_xM.shape) != xshape + (self.Mmax,):
            raise ValueError('coefficients of shape %s, expected %s'
                             % (tuple(c_xM.shape), xshape + (self.Mmax,)))
        nb = int(np.prod(xshape, dtype=np.int64))
        kq = self._kpoint(q)
        with Staged(c_xM, dtype=self.dtype) as c, \
                Staged(a_xG, writeback=True, dtype=self.dtype) as a:
            _lib.check(self._lib.gh_lfc_add(
                self.handle, c.data_ptr(), a.data_ptr(), nb, kq,
                stream_ptr()))

    # ---------------------------------------------- reference interface --
    def integrate(self, a_xG, c_axi, q=-1):
        """c_axi[a] = integral of a_xG times the functions of atom a
        (lfc.py:552-631)."""
        c_xM = self.integrate_M(a_xG, q=q)
        for a in self.my_atom_indices:
            if a in c_axi:
                M1 = self.M_a[a]
                src = c_xM[..., M1:M1 + self.Mmax_a[a]]
                dst = c_axi[a]
                if isinstance(dst, torch.Tensor):
                    dst.copy_(src)
                else:
                    dst[...] = src.cpu().numpy()

    def add(self, a_xG, c_axi=1.0, q=-1):
        """a_xG += sum_a,i c_axi[a][..., i] Phi_ai (lfc.py:392-472)."""
        xshape = tuple(a_xG.shape[:-3])
        td = torch_dtype(self.dtype)
        if isinstance(c_axi, float):
            assert len(xshape) == 0
            c_xM = torch.full((self.Mmax,), c_axi, dtype=td, device='cuda')
        else:
            c_xM = torch.zeros(xshape + (self.Mmax,), dtype=td, device='cuda')
            for a, c_xi in c_axi.items():
                M1 = self.M_a[a]
                c_xM[..., M1:M1 + self.Mmax_a[a]] = torch.as_tensor(
                    c_xi, device='cuda')
        self.add_M(a_xG, c_xM, q)


def LFC(gd, spline_aj, kd=None, cut=False, dtype=float, integral=None,
        forces=False, xp=None):
    return LocalizedFunctionsCollection(gd, spline_aj, kd, cut, dtype,
                                        integral, forces, xp)
