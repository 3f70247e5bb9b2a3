"""Loader for the reference C kernels built by oracle/Makefile (test only).

``oracle/_ref/libgpawref.so`` is GPAW's own c/bmgs, c/bc.c, c/operators.c,
c/transformers.c, c/fd_preconditioner.c, c/spline.c and c/lfc.c compiled from
the reference checkout (never copied), plus ``ref_glue.c``.  The functions
here build the reference's Python-level C objects exactly as the reference's
Python layer does:

* Operator     -- fd_operators.py:83-86 -> NewOperatorObject (operators.c:384)
* Transformer  -- transformers.py:68-76 -> NewTransformerObject (:202)
* Spline       -- spline.py:31 -> NewSplineObject (spline.c:151)
* spline_to_grid -- lfc.py:98-103 -> lfc.c:887
* LFC          -- lfc.py:368-369 -> NewLFCObject (lfc.c:146)
* fd_precond   -- preconditioner.py:66-71 -> fd_preconditioner.c:108

The library is built in this container from /root/reference and travels to
other machines as a prebuilt file; nothing here reads the reference sources
at run time.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, '_ref', 'libgpawref.so')
REFERENCE_C = '/root/reference/c'

_lib = None


def build():
    """Compile the reference kernels when the reference checkout exists."""
    if os.path.isdir(REFERENCE_C):
        subprocess.run(['make', '-s', '-C', HERE], check=True,
                       stdout=subprocess.DEVNULL)


def available():
    return os.path.exists(LIB)


def lib():
    global _lib
    if _lib is None:
        if not available():
            raise RuntimeError('reference build missing: ' + LIB)
        L = ctypes.PyDLL(LIB)
        L.gpawref_init.restype = ctypes.c_int
        rc = L.gpawref_init()
        if rc != 0:
            raise RuntimeError('gpawref_init failed: %d' % rc)
        for name in ['NewOperatorObject', 'NewTransformerObject',
                     'NewSplineObject', 'NewLFCObject', 'spline_to_grid',
                     'fd_precond']:
            f = getattr(L, name)
            f.restype = ctypes.py_object
            f.argtypes = [ctypes.py_object, ctypes.py_object]
        _lib = L
    return _lib


def _long(a):
    return np.ascontiguousarray(a, np.int64)


def Operator(coef_p, offset_pc, gd, real=True):
    """cgpaw.Operator for a serial grid (fd_operators.py:37-86)."""
    offset_pc = np.asarray(offset_pc)
    cfd = all(sum(o != 0 for o in off) < 2 for off in offset_pc)
    mp = int(np.abs(offset_pc).max())
    M_c = gd.n_c + 2 * mp
    stride_c = np.array([M_c[1] * M_c[2], M_c[2], 1])
    offset_p = _long(np.dot(offset_pc, stride_c))
    coef_p = np.ascontiguousarray(coef_p, float)
    args = (coef_p, offset_p, _long(gd.n_c), mp, _long(gd.neighbor_cd),
            int(real), None, int(cfd), 0)
    # bmgs_stencil and bc_init copy what they need (stencils.c:15-26)
    return lib().NewOperatorObject(None, args)


def Transformer(gdin, gdout, nn, real=True):
    """cgpaw.Transformer for serial grids (transformers.py:18-76)."""
    pad_cd = np.empty((3, 2), int)
    npad_cd = np.empty((3, 2), int)
    skip_cd = np.zeros((3, 2), int)
    if (gdin.N_c == 2 * gdout.N_c).all():
        pad_cd[:, 0] = 2 * nn - 1 - 2 * gdout.beg_c + gdin.beg_c
        pad_cd[:, 1] = 2 * nn - 2 + 2 * gdout.end_c - gdin.end_c
        npad_cd[:, 0] = 2 * nn - 2 + 2 * gdout.beg_c - gdin.beg_c
        npad_cd[:, 1] = 2 * nn - 1 - 2 * gdout.end_c + gdin.end_c
        interpolate = False
    else:
        pad_cd[:, 0] = nn - 1 - gdout.beg_c // 2 + gdin.beg_c
        pad_cd[:, 1] = nn + gdout.end_c // 2 - gdin.end_c
        npad_cd[:, 0] = nn + gdout.beg_c // 2 - gdin.beg_c
        npad_cd[:, 1] = nn - 1 - gdout.end_c // 2 + gdin.end_c
        skip_cd[:, 0] = gdout.beg_c % 2
        skip_cd[:, 1] = gdout.end_c % 2
        interpolate = True
    args = (_long(gdin.n_c), _long(gdout.n_c), 2 * nn, _long(pad_cd),
            _long(npad_cd), _long(skip_cd), _long(gdin.neighbor_cd),
            int(real), None, int(interpolate), 0)
    t = lib().NewTransformerObject(None, args)
    return t


def Spline(l, rcut, f_g):
    f_g = np.array(f_g, float)
    f_g[-1] = 0.0
    return lib().NewSplineObject(None, (int(l), float(rcut), f_g))


def spline_to_grid(spline, beg_c, end_c, pos_v, h_cv, n_c, gdcorner_c):
    return lib().spline_to_grid(None, (
        spline, _long(beg_c), _long(end_c),
        np.ascontiguousarray(pos_v, float), np.ascontiguousarray(h_cv, float),
        _long(n_c), _long(gdcorner_c)))


class LFC:
    """Reference LFC object for serial grids, built as LFC._update does
    (lfc.py:273-369): volumes per (atom, spline, box), event list G_B/W_B
    merge-sorted by G, phase_qW from sdisp."""

    def __init__(self, gd, spline_aj, spos_ac, k_c=None, dtype=float,
                 cut=False):
        self.gd = gd
        A_Wgm, M_W, G_wb, sdisp_W = [], [], [], []
        M_a = []
        M = 0
        self.splines = []
        for spos_c, spline_j in zip(np.asarray(spos_ac, float), spline_aj):
            M_a.append(M)
            Ma = 0
            for sp in spline_j:
                s = Spline(sp.l, sp.rcut, sp.f_g)
                self.splines.append(s)
                rcut = s.get_cutoff()
                for beg_c, end_c, sdisp_c in gd.get_boxes(spos_c, rcut,
                                                          cut=cut):
                    pos_v = np.dot(spos_c - sdisp_c, gd.cell_cv)
                    A_gm, G_b = spline_to_grid(s, beg_c, end_c, pos_v,
                                               gd.h_cv, gd.n_c, gd.beg_c)
                    if len(G_b) > 0:
                        A_Wgm.append(A_gm)
                        M_W.append(M + Ma)
                        G_wb.append(G_b)
                        sdisp_W.append(sdisp_c)
                Ma += 2 * sp.l + 1
            M += Ma
        self.Mmax = M
        self.M_a = M_a
        nW = len(A_Wgm)
        nB = sum(len(g) for g in G_wb)
        G_B = np.empty(nB, np.intc)
        W_B = np.empty(nB, np.intc)
        B1 = 0
        for W, G_b in enumerate(G_wb):
            B2 = B1 + len(G_b)
            G_B[B1:B2] = G_b
            W_B[B1:B2:2] = W
            W_B[B1 + 1:B2 + 1:2] = -W - 1
            B1 = B2
        if k_c is None:
            if dtype == float:
                phase_qW = np.empty((0, nW), complex)
            else:
                phase_qW = np.ones((1, nW), complex)
        else:
            phase_qW = np.exp(2j * np.pi * np.inner(
                np.asarray(k_c, float)[None], np.array(sdisp_W, float)))
        idx = np.argsort(G_B, kind='mergesort')
        self.G_B = np.ascontiguousarray(G_B[idx])
        self.W_B = np.ascontiguousarray(W_B[idx])
        self.A_Wgm = A_Wgm
        self.M_W = np.array(M_W, np.intc)
        self.phase_qW = np.ascontiguousarray(phase_qW)
        self.lfc = lib().NewLFCObject(None, (
            self.A_Wgm, self.M_W, self.G_B, self.W_B, float(gd.dv),
            self.phase_qW, 0))
        self.q = -1 if k_c is None and dtype == float else 0

    def integrate(self, a_xG):
        xshape = a_xG.shape[:-3]
        c_xM = np.zeros(xshape + (self.Mmax,), a_xG.dtype)
        self.lfc.integrate(np.ascontiguousarray(a_xG), c_xM, self.q)
        return c_xM

    def add(self, a_xG, c_xM):
        self.lfc.add(np.ascontiguousarray(c_xM), a_xG, self.q)
        return a_xG


def fd_precond(gd0, kin_coefs, kin_offsets, R_xG, phase_cd=None):
    """Run the reference C V-cycle (preconditioner.py:59-72)."""
    from .stencil import laplace_stencil
    real = not np.iscomplexobj(R_xG)
    dtype = float if real else complex
    gd1 = gd0.coarsen()
    gd2 = gd1.coarsen()
    kin0 = Operator(kin_coefs, kin_offsets, gd0, real)
    c1, o1 = laplace_stencil(gd1.h_cv, -0.5, 1)
    c2, o2 = laplace_stencil(gd2.h_cv, -0.5, 1)
    kin1 = Operator(c1, o1, gd1, real)
    kin2 = Operator(c2, o2, gd2, real)
    rest0 = Transformer(gd0, gd1, 1, real)
    rest1 = Transformer(gd1, gd2, 1, real)
    intp2 = Transformer(gd2, gd1, 1, real)
    intp1 = Transformer(gd1, gd0, 1, real)
    R = np.ascontiguousarray(R_xG)
    if R.ndim == 3:
        R = R[None]
    nb = len(R)
    d0 = gd0.zeros(nb, dtype)
    q0 = gd0.zeros(nb, dtype)
    r1, d1, q1 = [gd1.zeros(nb, dtype) for _ in range(3)]
    r2, d2, q2 = [gd2.zeros(nb, dtype) for _ in range(3)]
    diag = kin_coefs[np.flatnonzero((np.asarray(kin_offsets) == 0)
                                    .all(1))[-1]]
    step = 0.66666666 / diag
    ph = (np.ones((3, 2), complex) if phase_cd is None
          else np.ascontiguousarray(phase_cd, complex))
    lib().fd_precond(None, (rest0, rest1, intp1, intp2, kin0, kin1, kin2,
                            d0, q0, r1, d1, q1, r2, d2, q2, R,
                            np.ascontiguousarray(-R), float(step), ph))
    return d0 if R_xG.ndim == 4 else d0[0]
