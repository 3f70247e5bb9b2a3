"""Multigrid V-cycle preconditioner -- oracle restatement (test infrastructure).

Follows preconditioner.py:9-93 (the Python path, :74-93) which is the same
14-step sequence as fd_preconditioner.c:199-349:

  r1 = R0(-R); d1 = 4 s r1; q1 = L1 d1 - r1; r2 = R1(q1); d2 = 16 s r2;
  q2 = L2 d2 - r2; d2 -= 16 s q2; q1 = I2(d2); d1 -= q1; q1 = L1 d1 - r1;
  d1 -= 4 s q1; d0 = I1(-d1); q0 = L0 d0 - R; d0 = -(d0 - s q0)

with s = 0.66666666 / diag(L0), L1 = L2 = Laplace(gd_k, -0.5, n=1) and nn=1
transformers.

Pinned by tests/test_oracle_precond.py against the reference C build
(cgpaw.fd_precond) and, on a pure plane wave, the closed-form symbol of the
whole cycle built from the 1-D symbols of the primitives.
"""
import numpy as np

from .stencil import laplace_stencil, fd_apply
from .transformers import Transformer


class Preconditioner:
    def __init__(self, gd0, kin_coefs, kin_offsets):
        self.gd0 = gd0
        self.gd1 = gd1 = gd0.coarsen()
        self.gd2 = gd2 = gd1.coarsen()
        self.kin0 = (kin_coefs, kin_offsets)
        self.kin1 = laplace_stencil(gd1.h_cv, -0.5, 1)
        self.kin2 = laplace_stencil(gd2.h_cv, -0.5, 1)
        diag = kin_coefs[np.flatnonzero((kin_offsets == 0).all(1))[-1]]
        self.step = 0.66666666 / diag
        self.rest0 = Transformer(gd0, gd1, 1)
        self.rest1 = Transformer(gd1, gd2, 1)
        self.intp2 = Transformer(gd2, gd1, 1)
        self.intp1 = Transformer(gd1, gd0, 1)

    def __call__(self, R, phase_cd=None):
        pbc0, pbc1, pbc2 = self.gd0.pbc_c, self.gd1.pbc_c, self.gd2.pbc_c
        s = self.step

        def L(kin, a, pbc):
            return fd_apply(kin[0], kin[1], a, pbc, phase_cd)
        r1 = self.rest0.apply(-R, phase_cd)
        d1 = 4 * s * r1
        q1 = L(self.kin1, d1, pbc1)
        q1 -= r1
        r2 = self.rest1.apply(q1, phase_cd)
        d2 = 16 * s * r2
        q2 = L(self.kin2, d2, pbc2)
        q2 -= r2
        d2 -= 16 * s * q2
        q1 = self.intp2.apply(d2, phase_cd)
        d1 -= q1
        q1 = L(self.kin1, d1, pbc1)
        q1 -= r1
        d1 -= 4 * s * q1
        d0 = self.intp1.apply(-d1, phase_cd)
        q0 = L(self.kin0, d0, pbc0)
        q0 -= R
        d0 -= s * q0
        d0 *= -1.0
        return d0
