"""Pseudo-density accumulation -- oracle restatement (test infrastructure).

add_to_density follows c/utilities.c:313-356 as called band by band from
wavefunctions/fd.py:95-100: nt += f |psit|^2, bands added in order.
Pinned by tests/test_oracle_density.py against what the mathematics fixes:
the integral of the density of S-orthonormal (here plain orthonormal)
bands equals sum_n f_n, linearity in f, and the closed form |psi|^2 of a
plane wave (constant density 1/V).
"""
import numpy as np


def add_to_density(f_n, psit_nG, nt_G):
    for f, psit_G in zip(f_n, psit_nG):
        nt_G += f * (psit_G.conj() * psit_G).real
    return nt_G
