"""Radial spline (drop-in for gpaw.spline.Spline).

Spline.from_data(l, rmax, f_g) tabulates f on nbins+1 equidistant points of
[0, rmax] (last value forced to zero, spline.py:19-31); coefficients come
from gh_spline_coefs (the reference's bmgs/spline.c recurrence).  The
function carried is f(r); the localized function on the grid is
f(r) r^l Y_lm.
"""
import ctypes

import numpy as np

from . import _lib


class Spline:
    def __init__(self, l, rmax, f_g):
        f_g = np.array(f_g, float)
        f_g[-1] = 0.0
        self.l = int(l)
        self.rcut = float(rmax)
        self.nbins = len(f_g) - 1
        assert self.nbins >= 1 and self.rcut > 0
        self.dr = self.rcut / self.nbins
        self.f_g = f_g
        self.coefs = np.zeros(4 * (self.nbins + 1))
        L = _lib.lib()
        _lib.check(L.gh_spline_coefs(
            self.nbins, self.dr, f_g.ctypes.data_as(_lib.PD),
            self.coefs.ctypes.data_as(_lib.PD)))

    @classmethod
    def from_data(cls, l, rmax, f_g):
        assert rmax > 0.0
        return cls(l, rmax, f_g)

    def get_cutoff(self):
        return self.dr * self.nbins

    def get_angular_momentum_number(self):
        return self.l

    def get_npoints(self):
        return self.nbins + 1

    def __call__(self, r):
        return float(self.map(np.array([r], float))[0])

    def map(self, r_x):
        r_x = np.asarray(r_x, float)
        b = (r_x / self.dr).astype(int)
        inside = b < self.nbins
        b = np.where(inside, b, self.nbins)
        u = r_x - b * self.dr
        s = self.coefs.reshape(-1, 4)[b]
        v = s[..., 0] + u * (s[..., 1] + u * (s[..., 2] + u * s[..., 3]))
        return np.where(inside, v, 0.0)

    def __repr__(self):
        return 'Spline(l={}, rmax={:.2f}, ...)'.format(self.l,
                                                       self.get_cutoff())
