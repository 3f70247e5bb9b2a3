"""Radial cubic spline -- oracle restatement (test infrastructure only).

spline_coefs follows bmgs/spline.c:11-47 (tridiagonal second-derivative
recurrence with f2[0] = -0.5, four coefficients per bin, a zero bin appended);
spline_value follows bmgs/spline.c:50-58; Spline.from_data follows
spline.py:19-31 (last value forced to 0) and c/spline.c:151-165
(nbins = len(f) - 1, dr = rcut / nbins).

Pinned by tests/test_oracle_spline.py: exact interpolation at the knots,
C1 continuity between bins, against the reference C build
(cgpaw.Spline.__call__) and exact reproduction of a cubic polynomial whose
end conditions match the recurrence.
"""
import numpy as np


def spline_coefs(dr, f):
    f = np.asarray(f, float)
    nbins = len(f) - 1
    c = 3.0 / (dr * dr)
    f2 = np.empty(nbins + 1)
    u = np.empty(nbins)
    f2[0] = -0.5
    u[0] = (f[1] - f[0]) * c
    for b in range(1, nbins):
        p = 0.5 * f2[b - 1] + 2.0
        f2[b] = -0.5 / p
        u[b] = ((f[b + 1] - 2.0 * f[b] + f[b - 1]) * c - 0.5 * u[b - 1]) / p
    f2[nbins] = ((f[nbins - 1] * c - 0.5 * u[nbins - 1]) /
                 (0.5 * f2[nbins - 1] + 1.0))
    for b in range(nbins - 1, -1, -1):
        f2[b] = f2[b] * f2[b + 1] + u[b]
    data = np.zeros((nbins + 1, 4))
    for b in range(nbins):
        data[b, 0] = f[b]
        data[b, 1] = (f[b + 1] - f[b]) / dr - (f2[b] / 3 + f2[b + 1] / 6) * dr
        data[b, 2] = 0.5 * f2[b]
        data[b, 3] = (f2[b + 1] - f2[b]) / (6 * dr)
    return data


class Spline:
    def __init__(self, l, rcut, f_g):
        f_g = np.array(f_g, float)
        f_g[-1] = 0.0
        self.l = l
        self.rcut = float(rcut)
        self.nbins = len(f_g) - 1
        self.dr = self.rcut / self.nbins
        self.f_g = f_g
        self.data = spline_coefs(self.dr, f_g)

    def __call__(self, r):
        r = np.asarray(r, float)
        b = (r / self.dr).astype(int)
        inside = b < self.nbins
        b = np.where(inside, b, self.nbins)
        u = r - b * self.dr
        s = self.data[b]
        val = s[..., 0] + u * (s[..., 1] + u * (s[..., 2] + u * s[..., 3]))
        return np.where(inside, val, 0.0)

    def get_cutoff(self):
        return self.dr * self.nbins


Y_CONST = {0: 0.28209479177387814, 1: 0.48860251190291992}


def spherical_harmonics(l, f, x, y, z):
    """Real solid harmonics times f, l <= 2 (bmgs/spherical_harmonics.h:5-24).

    m order for l=1 is (y, z, x); for l=2 (xy, yz, 3z^2-r^2, xz, x^2-y^2)."""
    if l == 0:
        return (f * Y_CONST[0])[..., None]
    if l == 1:
        c = Y_CONST[1]
        return np.stack([f * c * y, f * c * z, f * c * x], axis=-1)
    if l == 2:
        r2 = x * x + y * y + z * z
        c1, c2, c3 = 1.0925484305920792, 0.31539156525252005, \
            0.54627421529603959
        return np.stack([f * c1 * x * y, f * c1 * y * z,
                         f * c2 * (-r2 + 3 * z * z), f * c1 * x * z,
                         f * c3 * (-y * y + x * x)], axis=-1)
    raise NotImplementedError('l > 2')
