"""Pin oracle/density.py: charge conservation, linearity, plane waves."""
import numpy as np

from oracle import density as ode
from oracle import grid as og


def test_charge_of_orthonormal_bands():
    rng = np.random.default_rng(1)
    gd = og.Grid((8, 10, 12), (4.0, 5.0, 6.0))
    nb = 5
    a = rng.standard_normal((nb, 8 * 10 * 12)) + \
        1j * rng.standard_normal((nb, 8 * 10 * 12))
    q, _ = np.linalg.qr(a.T)                 # orthonormal columns
    psit = (q.T / gd.dv**0.5).reshape((nb, 8, 10, 12))
    f = rng.uniform(0, 2, nb)
    nt = ode.add_to_density(f, psit, gd.zeros())
    assert abs(nt.sum() * gd.dv - f.sum()) < 1e-12
    assert nt.min() >= 0.0


def test_linear_in_occupations_and_plane_wave():
    rng = np.random.default_rng(2)
    gd = og.Grid((6, 6, 6), (3.0, 3.0, 3.0))
    psit = rng.standard_normal((3, 6, 6, 6))
    f1, f2 = rng.uniform(0, 2, 3), rng.uniform(0, 2, 3)
    n1 = ode.add_to_density(f1, psit, gd.zeros())
    n2 = ode.add_to_density(f2, psit, gd.zeros())
    n12 = ode.add_to_density(f1 + f2, psit, gd.zeros())
    assert abs(n12 - n1 - n2).max() < 1e-13 * abs(n12).max()
    r = np.indices((6, 6, 6)).transpose(1, 2, 3, 0) * 0.5
    pw = np.exp(2j * np.pi * (r @ np.array([1, 0, 2])) / 3.0) / 27**0.5
    nt = ode.add_to_density([2.0], pw[None], gd.zeros())
    assert abs(nt - 2.0 / 27).max() < 1e-14
