"""Grid restriction / interpolation -- oracle restatement (test infrastructure).

transformer_pads follows transformers.py:18-56 (pads, neighbour pads, skips)
for one domain.  restrict follows bmgs/restrict.c:17-62,136-155 and
interpolate follows bmgs/interpolate.c:9-46,70-100: three separable 1-D passes
(axis 2, then 1, then 0) over the padded input of bc.c.

1-D weights (restrict.c:36-56 / interpolate.c:27-40), K = 2*nn:
    K=2: 1/2            K=4: 9/16, -1/16
    K=6: 75/128, -25/256, 3/256
    K=8: 1225/2048, -245/2048, 49/2048, -5/2048

Pinned by tests/test_oracle_transformers.py against the reference C build
(cgpaw.Transformer), integral conservation (test_transformations.py:10-51),
exact reproduction of polynomials of degree < K by interpolation, and
brute-force dense 1-D filter matrices.
"""
import numpy as np

from .stencil import pad

WEIGHTS = {2: [0.5],
           4: [0.5625, -0.0625],
           6: [0.58593750, -0.09765625, 0.01171875],
           8: [0.59814453125, -0.11962890625, 0.02392578125,
               -0.00244140625]}


def transformer_pads(gdin, gdout, nn):
    """(interpolate, pad_cd, skip_cd) as in transformers.py:27-47."""
    pad_cd = np.empty((3, 2), int)
    skip_cd = np.zeros((3, 2), int)
    if (gdin.N_c == 2 * gdout.N_c).all():
        pad_cd[:, 0] = 2 * nn - 1 - 2 * gdout.beg_c + gdin.beg_c
        pad_cd[:, 1] = 2 * nn - 2 + 2 * gdout.end_c - gdin.end_c
        interpolate = False
    else:
        assert (gdout.N_c == 2 * gdin.N_c).all()
        pad_cd[:, 0] = nn - 1 - gdout.beg_c // 2 + gdin.beg_c
        pad_cd[:, 1] = nn + gdout.end_c // 2 - gdin.end_c
        skip_cd[:, 0] = gdout.beg_c % 2
        skip_cd[:, 1] = gdout.end_c % 2
        interpolate = True
    assert (pad_cd >= 0).all()
    return interpolate, pad_cd, skip_cd


def _restrict_1d(a, K, axis):
    """One restrict pass along `axis` of a padded array (restrict.c:17-62)."""
    w = WEIGHTS[K]
    n = a.shape[axis]
    nout = (n - 2 * K + 3) // 2
    a = np.moveaxis(a, axis, -1)
    c = K - 1
    # same association as the C: 0.5 * (((a0 + w1*(a1+a-1)) + w3*(a3+a-3)) ...)
    acc = a[..., c:c + 2 * nout:2]
    for j, wj in enumerate(w):
        o = 2 * j + 1
        acc = acc + wj * (a[..., c + o:c + o + 2 * nout:2] +
                          a[..., c - o:c - o + 2 * nout:2])
    b = 0.5 * acc
    return np.moveaxis(b, -1, axis)


def _interpolate_1d(a, K, axis, skip):
    """One interpolate pass along `axis` (interpolate.c:9-46)."""
    w = WEIGHTS[K]
    nn = K // 2
    nrow = a.shape[axis]
    n = nrow - (K - 1) + skip[1]
    a = np.moveaxis(a, axis, -1)
    base = nn - 1
    even = a[..., base:base + n]
    odd = 0.0
    for t, wt in enumerate(w):
        odd = odd + wt * (a[..., base - t:base - t + n] +
                          a[..., base + 1 + t:base + 1 + t + n])
    out = np.empty(a.shape[:-1] + (2 * n,), np.result_type(a, float))
    out[..., 0::2] = even
    out[..., 1::2] = odd
    out = out[..., skip[0]:2 * n - skip[1]]
    return np.moveaxis(out, -1, axis)


def restrict(a_xg, nn, pad_cd, pbc_c, phase_cd=None):
    K = 2 * nn
    b = pad(a_xg, pad_cd, pbc_c, phase_cd)
    nx = a_xg.ndim - 3
    for axis in (2, 1, 0):
        b = _restrict_1d(b, K, nx + axis)
    return b


def interpolate(a_xg, nn, pad_cd, skip_cd, pbc_c, phase_cd=None):
    K = 2 * nn
    b = pad(a_xg, pad_cd, pbc_c, phase_cd)
    nx = a_xg.ndim - 3
    for axis in (2, 1, 0):
        b = _interpolate_1d(b, K, nx + axis, skip_cd[axis])
    return b


class Transformer:
    """Oracle counterpart of gpaw.transformers.Transformer for one domain."""

    def __init__(self, gdin, gdout, nn=1):
        self.gdin, self.gdout, self.nn = gdin, gdout, nn
        self.interpolate, self.pad_cd, self.skip_cd = transformer_pads(
            gdin, gdout, nn)

    def apply(self, a_xg, phase_cd=None):
        if self.interpolate:
            return interpolate(a_xg, self.nn, self.pad_cd, self.skip_cd,
                               self.gdin.pbc_c, phase_cd)
        return restrict(a_xg, self.nn, self.pad_cd, self.gdin.pbc_c, phase_cd)
