"""C ABI checks that need no GPU: the library loads, exports every function
include/gpaw_amd.h declares, and host-side validation returns the documented
error codes without touching the device."""
import ctypes
import os
import re

import numpy as np
import pytest

from gpaw_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, 'include', 'gpaw_amd.h')


def declared():
    text = open(HEADER).read()
    return sorted(set(re.findall(r'\b(gh_[a-z0-9_]+)\s*\(', text)))


@pytest.fixture(scope='module')
def L():
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip('libgpawhip.so not built')
    return _lib.lib()


def test_every_declared_symbol_exported(L):
    names = declared()
    assert len(names) >= 20
    for name in names:
        assert hasattr(L, name), name
    assert set(names) <= set(_lib._SIGNATURES), \
        set(names) - set(_lib._SIGNATURES)


def test_version(L):
    assert b'gfx950' in L.gh_version()


def test_stencil_create_validation(L):
    h = ctypes.c_void_p()
    coefs, cp = _lib.f64([1.0, 2.0])
    offs, op = _lib.i64([0, 0, 0, 0, 0, 9])
    n, npp = _lib.i64([8, 8, 8])
    pbc, pp = _lib.i32([1, 1, 1])
    # offset longer than the domain
    assert L.gh_stencil_create(ctypes.byref(h), 2, cp, op, npp, pp, 0) == -2
    assert b'longer' in L.gh_last_error()
    offs, op = _lib.i64([0, 0, 0, 0, 0, 1])
    assert L.gh_stencil_create(ctypes.byref(h), 2, cp, op, npp, pp, 9) == -1
    assert L.gh_stencil_create(ctypes.byref(h), 2, cp, op, npp, pp, 0) == 0
    assert L.gh_stencil_diagonal(h) == 1.0
    # in == out is rejected before any device work
    assert L.gh_stencil_apply(h, 4096, 4096, 1, None, None, None, 0.0,
                              None) == -1
    assert L.gh_stencil_destroy(h) == 0


def test_transformer_create_validation(L):
    h = ctypes.c_void_p()
    nin, a = _lib.i64([16, 16, 16])
    nout, b = _lib.i64([8, 8, 8])
    pad, p = _lib.i64([1, 0, 1, 0, 1, 0])
    skip, s = _lib.i64([0] * 6)
    pbc, q = _lib.i32([1, 1, 1])
    assert L.gh_transformer_create(ctypes.byref(h), a, b, 1, p, s, q, 0,
                                   0) == 0
    assert L.gh_transformer_destroy(h) == 0
    bad, bb = _lib.i64([9, 8, 8])
    assert L.gh_transformer_create(ctypes.byref(h), a, bb, 1, p, s, q, 0,
                                   0) == -2
    assert L.gh_transformer_create(ctypes.byref(h), a, b, 5, p, s, q, 0,
                                   0) == -1


def test_spline_host_functions_match_oracle():
    """gh_spline_coefs / gh_spline_to_grid are host code: check them
    against the oracle (itself pinned to the reference build)."""
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip('libgpawhip.so not built')
    from oracle.spline import Spline as OSpline
    from oracle.lfc import spline_to_grid as o_s2g
    from gpaw_amd.spline import Spline
    r = np.linspace(0, 2.0, 81)
    for l in (0, 1, 2):
        f = np.exp(-r**2 / 0.6)
        s = Spline.from_data(l, 2.0, f)
        o = OSpline(l, 2.0, f)
        np.testing.assert_array_equal(s.coefs.reshape(-1, 4), o.data)
        h_cv = np.diag([0.3, 0.32, 0.35])
        beg = np.array([0, 2, 5])
        end = np.array([9, 15, 16])
        n_c = np.array([20, 20, 20])
        pos = np.array([0.5, 2.6, 3.1])
        npts = int(np.prod(end - beg))
        A = np.empty((npts, 2 * l + 1))
        G = np.empty(npts, np.int32)
        args = (_lib.i64(beg), _lib.i64(end), _lib.f64(pos),
                _lib.f64(h_cv), _lib.i64(n_c), _lib.i64([0, 0, 0]))
        n = _lib.lib().gh_spline_to_grid(
            l, s.nbins, s.dr, s.coefs.ctypes.data_as(_lib.PD), args[0][1],
            args[1][1], args[2][1], args[3][1], args[4][1], args[5][1],
            A.ctypes.data_as(_lib.PD), G.ctypes.data_as(_lib.PI32))
        A_ref, G_B, G_ref = o_s2g(o, beg, end, pos, h_cv, n_c, [0, 0, 0])
        assert n == len(G_ref)
        np.testing.assert_array_equal(G[:n], G_ref)
        np.testing.assert_allclose(A[:n], A_ref, rtol=0, atol=1e-16)
