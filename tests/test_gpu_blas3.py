"""BLAS-3 drop-ins (gpaw_amd.utilities.blas: mmm, rk, r2k) and
GridDescriptor.integrate against the BLAS routines the reference binds.
c/blas.c:48-170 hands its row-major arrays to dgemm / dsyrk / zherk /
dsyr2k / zher2k as their Fortran transposes with uplo 'u'; the expected
values here are scipy's BLAS wrappers called the same way (Fortran view of
each array, the reference's trans flag), so the conjugation order of r2k
with a complex alpha is pinned by the library, not restated.  Only the
lower triangle (the reference's defined part) is compared for rk / r2k."""
import numpy as np
import pytest
import scipy.linalg.blas as sb

torch = pytest.importorskip('torch')

pytestmark = pytest.mark.gpu

if not torch.cuda.is_available():
    pytest.skip('no GPU', allow_module_level=True)

from gpaw_amd.grid_descriptor import GridDescriptor  # noqa: E402
from gpaw_amd.utilities.blas import mmm, r2k, rk  # noqa: E402


def rnd(rng, shape, cplx):
    x = rng.standard_normal(shape)
    if cplx:
        x = x + 1j * rng.standard_normal(shape)
    return x


def lower(x):
    return np.tril(x)


def close(got, ref, tol=1e-12):
    return abs(got - ref).max() <= tol * max(abs(ref).max(), 1.0)


@pytest.mark.parametrize('cplx', [False, True])
@pytest.mark.parametrize('n,k', [(200, 4096), (7, 9)])
def test_rk(cplx, n, k):
    rng = np.random.default_rng(n + k)
    a = rnd(rng, (n, k), cplx)
    c0 = rnd(rng, (n, n), cplx)
    c0 = c0 + c0.conj().T
    f = sb.zherk if cplx else sb.dsyrk
    ref = f(0.7, a.T, beta=0.3, c=c0.T.copy(), trans=2 if cplx else 1,
            lower=0).T
    c = torch.from_numpy(c0.copy()).cuda()
    rk(0.7, torch.from_numpy(a).cuda(), 0.3, c)
    assert close(lower(c.cpu().numpy()), lower(ref))
    # trans='n': a^H a over the first axis
    refn = f(0.7, a.T.copy(), beta=0.0, trans=0, lower=0).T
    cn = torch.zeros((k, k), dtype=c.dtype, device='cuda')
    rk(0.7, torch.from_numpy(a).cuda(), 0.0, cn, trans='n')
    assert close(lower(cn.cpu().numpy()), lower(refn))


@pytest.mark.parametrize('cplx', [False, True])
@pytest.mark.parametrize('n,k', [(130, 6912), (5, 8)])
def test_r2k(cplx, n, k):
    rng = np.random.default_rng(3 * n + k)
    a = rnd(rng, (n, k), cplx)
    b = rnd(rng, (n, k), cplx)
    alpha = (0.6 - 0.4j) if cplx else 0.6
    if cplx:
        ref = sb.zher2k(alpha, a.T, b.T, trans=2, lower=0).T
    else:
        ref = sb.dsyr2k(alpha, a.T, b.T, trans=1, lower=0).T
    c = torch.zeros((n, n), dtype=torch.complex128 if cplx else
                    torch.float64, device='cuda')
    r2k(alpha, torch.from_numpy(a).cuda(), torch.from_numpy(b).cuda(), 0.0,
        c)
    assert close(lower(c.cpu().numpy()), lower(ref))


@pytest.mark.parametrize('cplx', [False, True])
def test_mmm(cplx):
    rng = np.random.default_rng(11)
    for opa, opb, sa, sb_ in (('N', 'N', (6, 5), (5, 4)),
                              ('T', 'N', (5, 6), (5, 4)),
                              ('N', 'C', (6, 5), (4, 5)),
                              ('C', 'T', (5, 6), (4, 5)),
                              ('N', 'C', (64, 8192), (48, 8192))):
        a, b = rnd(rng, sa, cplx), rnd(rng, sb_, cplx)
        A = a if opa == 'N' else a.T if opa == 'T' else a.conj().T
        B = b if opb == 'N' else b.T if opb == 'T' else b.conj().T
        c0 = rnd(rng, (A.shape[0], B.shape[1]), cplx)
        ref = 0.5 * A @ B + 2.0 * c0
        c = torch.from_numpy(c0.copy()).cuda()
        mmm(0.5, torch.from_numpy(a).cuda(), opa, torch.from_numpy(b).cuda(),
            opb, 2.0, c)
        assert close(c.cpu().numpy(), ref), (opa, opb)


@pytest.mark.parametrize('dtype', [np.float64, np.complex128, np.float32])
def test_grid_integrate(dtype):
    """grid_descriptor.py:275-326: one array, the rk / r2k / mmm forms."""
    gd = GridDescriptor((16, 16, 12), (5.0, 5.0, 4.0))
    rng = np.random.default_rng(5)
    cplx = np.dtype(dtype).kind == 'c'
    a = rnd(rng, (6, 16, 16, 12), cplx).astype(dtype)
    b = rnd(rng, (6, 16, 16, 12), cplx).astype(dtype)
    ta, tb = torch.from_numpy(a).cuda(), torch.from_numpy(b).cuda()
    tol = 1e-12 if dtype != np.float32 else 1e-5
    A = a.reshape(6, -1).astype(np.complex128 if cplx else np.float64)
    B = b.reshape(6, -1).astype(A.dtype)
    one = gd.integrate(ta)
    assert close(one.cpu().numpy(), A.sum(1) * gd.dv, tol)
    s = gd.integrate(ta, ta).cpu().numpy()
    ref = gd.dv * A.conj() @ A.T
    assert close(np.triu(s), np.triu(ref), tol)
    g = gd.integrate(ta, tb).cpu().numpy()
    assert close(g, gd.dv * A.conj() @ B.T, tol)
    h = gd.integrate(ta, tb, hermitian=True).cpu().numpy()
    G = gd.dv * A.conj() @ B.T
    assert close(h, 0.5 * (G + G.conj().T), tol)
    x = gd.integrate(ta[0], tb[0])
    assert abs(x - gd.dv * (A[0].conj() @ B[0])) < tol * abs(x) + 1e-12
