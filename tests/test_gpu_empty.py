"""Empty band sets through every drop-in (run on an MI355X).

The reference's C loops run zero times for zero bands (operators.c,
transformers.c, lfc.c and blas.c all loop over the leading axis), so a call
with an empty leading axis is a valid no-op there.  Empty torch tensors carry
null data pointers; these tests check that such calls return the empty
result, touch no other memory, and raise nothing.
"""
import numpy as np
import pytest

torch = pytest.importorskip('torch')

pytestmark = pytest.mark.gpu

if not torch.cuda.is_available():
    pytest.skip('no GPU', allow_module_level=True)

from gpaw_amd import blas  # noqa: E402
from gpaw_amd.density import add_to_density_bands  # noqa: E402
from gpaw_amd.fd_operators import Laplace  # noqa: E402
from gpaw_amd.grid_descriptor import GridDescriptor  # noqa: E402
from gpaw_amd.lfc import LFC  # noqa: E402
from gpaw_amd.preconditioner import Preconditioner  # noqa: E402
from gpaw_amd.spline import Spline  # noqa: E402
from gpaw_amd.transformers import Transformer  # noqa: E402
from gpaw_amd.utilities.blas import mmm, r2k, rk  # noqa: E402

from oracle import grid as og  # noqa: E402

N_c = (16, 20, 24)
TD = {np.float64: torch.float64, np.complex128: torch.complex128,
      np.float32: torch.float32, np.complex64: torch.complex64}


def empty(shape, dtype):
    t = torch.empty(shape, dtype=TD[dtype], device='cuda')
    assert t.numel() == 0
    return t


@pytest.mark.parametrize('dtype', [np.float64, np.complex128, np.float32,
                                   np.complex64])
def test_stencil_empty(dtype):
    gd = GridDescriptor(N_c, np.array(N_c) * 0.3)
    op = Laplace(gd, -0.5, 3, dtype)
    ph = None
    if np.dtype(dtype).kind == 'c':
        ph = og.Grid(N_c, np.array(N_c) * 0.3).phase_cd((0.25, 0, 0))
    a = empty((0,) + N_c, dtype)
    out = empty((0,) + N_c, dtype)
    vt = torch.zeros(N_c, dtype=torch.float64 if dtype in (
        np.float64, np.complex128) else torch.float32, device='cuda')
    op.apply(a, out, ph, vt_g=vt)
    op.apply_shifted(a, out, torch.zeros(0, dtype=torch.float64,
                                         device='cuda'), ph, vt_g=vt)
    assert out.shape == (0,) + N_c


def test_generic_stencil_empty():
    """A non-orthogonal cell: plain apply, then the band-wise axpy."""
    cell_cv = 4.0 * np.array([(0, 1, 1), (1, 0, 1), (1, 1, 0)])
    gd = GridDescriptor((16, 16, 16), cell_cv)
    op = Laplace(gd, -0.5, 2, complex)
    assert not op.cfd
    a = empty((0, 16, 16, 16), np.complex128)
    out = empty((0, 16, 16, 16), np.complex128)
    op.apply_shifted(a, out, torch.zeros(0, dtype=torch.float64,
                                         device='cuda'))


@pytest.mark.parametrize('dtype', [np.float64, np.complex64])
def test_transformers_empty(dtype):
    gd0 = GridDescriptor(N_c, np.array(N_c) * 0.3)
    gd1 = gd0.coarsen()
    for nn in (1, 3):
        r = Transformer(gd0, gd1, nn, dtype).apply(empty((0,) + N_c, dtype))
        assert tuple(r.shape) == (0,) + tuple(gd1.n_c)
        i = Transformer(gd1, gd0, nn, dtype).apply(r)
        assert tuple(i.shape) == (0,) + N_c


@pytest.mark.parametrize('dtype', [np.float64, np.complex128])
def test_lfc_empty(dtype):
    gd = GridDescriptor(N_c, np.array(N_c) * 0.3)
    r = np.linspace(0, 2.0, 101)
    spl = [[Spline.from_data(l, 2.0, np.exp(-r**2)) for l in (0, 1)]
           for _ in range(2)]
    kd = None if dtype is np.float64 else np.array([(0.25, 0, 0)])
    lfc = LFC(gd, spl, kd, dtype=dtype)
    lfc.set_positions(np.array([(0.1, 0.2, 0.3), (0.6, 0.7, 0.8)]))
    a = empty((0,) + N_c, dtype)
    c = lfc.integrate_M(a)
    assert tuple(c.shape) == (0, lfc.Mmax)
    lfc.add_M(a, c)


@pytest.mark.parametrize('dtype', [np.float64, np.complex64])
def test_preconditioner_empty(dtype):
    gd = GridDescriptor(N_c, np.array(N_c) * 0.3)
    P = Preconditioner(gd, Laplace(gd, -0.5, 3, dtype), dtype)
    d = P(empty((0,) + N_c, dtype))
    assert tuple(d.shape) == (0,) + N_c


def test_grid_integrals_empty():
    gd = GridDescriptor(N_c, np.array(N_c) * 0.3)
    a = empty((0,) + N_c, np.float64)
    assert tuple(gd.integrate(a).shape) == (0,)
    assert tuple(gd.integrate(a, a).shape) == (0, 0)
    b = torch.ones((3,) + N_c, dtype=torch.float64, device='cuda')
    assert tuple(gd.integrate(a, b).shape) == (0, 3)


def test_subspace_blas_empty():
    a = empty((0, 50), np.float64)
    c = torch.empty((0, 0), dtype=torch.float64, device='cuda')
    rk(1.0, a, 0.0, c)
    r2k(0.5, a, a, 0.0, c)
    mmm(1.0, a, 'N', a, 'T', 0.0, c)
    b = torch.ones((4, 50), dtype=torch.float64, device='cuda')
    c2 = torch.empty((0, 4), dtype=torch.float64, device='cuda')
    mmm(1.0, a, 'N', b, 'T', 0.0, c2)


@pytest.mark.parametrize('dtype', [np.float64, np.complex64])
def test_band_blas_empty(dtype):
    """The per-band reductions and updates, and the density sum, with zero
    bands: the density must come back unchanged."""
    x = empty((0,) + N_c, dtype)
    y = empty((0,) + N_c, dtype)
    z0 = torch.zeros(0, dtype=torch.float64, device='cuda')
    blas.multi_axpy(z0, x, y)
    assert blas.multi_dotc(x, y).shape == (0,)
    assert tuple(blas.multi_gram([x, y]).shape) == (0, 2, 2)
    coef = torch.zeros((0, 2), dtype=torch.float64, device='cuda')
    assert blas.multi_lincomb([x, y], coef).shape == x.shape
    nt = torch.full(N_c, 0.5, dtype=torch.float64, device='cuda')
    add_to_density_bands(np.zeros(0), x, nt)
    torch.cuda.synchronize()
    assert bool((nt == 0.5).all())
