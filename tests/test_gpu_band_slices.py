"""Band sets larger than one launch takes (run on an MI355X).

The band-wise reductions and updates (mblas.py:13-32 and the DIIS sums of
rmmdiis.py) and the non-local P dH (hamiltonian.py:23-44) loop over any
number of bands in the reference.  Here each launch puts one band on a grid
row, so the entry points pass more than 32768 bands in slices; these tests
run 40000 bands (two slices, the second ragged) against torch FP64 on the
same data.
"""
import numpy as np
import pytest

torch = pytest.importorskip('torch')

pytestmark = pytest.mark.gpu

if not torch.cuda.is_available():
    pytest.skip('no GPU', allow_module_level=True)

from gpaw_amd import _lib, blas  # noqa: E402
from gpaw_amd._arrays import stream_ptr  # noqa: E402
from gpaw_amd.grid_descriptor import GridDescriptor  # noqa: E402
from gpaw_amd.lfc import LFC  # noqa: E402
from gpaw_amd.spline import Spline  # noqa: E402

NB = 40000


def randn(shape, dtype, seed):
    g = torch.Generator(device='cuda')
    g.manual_seed(seed)
    return torch.randn(shape, dtype=dtype, device='cuda', generator=g)


def rel(a, b):
    return float((a - b).abs().max() / b.abs().max())


@pytest.mark.parametrize('dtype', [torch.float64, torch.complex128])
def test_multi_dotc_slices(dtype):
    a = randn((NB, 3, 4, 5), dtype, 1)
    b = randn((NB, 3, 4, 5), dtype, 2)
    d = blas.multi_dotc(a, b, scale=0.5)
    ref = 0.5 * (a.conj() * b).reshape(NB, -1).sum(1).real
    assert rel(d, ref) < 1e-13


@pytest.mark.parametrize('dtype', [torch.float64, torch.complex64])
def test_multi_lincomb_slices(dtype):
    vs = [randn((NB, 2, 4, 8), dtype, 3 + i) for i in range(3)]
    c = randn((NB, 3), torch.complex128 if dtype.is_complex
              else torch.float64, 9)
    out = blas.multi_lincomb(vs, c)
    wide = torch.complex128 if dtype.is_complex else torch.float64
    ref = sum(c[:, i].reshape(-1, 1, 1, 1) * vs[i].to(wide) for i in range(3))
    assert rel(out.to(wide), ref) < (1e-6 if dtype == torch.complex64
                                     else 1e-14)


def test_multi_gram_slices():
    vs = [randn((NB, 64), torch.complex128, 11 + i) for i in range(2)]
    G = blas.multi_gram(vs, scale=0.25)
    V = torch.stack(vs)                                   # (2, NB, 64)
    ref = 0.25 * torch.einsum('ixg,jxg->xij', V.conj(), V)
    assert G.shape == (NB, 2, 2)
    assert rel(G, ref) < 1e-13


def test_dH_slices():
    N_c = (16, 16, 16)
    gd = GridDescriptor(N_c, np.array(N_c) * 0.3)
    r = np.linspace(0, 1.5, 101)
    spl = [[Spline.from_data(l, 1.5, np.exp(-r**2)) for l in (0, 1)]
           for _ in range(2)]
    lfc = LFC(gd, spl)
    lfc.set_positions(np.array([(0.2, 0.3, 0.4), (0.7, 0.6, 0.5)]))
    Mmax = lfc.Mmax
    rng = np.random.default_rng(5)
    dH = [rng.standard_normal((4, 4)) for _ in range(2)]
    dS = [rng.standard_normal((4, 4)) for _ in range(2)]
    dH = [0.5 * (h + h.T) for h in dH]
    dS = [0.5 * (s + s.T) for s in dS]
    dH_flat = torch.from_numpy(np.concatenate([h.ravel() for h in dH])).cuda()
    dS_flat = torch.from_numpy(np.concatenate([s.ravel() for s in dS])).cuda()
    P = randn((NB, Mmax), torch.float64, 21)
    eps = randn((NB,), torch.float64, 22)
    out = torch.empty_like(P)
    _lib.check(_lib.lib().gh_lfc_apply_dH(
        lfc.handle, P.data_ptr(), dH_flat.data_ptr(), eps.data_ptr(),
        dS_flat.data_ptr(), out.data_ptr(), NB, stream_ptr()))
    ref = torch.empty_like(P)
    for a in range(2):
        M0 = int(lfc.M_a[a])
        Pa = P[:, M0:M0 + 4]
        ref[:, M0:M0 + 4] = (Pa @ torch.from_numpy(dH[a]).cuda()
                             - eps[:, None] * (Pa @ torch.from_numpy(
                                 dS[a]).cuda()))
    assert rel(out, ref) < 1e-13
