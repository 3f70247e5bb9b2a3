"""gh_multi_gram_noalloc (the allocation-free variant of gh_multi_gram,
whose partial sums live in a fixed device array) against gh_multi_gram with
a caller workspace, called through the C ABI: two calls in a row, the
second smaller and of another dtype, must agree bit for bit (same kernels,
same summation order); a call beyond the fixed scratch is refused on the
host before anything is launched."""
import ctypes

import pytest

torch = pytest.importorskip('torch')

pytestmark = pytest.mark.gpu

if not torch.cuda.is_available():
    pytest.skip('no GPU', allow_module_level=True)

from gpaw_amd import _lib  # noqa: E402
from gpaw_amd._arrays import dtype_code, stream_ptr  # noqa: E402


def run_both(vecs, scale):
    L = _lib.lib()
    k = len(vecs)
    v0 = vecs[0]
    nb, n = v0.shape[0], v0[0].numel()
    cplx = v0.is_complex()
    shape = (nb, k, k, 2) if cplx else (nb, k, k)
    out_ref = torch.full(shape, float('nan'), dtype=torch.float64,
                         device='cuda')
    out_na = torch.full(shape, float('nan'), dtype=torch.float64,
                        device='cuda')
    work = torch.empty(int(L.gh_multi_gram_work(k, nb)), dtype=torch.float64,
                       device='cuda')
    ptrs = (_lib.P * k)(*[v.data_ptr() for v in vecs])
    code = dtype_code(v0.dtype)
    _lib.check(L.gh_multi_gram(code, nb, n, k, ptrs, scale,
                               out_ref.data_ptr(), work.data_ptr(),
                               stream_ptr()))
    _lib.check(L.gh_multi_gram_noalloc(code, nb, n, k, ptrs, scale,
                                       out_na.data_ptr(), stream_ptr()))
    torch.cuda.synchronize()
    return out_ref, out_na


def test_noalloc_matches_workspace_variant():
    g = torch.Generator(device='cuda')
    g.manual_seed(7)
    # first call: 300 bands, k = 3, float64
    a = [torch.randn((300, 24, 24, 24), dtype=torch.float64, device='cuda',
                     generator=g) for _ in range(3)]
    ref, na = run_both(a, 0.125)
    assert torch.isfinite(ref).all()
    assert torch.equal(ref, na)
    # second call, smaller and complex: 17 bands, k = 2
    b = [torch.randn((17, 16, 16, 16), dtype=torch.complex128, device='cuda',
                     generator=g) for _ in range(2)]
    ref, na = run_both(b, 2.0)
    assert torch.isfinite(ref).all()
    assert torch.equal(ref, na)
    # the band Gram matrix is Hermitian: G[x, 0, 1] = conj(G[x, 1, 0])
    assert torch.equal(na[:, 0, 1, 0], na[:, 1, 0, 0])
    assert torch.equal(na[:, 0, 1, 1], -na[:, 1, 0, 1])


def test_noalloc_refuses_oversize():
    L = _lib.lib()
    ptrs = (_lib.P * 6)(*([1] * 6))   # never dereferenced: refused first
    rc = L.gh_multi_gram_noalloc(_lib.GH_F64, 100000, 64, 6, ptrs, 1.0,
                                 ctypes.c_void_p(1), None)
    assert rc == -1
    assert b'scratch' in L.gh_last_error()
