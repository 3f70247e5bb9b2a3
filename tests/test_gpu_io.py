"""Wave functions on the GPU through a .gpw file and back (run on an
MI355X): device tensors stream to the file band block by band block and
are read back onto the device (gpaw/new/ibzwfs.py:370-408,
gpaw/new/fd/builder.py:89-124)."""
import numpy as np
import pytest

torch = pytest.importorskip('torch')

pytestmark = pytest.mark.gpu

if not torch.cuda.is_available():
    pytest.skip('no GPU', allow_module_level=True)

from gpaw_amd.io import read_wave_functions, write_wave_functions  # noqa


@pytest.mark.parametrize('dtype', [torch.float64, torch.complex64])
def test_device_wave_functions_round_trip(tmp_path, dtype):
    g = torch.Generator(device='cuda')
    g.manual_seed(3)
    psit = [[torch.randn((40, 24, 24, 24), dtype=dtype, device='cuda',
                         generator=g) for k in range(2)]]
    f = tmp_path / 'wf.gpw'
    write_wave_functions(f, psit, block=16)
    for k in range(2):
        t = read_wave_functions(f, 0, k, device='cuda', block=16)
        assert t.is_cuda and t.dtype == dtype
        tol = 1e-6 if dtype == torch.complex64 else 1e-14
        assert torch.allclose(t, psit[0][k], rtol=tol, atol=0)
