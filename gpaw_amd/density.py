"""Pseudo-density accumulation on the GPU.

add_to_density(f, psit_G, nt_G): drop-in for cgpaw.add_to_density
(c/utilities.c:313-356), nt_G += f |psit_G|^2 for one band.
add_to_density_bands(f_n, psit_nG, nt_G): all bands in one pass, as
FDWaveFunctions.add_to_density_from_k_point_with_occupation
(wavefunctions/fd.py:95-100) does band by band.  nt is FP64.
"""
import numpy as np
import torch

from . import _lib
from ._arrays import dtype_code, stream_ptr


def add_to_density_bands(f_n, psit_nG, nt_G):
    assert nt_G.dtype == torch.float64 and nt_G.is_contiguous()
    assert psit_nG.is_contiguous()
    assert tuple(psit_nG.shape[1:]) == tuple(nt_G.shape)
    nb = psit_nG.shape[0]
    f = torch.as_tensor(np.asarray(f_n, float) if not isinstance(
        f_n, torch.Tensor) else f_n, dtype=torch.float64,
        device='cuda').reshape(-1).contiguous()
    assert f.numel() == nb
    _lib.check(_lib.lib().gh_add_to_density(
        dtype_code(psit_nG.dtype), nb, nt_G.numel(), f.data_ptr(),
        psit_nG.data_ptr(), nt_G.data_ptr(), stream_ptr()))
    return nt_G


def add_to_density(f, psit_G, nt_G):
    return add_to_density_bands([float(f)], psit_G.contiguous()[None], nt_G)
