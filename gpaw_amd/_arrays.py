"""Device-array plumbing between the drop-in classes and libgpawhip.

Arrays passed to the operators are either torch tensors already on the GPU
(the fast path: only their pointers cross the C ABI) or NumPy arrays, which
are copied to the GPU, processed there and copied back, so that callers
written against the reference's NumPy interface keep working.  There is no
host compute path.
"""
import numpy as np
import torch

from . import _lib

_TORCH = {np.dtype(np.float64): torch.float64,
          np.dtype(np.complex128): torch.complex128,
          np.dtype(np.float32): torch.float32,
          np.dtype(np.complex64): torch.complex64}

_CODE = {torch.float64: _lib.GH_F64, torch.complex128: _lib.GH_C128,
         torch.float32: _lib.GH_F32, torch.complex64: _lib.GH_C64}

_REAL = {torch.float64: torch.float64, torch.complex128: torch.float64,
         torch.float32: torch.float32, torch.complex64: torch.float32}


def torch_dtype(dtype):
    if isinstance(dtype, torch.dtype):
        return dtype
    return _TORCH[np.dtype(dtype)]


def dtype_code(dtype):
    return _CODE[torch_dtype(dtype)]


def real_dtype(dtype):
    return _REAL[torch_dtype(dtype)]


def is_complex(dtype):
    return torch_dtype(dtype) in (torch.complex128, torch.complex64)


def require_gpu():
    if not torch.cuda.is_available():
        raise _lib.GPAWAMDError('gpaw_amd needs a GPU (HIP device); none is '
                                'visible')


def stream_ptr():
    return torch.cuda.current_stream().cuda_stream


class Staged:
    """Context giving a device tensor view of an array argument.

    torch CUDA tensors pass through (must be contiguous); NumPy arrays are
    staged to the device and, if `writeback`, copied back on exit."""

    def __init__(self, a, writeback=False, dtype=None):
        self.orig = a
        self.writeback = writeback
        if isinstance(a, torch.Tensor):
            if not a.is_cuda:
                raise ValueError('tensor must live on the GPU')
            if not a.is_contiguous():
                raise ValueError('array must be C-contiguous')
            self.t = a
            self.host = False
        else:
            require_gpu()
            arr = np.asarray(a)
            if not arr.flags.c_contiguous:
                raise ValueError('array must be C-contiguous')
            self.t = torch.from_numpy(arr).to('cuda')
            self.host = True
        if dtype is not None and self.t.dtype != torch_dtype(dtype):
            raise ValueError('dtype %s does not match operator dtype %s'
                             % (self.t.dtype, torch_dtype(dtype)))

    def __enter__(self):
        return self.t

    def __exit__(self, *exc):
        if self.host and self.writeback and exc[0] is None:
            torch.cuda.current_stream().synchronize()
            self.orig[...] = self.t.cpu().numpy()
        return False


def ptr(t):
    return t.data_ptr() if t is not None else None
