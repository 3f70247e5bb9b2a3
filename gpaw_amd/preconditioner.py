"""Multigrid V-cycle preconditioner (drop-in for gpaw.preconditioner).

Preconditioner(gd0, kin0, dtype, block) builds the same three-level cycle as
gpaw/preconditioner.py:9-33 (7-point Laplacians on the coarse levels, nn=1
transformers, step = 0.66666666 / diag(kin0)); __call__ runs the whole
14-step cycle of preconditioner.py:35-93 / c/fd_preconditioner.c:108-353 in
one gh_precond_apply call.
"""
import numpy as np
import torch

from . import _lib
from ._arrays import Staged, stream_ptr, torch_dtype
from .fd_operators import FDOperator, Laplace
from .transformers import Transformer


SINGLE = {np.dtype(float): np.float32, np.dtype(complex): np.complex64}


class Preconditioner:
    """precision='single' (an extension, not in the reference): for FP64
    data the whole cycle runs in FP32 / complex64 -- residuals converted on
    entry, the correction converted back -- while the caller's residuals,
    subspace algebra and orthonormalisation stay in FP64.  The V-cycle only
    supplies a search direction, so the converged states do not change."""

    def __init__(self, gd0, kin0, dtype=float, block=1, xp=None,
                 precision=None):
        gd1 = gd0.coarsen()
        gd2 = gd1.coarsen()
        self.gd0 = gd0
        self.kin0 = kin0
        self.step = 0.66666666 / kin0.get_diagonal_element()
        self.outer_dtype = dtype
        self.single = (precision == 'single' and
                       np.dtype(dtype) in SINGLE)
        if self.single:
            dtype = SINGLE[np.dtype(dtype)]
            kin0 = FDOperator(kin0.coef_p, kin0.offset_pc, gd0, dtype)
            self.kin0_single = kin0
        self.kin1 = Laplace(gd1, -0.5, 1, dtype, xp=xp)
        self.kin2 = Laplace(gd2, -0.5, 1, dtype, xp=xp)
        self.restrictor_object0 = Transformer(gd0, gd1, 1, dtype, xp=xp)
        self.restrictor_object1 = Transformer(gd1, gd2, 1, dtype, xp=xp)
        self.interpolator_object2 = Transformer(gd2, gd1, 1, dtype, xp=xp)
        self.interpolator_object1 = Transformer(gd1, gd0, 1, dtype, xp=xp)
        self.dtype = dtype
        self.block = block
        self.xp = xp
        self._scratch = None
        self._lib = _lib.lib()

    def calculate_kinetic_energy(self, psit_xG, kpt):
        return None

    def _kin0(self):
        return self.kin0_single if self.single else self.kin0

    def _scratch_for(self, nb):
        nbytes = self._lib.gh_precond_scratch_bytes(
            self._kin0().handle, self.kin1.handle, self.kin2.handle, nb)
        if self._scratch is None or self._scratch.numel() < nbytes:
            self._scratch = torch.empty(int(nbytes), dtype=torch.uint8,
                                        device='cuda')
        return self._scratch

    def __call__(self, residuals, kpt=None, ekin=None, out=None):
        """d0 = P(residuals); kpt supplies phase_cd (None: Gamma point)."""
        phases = getattr(kpt, 'phase_cd', None) if kpt is not None else None
        if residuals.ndim == 3:
            r = residuals[None]
            o = None if out is None else out[None]
            return self.__call__(r, kpt, ekin, o)[0]
        nb = residuals.shape[0]
        if out is None:
            out = self.gd0.empty(nb, self.outer_dtype)
            if not hasattr(residuals, 'data_ptr'):
                out = out.cpu().numpy()
        ph, php = _lib.phase_arg(phases)
        with Staged(residuals, dtype=self.outer_dtype) as R, \
                Staged(out, writeback=True, dtype=self.outer_dtype) as d0:
            scratch = self._scratch_for(nb)
            if self.single:
                r = R.to(torch_dtype(self.dtype))
                d = torch.empty_like(r)
            else:
                r, d = R, d0
            _lib.check(self._lib.gh_precond_apply(
                self._kin0().handle, self.kin1.handle, self.kin2.handle,
                self.restrictor_object0.handle,
                self.restrictor_object1.handle,
                self.interpolator_object2.handle,
                self.interpolator_object1.handle, float(self.step),
                r.data_ptr(), d.data_ptr(), nb, php, scratch.data_ptr(),
                stream_ptr()))
            if self.single:
                d0.copy_(d)
        return out
