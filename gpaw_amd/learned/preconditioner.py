"""LearnedPreconditioner: the U-Net in place of the multigrid V-cycle
(gpaw/preconditioner.py:9-93, applied to the residuals at
gpaw/eigensolvers/rmmdiis.py:138-143), with the V-cycle as fallback.

__call__(R, kpt) returns d = f(R, v) per band (the U-Net of
gpaw_amd.learned.unet on the matrix cores).  The eigensolver judges each
band's preconditioned step by the residual norm: with dR = (H - eps S) d
and the line-search step lambda, rho = |R + lambda dR|^2 / |R|^2 is the
fraction of the residual left.  judge(rho) returns the bands whose rho
exceeds `max_rho`; the eigensolver redoes their step with the V-cycle, and
the later preconditioner calls of the same sweep give those bands the
V-cycle too.  When fewer than `min_keep` of the bands keep the network's
step, the V-cycle takes over for the rest of the run (`fallback_active`).
"""
import numpy as np
import torch

from .. import blas
from ..preconditioner import Preconditioner
from .unet import HIPUNet, load_model


class LearnedPreconditioner:
    def __init__(self, gd, kin, vt_G, weights, dtype=float, max_rho=0.6,
                 fallback=None, chunk=128, min_keep=0.1):
        if np.dtype(dtype).kind == 'c':
            raise NotImplementedError('the learned preconditioner is built '
                                      'for real wave functions')
        self.gd = gd
        self.dtype = dtype
        self.net = HIPUNet(load_model(weights) if isinstance(weights, str)
                           else weights)
        tdt = torch.float32 if np.dtype(dtype) == np.float32 else \
            torch.float64
        self.vt = torch.as_tensor(vt_G, device='cuda').to(tdt).contiguous()
        self.fallback = fallback if fallback is not None else \
            Preconditioner(gd, kin, dtype)
        self.max_rho = max_rho
        self.chunk = chunk
        self.fallback_active = False
        self.min_keep = min_keep
        self.reject = None
        self.calls = 0
        self.rejected = 0
        self.judged = 0

    def calculate_kinetic_energy(self, psit_xG, kpt):
        return None

    def reset(self):
        """Forget the judge's decisions (a new run from a new start)."""
        self.fallback_active = False
        self.reject = None
        self.calls = 0
        self.rejected = 0

    def __call__(self, residuals, kpt=None, ekin=None, out=None):
        if self.fallback_active:
            return self.fallback(residuals, kpt, ekin, out)
        self.calls += 1
        R = residuals if residuals.ndim == 4 else residuals[None]
        mask = self.reject
        if mask is not None and mask.numel() == R.shape[0] and \
                bool(mask.any()):
            # bands judged this sweep: the V-cycle for the rejected ones
            d = self._net(R)
            bad = mask.nonzero()[:, 0]
            d[bad] = self.fallback(R[bad].contiguous(), kpt)
            if out is None:
                return d if residuals.ndim == 4 else d[0]
            out.copy_(d if residuals.ndim == 4 else d[0])
            return out
        return self._net(residuals, out)

    def _net(self, residuals, out=None):
        R = residuals if residuals.ndim == 4 else residuals[None]
        nb = R.shape[0]
        n = R[0].numel()
        rms = torch.sqrt(blas.multi_dotc(R, R) / n)
        scale = (1.0 / torch.clamp(rms, min=1e-300)).to(torch.float32)
        if out is None:
            out = torch.empty_like(R)
        o = out if out.ndim == 4 else out[None]
        # bands in chunks: the activations of a chunk are ~20 bytes per
        # grid point and band
        for b0 in range(0, nb, self.chunk):
            b1 = min(nb, b0 + self.chunk)
            self.net(R[b0:b1], self.vt, scale[b0:b1].contiguous(), o[b0:b1])
        return out

    def judge(self, rho):
        """rho: per-band fraction of the residual norm^2 left after the
        line search along the preconditioned direction.  Returns the
        rejected bands (bool per band; None: all accepted)."""
        if self.fallback_active:
            return None
        rho = torch.nan_to_num(rho, nan=1.0)   # a zero step leaves all of R
        bad = rho > self.max_rho
        nbad = int(bad.sum())
        self.rejected += nbad
        self.judged += int(bad.numel())
        if bad.numel() - nbad < self.min_keep * bad.numel():
            self.fallback_active = True
            self.reject = None
            return torch.ones_like(bad)
        self.reject = bad if nbad else None
        return self.reject
