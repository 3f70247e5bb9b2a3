"""Finite-difference Hamiltonian applied to wave functions on the GPU.

H psi = T psi + v psi + sum_a |p_a> dH_a <p_a|psi>, as in the reference's
Hamiltonian.apply (gpaw/hamiltonian.py:445-479): kinetic stencil
(wfs.kin.apply), local potential (apply_local_potential :437-443, fused into
the stencil epilogue here), projections P = <p|psi> (pt.integrate), the
non-local matrices (apply_non_local_hamilton :23-44) and pt.add.
"""
import numpy as np
import torch

from . import _lib
from ._arrays import Staged, real_dtype, stream_ptr, torch_dtype


class FDHamiltonian:
    def __init__(self, gd, kin, vt_G, pt=None, dH_aii=None, dtype=float,
                 dS_aii=None):
        self.gd = gd
        self.kin = kin
        self.dtype = dtype
        vt64 = torch.as_tensor(np.asarray(vt_G) if not isinstance(
            vt_G, torch.Tensor) else vt_G, device='cuda').to(torch.float64)
        rdt = real_dtype(dtype)
        self.stencil_offset = 0.0
        if rdt == torch.float32:
            # The single-precision stencil's rounded coefficients do not sum
            # to zero: T_32 = T + delta on constants, a uniform shift of every
            # eigenvalue by delta (~1e-6 Ha for h = 0.2 A).  delta is known
            # exactly (the FP64 sum of the FP32 coefficients), and the centre
            # coefficient multiplies psi(g) as the potential does, so it is
            # removed through the potential before that is rounded.
            coef32 = np.asarray(kin.coef_p, np.float64).astype(np.float32)
            self.stencil_offset = float(coef32.astype(np.float64).sum())
            vt64 = vt64 - self.stencil_offset
        self.vt_G = vt64.to(rdt).contiguous()
        self.pt = pt
        self._lib = _lib.lib()
        self.dS_flat = None
        if pt is not None:
            assert dH_aii is not None
            self.dH_flat = self._blocks(dH_aii)
            self._dH_host = dH_aii
            if dS_aii is not None:
                self.dS_flat = self._blocks(dS_aii)

    def _blocks(self, X_aii):
        """Per-atom (nm_a, nm_a) matrices, packed atom after atom."""
        blocks = []
        for a in range(len(self.pt.Mmax_a)):
            X = np.asarray(X_aii[a], float)
            assert X.shape == (self.pt.Mmax_a[a],) * 2
            blocks.append(X.ravel())
        return torch.as_tensor(np.concatenate(blocks), device='cuda')

    def dH(self, P_xM, out=None, eps_x=None):
        """out[x, M_a + i] = sum_j P[x, M_a + j] dH_a[j, i]
        (- eps_x[x] sum_j P[x, M_a + j] dS_a[j, i] when eps_x is given and
        the Hamiltonian has overlap corrections dS)."""
        return self._apply_blocks(self.dH_flat, P_xM, out, eps_x,
                                  self.dS_flat)

    def dS(self, P_xM, out=None):
        """out[x, M_a + i] = sum_j P[x, M_a + j] dS_a[j, i] (None: dS = 0)."""
        if self.dS_flat is None:
            return None
        return self._apply_blocks(self.dS_flat, P_xM, out, None, None)

    def _apply_blocks(self, X_flat, P_xM, out, eps_x, dS_flat):
        P_xM = P_xM.contiguous()
        if out is None:
            out = torch.empty_like(P_xM)
        nb = int(np.prod(P_xM.shape[:-1], dtype=np.int64))
        eps = None
        if eps_x is not None and dS_flat is not None:
            eps = torch.as_tensor(eps_x, dtype=torch.float64,
                                  device='cuda').contiguous()
        _lib.check(self._lib.gh_lfc_apply_dH(
            self.pt.handle, P_xM.data_ptr(), X_flat.data_ptr(),
            eps.data_ptr() if eps is not None else None,
            dS_flat.data_ptr() if eps is not None else None,
            out.data_ptr(), nb, stream_ptr()))
        return out

    def apply_local(self, psit_xG, out_xG, phase_cd=None):
        """out = T psi + v psi (one fused stencil pass)."""
        self.kin.apply(psit_xG, out_xG, phase_cd, vt_g=self.vt_G)

    def apply_local_shifted(self, psit_xG, out_xG, eps_x, phase_cd=None):
        """out = (T + v - eps_x) psi per band (one fused stencil pass)."""
        self.kin.apply_shifted(psit_xG, out_xG, eps_x, phase_cd,
                               vt_g=self.vt_G)

    def apply(self, psit_xG, out_xG, phase_cd=None, P_xM=None):
        """out = H psi; returns the projections P = <p|psi>."""
        self.apply_local(psit_xG, out_xG, phase_cd)
        if self.pt is None:
            return None
        if P_xM is None:
            P_xM = self.pt.integrate_M(psit_xG)
        self.pt.add_M(out_xG, self.dH(P_xM))
        return P_xM
