"""Pseudo-density accumulation on the GPU.

add_to_density(f, psit_G, nt_G): drop-in for cgpaw.add_to_density
(c/utilities.c:313-356), nt_G += f |psit_G|^2 for one band.
add_to_density_bands(f_n, psit_nG, nt_G): all bands in one pass, as
FDWaveFunctions.add_to_density_from_k_point_with_occupation
(wavefunctions/fd.py:95-100) does band by band.  nt is FP64.

The steps that follow on the fine grid (RealSpaceDensity,
gpaw/density.py:686-766, and RealSpaceHamiltonian.restrict_and_collect,
gpaw/hamiltonian.py:692-698) for one domain: FineGrids.interpolate /
interpolate_pseudo_density (nn = 3 interpolation coarse -> fine, with the
charge rescaling of non-periodic cells), ghat / calculate_pseudo_charge
(compensation charges: Gaussian shape functions r^l exp(-(r/rc)^2) per l,
normalised on the fine grid with integral sqrt(4 pi), added with the
multipole moments Q_aL) and restrict_and_collect (nn = 3 restriction fine
-> coarse).
"""
from math import gamma, pi, sqrt
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


def gauss_shape_functions(rc, lmax, rcut, npts=201):
    """Radial parts f_l(r) of the Gaussian compensation-charge shape
    functions g_l(r) = f_l(r) r^l (gpaw/atom/shapefunc.py 'gauss'):
    f_l = c_l exp(-(r/rc)^2) with c_l fixing int_0^inf g_l r^l r^2 dr = 1,
    i.e. c_l = 2 / (rc^(2l+3) Gamma(l + 3/2)).  Returns (r, [f_l]) on npts
    points up to rcut (f = 0 at rcut)."""
    r = np.linspace(0, rcut, npts)
    out = []
    for l in range(lmax + 1):
        c = 2.0 / (rc**(2 * l + 3) * gamma(l + 1.5))
        f = c * np.exp(-(r / rc)**2)
        f[-1] = 0.0
        out.append(f)
    return r, out


class FineGrids:
    """Coarse / fine grid pair of the density and the potential (one
    domain): the nn = 3 transformers (density.py:701, hamiltonian.py:682)
    and the compensation charges ghat (density.py:712-713)."""

    def __init__(self, gd, finegd, stencil=3, dtype=float):
        from .transformers import Transformer
        self.gd = gd
        self.finegd = finegd
        self.interpolator = Transformer(gd, finegd, stencil, dtype)
        self.restrictor = Transformer(finegd, gd, stencil, dtype)
        self.ghat = None

    def set_compensation_charges(self, spos_ac, rc_a, lmax_a, rcut_a=None):
        """ghat = LFC(finegd, [ghat_l per atom], integral=sqrt(4 pi)) with
        Gaussian shape functions of width rc_a[a] and l <= lmax_a[a]."""
        from .lfc import LFC
        from .spline import Spline
        spl = []
        for a, (rc, lmax) in enumerate(zip(rc_a, lmax_a)):
            rcut = rcut_a[a] if rcut_a is not None else 5.0 * rc
            r, f_l = gauss_shape_functions(rc, lmax, rcut)
            spl.append([Spline.from_data(l, rcut, f) for l, f in
                        enumerate(f_l)])
        self.ghat = LFC(self.finegd, spl, integral=sqrt(4 * pi))
        self.ghat.set_positions(spos_ac)

    def interpolate(self, in_xR, out_xR=None):
        """Interpolate array(s) to the fine grid (density.py:733-748)."""
        if out_xR is None:
            out_xR = self.finegd.empty(tuple(in_xR.shape[:-3]))
        a = in_xR.reshape((-1,) + tuple(in_xR.shape[-3:]))
        b = out_xR.reshape((-1,) + tuple(out_xR.shape[-3:]))
        self.interpolator.apply(a, b)
        return out_xR

    def interpolate_pseudo_density(self, nt_sG, nt_sg=None,
                                   pseudo_charge=None):
        """nt_sg = interpolated nt_sG (density.py:715-731); with a
        non-periodic cell and the expected pseudo charge given, nt_sg is
        rescaled so that it integrates to it."""
        nt_sg = self.interpolate(nt_sG, nt_sg)
        if not self.gd.pbc_c.all() and pseudo_charge is not None:
            if abs(pseudo_charge) > 1.0e-14:
                x = pseudo_charge / float(self.finegd.integrate(nt_sg).sum())
                nt_sg *= x
        return nt_sg

    def calculate_pseudo_charge(self, nt_sg, Q_aL):
        """rhot_g = sum_s nt_sg + sum_aL Q_aL ghat_aL (density.py:756-760;
        no background charge)."""
        rhot_g = nt_sg.sum(dim=0).contiguous()
        self.ghat.add(rhot_g, Q_aL)
        return rhot_g

    def restrict_and_collect(self, a_xg, b_xg=None, phases=None):
        """b_xG = restriction of a_xg (hamiltonian.py:692-698, one
        domain)."""
        return self.restrictor.apply(a_xg, b_xg, phases)
