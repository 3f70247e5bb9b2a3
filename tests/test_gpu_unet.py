"""The learned preconditioner's HIP U-Net layers against torch conv3d in
FP32 on the same BF16 values (weights and inputs rounded to BF16 as the
kernels take them): each 3x3x3 convolution mode (first layer from R and v,
plain, upsample-plus-skip), pooling and the 1x1 maps within 2 BF16 units
(only the FP32 summation order differs), and the whole network against
UNet3.forward(bf16=True) on periodic grids, cubic and not."""
import numpy as np
import pytest

torch = pytest.importorskip('torch')

pytestmark = pytest.mark.gpu

if not torch.cuda.is_available():
    pytest.skip('no GPU', allow_module_level=True)

import torch.nn.functional as F  # noqa: E402

from gpaw_amd import _lib  # noqa: E402
from gpaw_amd._arrays import stream_ptr  # noqa: E402
from gpaw_amd.learned.unet import HIPUNet, UNet3, VSCALE  # noqa: E402


def bf(t):
    return t.to(torch.bfloat16).to(torch.float32)


def to_cl(t):
    """(B, C, ...) float -> channel-last bf16 bits (int16)."""
    return bf(t).permute(0, 2, 3, 4, 1).contiguous().to(
        torch.bfloat16).view(torch.int16)


def from_cl(a, shape):
    B, C, n0, n1, n2 = shape
    return a.view(torch.bfloat16).reshape(B, n0, n1, n2, C).permute(
        0, 4, 1, 2, 3).to(torch.float32)


def conv_ref(x, W, b, relu=True):
    y = F.conv3d(F.pad(x, (1, 1, 1, 1, 1, 1), mode='circular'), bf(W), b)
    return F.relu(y) if relu else y


def close_bf16(got, ref, units=2, atol=1e-6):
    """Within `units` BF16 units of ref, or within atol -- the FP32 sums'
    own error, which decides the BF16 rounding (and the ReLU cut) of
    values close to a boundary or to zero."""
    tol = units * 2.0 ** -8 * ref.abs() + atol
    return bool(((got - ref).abs() <= tol).all())


def fp32_atol(ci, W, x):
    """A bound on the FP32 rounding of 27 ci products |w| |x| (u = 2^-24,
    times the number of terms)."""
    n = 27 * ci
    return n * 2.0 ** -24 * n * float(W.abs().max()) * float(x.abs().max())


@pytest.mark.parametrize('shape', [(24, 24, 24), (16, 20, 28)])
def test_conv_layers(shape):
    L = _lib.lib()
    g = torch.Generator(device='cuda')
    g.manual_seed(3)
    nb = 3
    n = np.array(shape, np.int64)
    half = tuple(s // 2 for s in shape)
    for ci, co, mode in ((8, 16, 1), (16, 32, 1), (32, 32, 1), (16, 16, 1),
                         (8, 8, 1), (16, 16, 2), (8, 8, 2)):
        W = torch.randn((co, ci, 3, 3, 3), device='cuda', generator=g) * 0.2
        b = torch.randn(co, device='cuda', generator=g) * 0.1
        packed = np.zeros(int(L.gh_unet_conv3_weight_elems(ci, co)), np.uint16)
        Wc = W.cpu().numpy().astype(np.float32)
        _lib.check(L.gh_unet_pack_conv3(ci, co, Wc.ctypes.data,
                                        packed.ctypes.data))
        Wd = torch.from_numpy(packed.view(np.int16)).cuda()
        x = bf(torch.randn((nb, ci) + shape, device='cuda', generator=g))
        out = torch.empty(nb * int(np.prod(shape)) * co, dtype=torch.int16,
                          device='cuda')
        if mode == 1:
            inp, skip = to_cl(x), None
            ref_in = x
        else:
            c = bf(torch.randn((nb, ci) + half, device='cuda', generator=g))
            inp, skip = to_cl(c), to_cl(x)
            ref_in = bf(F.interpolate(c, scale_factor=2, mode='nearest') + x)
        _lib.check(L.gh_unet_conv3(
            ci, co, mode, 0, nb, n.ctypes.data_as(_lib.PI64), inp.data_ptr(),
            skip.data_ptr() if skip is not None else None, None, None, None,
            0.0, Wd.data_ptr(), b.data_ptr(), 1, out.data_ptr(),
            stream_ptr()))
        got = from_cl(out, (nb, co) + shape)
        ref = bf(conv_ref(ref_in, W, b))
        atol = fp32_atol(ci, W, ref_in)
        assert close_bf16(got, ref, atol=atol), (ci, co, mode,
                                                 float((got - ref).abs().max()))


@pytest.mark.parametrize('dtype', [torch.float32, torch.float64])
def test_first_layer_and_network(dtype):
    torch.manual_seed(1)
    model = UNet3().cuda()
    net = HIPUNet(model)
    g = torch.Generator(device='cuda')
    g.manual_seed(5)
    for shape in ((24, 24, 24), (16, 20, 28)):
        nb = 3
        R = torch.randn((nb,) + shape, device='cuda', generator=g,
                        dtype=torch.float64).to(dtype)
        v = (-torch.rand(shape, device='cuda', generator=g,
                         dtype=torch.float64) * 2).to(dtype)
        scale = (1.0 / R.double().pow(2).mean(dim=(1, 2, 3)).sqrt()).to(
            torch.float32)
        out = torch.empty_like(R)
        net(R, v, scale, out)
        x = torch.stack([R.float() * scale[:, None, None, None],
                         (v.float() * VSCALE).expand(nb, *shape)], 1)
        with torch.no_grad():
            ref = model(x, bf16=True)[:, 0] / scale[:, None, None, None]
        err = (out.float() - ref).abs()
        # BF16 activations: a value at a rounding boundary may round the
        # other way and carry its 2^-8 difference on through the network
        assert float(err.max()) < 3e-2 * float(ref.abs().max())
        assert float(err.mean()) < 3e-3 * float(ref.abs().mean())
        # first layer alone, tight
        e0 = net._buf['e0'][:nb * int(np.prod(shape)) * 8]
        with torch.no_grad():
            r0 = bf(conv_ref(bf(x), model.c0.weight, model.c0.bias))
        assert close_bf16(from_cl(e0, (nb, 8) + shape), r0,
                          atol=fp32_atol(2, model.c0.weight, x))


def test_learned_preconditioner_runs_and_falls_back():
    """The learned preconditioner inside RMM-DIIS on a small cell: an
    untrained network's steps are judged by the residual norm, rejected,
    and the V-cycle takes over; the run then converges as the plain one."""
    from gpaw_amd import synthetic as syn
    from gpaw_amd.eigensolvers import RMMDIIS, FDWaveFunctions
    from gpaw_amd.grid_descriptor import GridDescriptor
    from gpaw_amd.learned import LearnedPreconditioner
    w = syn.full_hpsi_geometry(24, 4)
    gd = GridDescriptor((24,) * 3, (w.L,) * 3)
    vt = syn.potential(gd, w.pos_av)
    sysm = syn.System(w, gd, vt, np.float64)
    torch.manual_seed(0)
    model = UNet3()
    with torch.no_grad():
        for p in model.parameters():
            p.mul_(0.0)         # zero network: d = 0, rho = 1 -> rejected
    lp = LearnedPreconditioner(gd, sysm.kin, vt, model)
    psi0 = syn.lcao_start(gd, w, sysm.H, 12)
    wfs = FDWaveFunctions(gd, psi0, sysm.pt, lp)
    es = RMMDIIS()
    wts = [1.0] * 8 + [0.0] * 4
    errs = [es.iterate(sysm.H, wfs, wts) for _ in range(12)]
    assert lp.fallback_active and lp.rejected == 12 and lp.judged == 12
    assert errs[-1] < 1e-6 * errs[0]


def test_training_operator_matches_hamiltonian():
    """The torch (H - eps) used in the training loss equals the HIP
    residual operator of the same cell (FP32 storage of the pieces)."""
    from gpaw_amd import synthetic as syn
    from gpaw_amd.eigensolvers import RMMDIIS, FDWaveFunctions
    from gpaw_amd.grid_descriptor import GridDescriptor
    from gpaw_amd.learned import datagen
    data = datagen.generate(2, seed0=77, N=16, nbands=4)
    idx = torch.tensor([0, 5], device='cuda')
    f = data['y'][idx]
    got = datagen.apply_shifted(data, idx, f)
    for k, i in enumerate(idx.tolist()):
        c = int(data['cell'][i])
        w = datagen.cell_geometry(77 + c, 16)
        gd = GridDescriptor((16,) * 3, (w.L,) * 3)
        vt = datagen.cell_potential(gd, w)
        sysm = syn.System(w, gd, vt, np.float64)
        wfs = FDWaveFunctions(gd, f[k].double().contiguous(), sysm.pt,
                              sysm.pc)
        ref = torch.empty_like(wfs.psit)
        RMMDIIS().residual_of(sysm.H, wfs, wfs.psit,
                              data['eps'][i:i + 1].double(), ref)
        err = (got[k].double() - ref).abs().max() / ref.abs().max()
        assert float(err) < 1e-5


def test_learned_preconditioner_per_band_judge():
    """Bands whose network step leaves more than max_rho of |R|^2 take the
    V-cycle's step; the rest keep the network's.  With max_rho = 1 every
    band is kept, with max_rho = -1 every band is redone (and then the
    V-cycle takes over), and the iteration equals the plain V-cycle one."""
    from gpaw_amd import synthetic as syn
    from gpaw_amd.eigensolvers import RMMDIIS, FDWaveFunctions
    from gpaw_amd.grid_descriptor import GridDescriptor
    from gpaw_amd.learned import LearnedPreconditioner
    from gpaw_amd.learned.unet import load_model
    import os
    w = syn.full_hpsi_geometry(24, 4)
    gd = GridDescriptor((24,) * 3, (w.L,) * 3)
    vt = syn.potential(gd, w.pos_av)
    sysm = syn.System(w, gd, vt, np.float64)
    weights = os.path.join(os.path.dirname(os.path.dirname(
        os.path.abspath(__file__))), 'gpaw_amd', 'data', 'unet_precond.npz')
    model = load_model(weights)
    psi0 = syn.lcao_start(gd, w, sysm.H, 12)
    errs = {}
    for max_rho in (1.0, -1.0, None):
        pc = (sysm.pc if max_rho is None else
              LearnedPreconditioner(gd, sysm.kin, vt, model, max_rho=max_rho,
                                    fallback=sysm.pc))
        wfs = FDWaveFunctions(gd, psi0.clone(), sysm.pt, pc)
        es = RMMDIIS()
        errs[max_rho] = [es.iterate(sysm.H, wfs, [1.0] * 8 + [0.0] * 4)
                         for _ in range(6)]
        if max_rho == 1.0:
            assert pc.rejected == 0 and not pc.fallback_active
        if max_rho == -1.0:
            assert pc.fallback_active
    np.testing.assert_allclose(errs[-1.0], errs[None], rtol=1e-10)
    assert errs[1.0][-1] < errs[1.0][0]
