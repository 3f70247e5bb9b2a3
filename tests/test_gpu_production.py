"""Parity at the benchmark's own shapes (run on an MI355X): 256 real FP64
bands on the 112^3 grid with 64 atoms x 8 projectors, built by
bench.build_workload -- the launch shapes the benchmark times (the
point-major LFC add in 16-band tiles, integrate tiles of 16 bands, the LDS-ring stencil with band
groups and the XCD-ordered block map, the V-cycle at 256 bands) compared
with the oracle on sampled bands: H psi (gpaw/hamiltonian.py:445-479), the
residual (H - eps S) psi (eigensolver.py:153-172) and the preconditioner
(preconditioner.py:35-93); and the single-precision and Bloch (k = (0.25, 0,
0)) forms at the same shape."""
import numpy as np
import pytest

torch = pytest.importorskip('torch')

pytestmark = pytest.mark.gpu

if not torch.cuda.is_available():
    pytest.skip('no GPU', allow_module_level=True)

from oracle import grid as og  # noqa: E402
from oracle import lfc as olfc  # noqa: E402
from oracle import precond as opc  # noqa: E402
from oracle import rmmdiis as orm  # noqa: E402
from oracle import spline as osp  # noqa: E402
from oracle import stencil as ost  # noqa: E402

SAMPLES = [0, 131, 255]


def relerr(a, b):
    return abs(a - b).max() / abs(b).max()


@pytest.fixture(scope='module')
def workload():
    import bench
    w, gd, pt, vt, kin, H, psi, out = bench.build_workload(112, 64, 256)
    ogd = og.Grid((112,) * 3, (w.L,) * 3)
    opt = olfc.LFC(ogd, [[osp.Spline(l, rc, f) for l, rc, f in p]
                         for p in w.proj])
    opt.set_positions(w.spos_ac)
    c, o = ost.laplace_stencil(ogd.h_cv, -0.5, 3)
    oham = orm.Hamiltonian(ogd, (c, o), vt, opt, w.dH)
    yield w, gd, pt, kin, H, psi, out, oham, ogd
    del psi, out
    torch.cuda.empty_cache()


def test_full_hpsi_256_bands(workload):
    w, gd, pt, kin, H, psi, out, oham, ogd = workload
    assert pt.Mmax == 512
    P = H.apply(psi, out)
    torch.cuda.synchronize()
    x = psi[SAMPLES].cpu().numpy()
    ref_P = oham.projections(x)
    assert relerr(P[SAMPLES].cpu().numpy(), ref_P) < 1e-12
    ref = oham.apply(x)
    assert relerr(out[SAMPLES].cpu().numpy(), ref) < 1e-12


def test_residual_256_bands(workload):
    w, gd, pt, kin, H, psi, out, oham, ogd = workload
    eps = torch.linspace(-1.5, 0.5, 256, dtype=torch.float64, device='cuda')
    H.apply_local_shifted(psi, out, eps)
    P = pt.integrate_M(psi)
    pt.add_M(out, H.dH(P, eps_x=eps))
    torch.cuda.synchronize()
    x = psi[SAMPLES].cpu().numpy()
    ref = oham.residual(x, eps[SAMPLES].cpu().numpy())
    assert relerr(out[SAMPLES].cpu().numpy(), ref) < 1e-12


def test_vcycle_256_bands(workload):
    from gpaw_amd.preconditioner import Preconditioner
    w, gd, pt, kin, H, psi, out, oham, ogd = workload
    d = Preconditioner(gd, kin)(psi, out=out)
    torch.cuda.synchronize()
    c, o = ost.laplace_stencil(ogd.h_cv, -0.5, 3)
    ref = opc.Preconditioner(ogd, c, o)(psi[SAMPLES].cpu().numpy())
    assert relerr(d[SAMPLES].cpu().numpy(), ref) < 1e-12


def test_single_precision_hpsi_and_vcycle_256_bands(workload):
    """The FP32 forms at the benchmark's shape -- the four-point ring
    stencil (28-row tiles), the FP32 point-major projector add with FP64
    sums, the FP32 V-cycle -- against the FP64 oracle on sampled bands
    (relative 2e-5 of the result's max-abs)."""
    from gpaw_amd import synthetic as syn
    w, gd, pt, kin, H, psi, out, oham, ogd = workload
    s32 = syn.System(w, gd, oham.vt_G, np.float32)
    x32 = psi[SAMPLES].float().contiguous()
    # all 256 bands through the kernels, sampled bands compared
    p32 = psi.float()
    o32 = torch.empty_like(p32)
    s32.H.apply(p32, o32)
    torch.cuda.synchronize()
    x = x32.double().cpu().numpy()
    ref = oham.apply(x)
    assert relerr(o32[SAMPLES].double().cpu().numpy(), ref) < 2e-5
    d32 = s32.pc(p32, out=o32)
    torch.cuda.synchronize()
    c, o = ost.laplace_stencil(ogd.h_cv, -0.5, 3)
    refd = opc.Preconditioner(ogd, c, o)(x)
    assert relerr(d32[SAMPLES].double().cpu().numpy(), refd) < 2e-5
    del p32, o32


@pytest.mark.parametrize('dtype', [np.complex64, np.complex128])
def test_complex_k_hpsi_and_vcycle_256_bands(workload, dtype):
    """The Bloch (k = (0.25, 0, 0)) forms at the benchmark's shape: the
    complex64 two-point ring path / complex128 one-point path with the
    x-wrap phase, the projector sweeps with per-image phases, and the
    V-cycle, against the complex128 oracle on sampled bands (2e-5 in
    single precision, 1e-12 in double)."""
    from gpaw_amd import synthetic as syn
    w, gd, pt, kin, H, psi, out, oham, ogd = workload
    k_c = (0.25, 0.0, 0.0)
    sc = syn.System(w, gd, oham.vt_G, dtype, k_c)
    td = torch.complex64 if dtype == np.complex64 else torch.complex128
    g = torch.Generator(device='cuda')
    g.manual_seed(5)
    p = torch.randn((256, 112, 112, 112), dtype=td, device='cuda',
                    generator=g)
    o = torch.empty_like(p)
    sc.H.apply(p, o, sc.phase_cd)
    torch.cuda.synchronize()
    x = p[SAMPLES].to(torch.complex128).cpu().numpy()
    ph = ogd.phase_cd(k_c)
    c, oo = ost.laplace_stencil(ogd.h_cv, -0.5, 3)
    olf = olfc.LFC(ogd, [[osp.Spline(l, rc, f) for l, rc, f in pp]
                         for pp in w.proj], k_c, complex)
    olf.set_positions(w.spos_ac)
    ok = orm.Hamiltonian(ogd, (c, oo), oham.vt_G, olf, w.dH, phase_cd=ph)
    ref = ok.apply(x)
    tol = 2e-5 if dtype == np.complex64 else 1e-12
    assert relerr(o[SAMPLES].to(torch.complex128).cpu().numpy(), ref) < tol

    class K:
        phase_cd = sc.phase_cd
    d = sc.pc(p, K, out=o)
    torch.cuda.synchronize()
    refd = opc.Preconditioner(ogd, c, oo)(x, ph)
    assert relerr(d[SAMPLES].to(torch.complex128).cpu().numpy(), refd) < tol
    del p, o
