"""GPU parity of the HIP path against the oracle (run on an MI355X).

Every test compares libgpawhip results (through the drop-in classes and the
C ABI) with oracle/ on the same seeded inputs.  Tolerances: relative 1e-12
of the result's max-abs for FP64 / complex128 (the reference's own GPU-vs-CPU
tolerance for the Laplacian, test_fd_laplace.py:54), 1e-13 for transformers
and LFC, 2e-5 for FP32.
"""
import numpy as np
import pytest

torch = pytest.importorskip('torch')

pytestmark = pytest.mark.gpu

if not torch.cuda.is_available():
    pytest.skip('no GPU', allow_module_level=True)

from gpaw_amd.fd_operators import FDOperator, Laplace  # noqa: E402
from gpaw_amd.grid_descriptor import GridDescriptor  # noqa: E402
from gpaw_amd.transformers import Transformer  # noqa: E402
from gpaw_amd.preconditioner import Preconditioner  # noqa: E402
from gpaw_amd.lfc import LFC  # noqa: E402
from gpaw_amd.spline import Spline  # noqa: E402
from gpaw_amd.hamiltonian import FDHamiltonian  # noqa: E402

from oracle import grid as og  # noqa: E402
from oracle import stencil as ost  # noqa: E402
from oracle import transformers as otr  # noqa: E402
from oracle import precond as opc  # noqa: E402
from oracle import lfc as olfc  # noqa: E402
from oracle import spline as osp  # noqa: E402


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


def relerr(a, b):
    return abs(a - b).max() / max(abs(b).max(), 1e-300)


def rand(rng, shape, dtype):
    a = rng.standard_normal(shape)
    if np.dtype(dtype).kind == 'c':
        a = a + 1j * rng.standard_normal(shape)
    return a.astype(dtype)


STENCIL_CASES = [
    # N_c, h, pbc, n, dtype, k_c
    ((48, 48, 48), 0.2 / 0.52917721067, True, 3, np.float64, None),
    ((20, 24, 28), 0.35, True, 3, np.complex128, (0.25, 0, 0)),
    ((20, 24, 28), 0.35, True, 3, np.complex128, (0.1, -0.3, 0.45)),
    ((16, 17, 18), 0.4, (True, False, True), 3, np.float64, None),
    ((16, 18, 20), 0.4, False, 2, np.float64, None),
    ((12, 14, 16), 0.4, True, 1, np.float64, None),
    ((14, 12, 30), 0.3, True, 6, np.float64, None),
    ((14, 12, 30), 0.3, True, 5, np.complex128, (0.2, 0.1, 0.3)),
    ((10, 12, 200), 0.3, True, 4, np.float64, None),
    ((20, 24, 28), 0.35, True, 3, np.float32, None),
    ((20, 24, 28), 0.35, True, 3, np.complex64, (0.25, 0, 0)),
    ((20, 24, 28), 0.35, True, 3, np.complex64, (0.1, -0.3, 0.45)),
    ((16, 18, 20), 0.35, (True, False, True), 3, np.complex64,
     (0.25, 0, 0)),
    ((16, 18, 20), 0.35, (False, True, False), 2, np.complex64,
     (0, 0.2, 0)),
    ((18, 16, 22), 0.35, (True, True, False), 3, np.float32, None),
]


@pytest.mark.parametrize('N_c,h,pbc,n,dtype,k_c', STENCIL_CASES)
def test_laplace_vs_oracle(N_c, h, pbc, n, dtype, k_c):
    rng = np.random.default_rng(1)
    gd = GridDescriptor(N_c, np.array(N_c) * h, pbc)
    ogd = og.Grid(N_c, np.array(N_c) * h, pbc)
    nb = 4
    a = rand(rng, (nb,) + tuple(gd.n_c), dtype)
    ph = None if k_c is None else ogd.phase_cd(k_c)
    op = Laplace(gd, -0.5, n, dtype)
    out = torch.empty_like(dev(a))
    op.apply(dev(a), out, ph)
    c, o = ost.laplace_stencil(ogd.h_cv, -0.5, n)
    ref = ost.fd_apply(c, o, a.astype(np.complex128 if np.dtype(dtype).kind
                                      == 'c' else np.float64), ogd.pbc_c, ph)
    tol = 2e-5 if dtype in (np.float32, np.complex64) else 1e-12
    assert relerr(host(out), ref) < tol


@pytest.mark.parametrize('kernel', ['ring', 'reg', 'persist'])
@pytest.mark.parametrize('N_c,pbc,dtype,k_c',
                         [((24, 20, 28), True, np.float64, None),
                          ((24, 20, 28), True, np.complex128, (0.25, 0, 0.1)),
                          ((16, 17, 18), (False, False, True), np.float64,
                           None),
                          ((18, 14, 16), True, np.float32, None)])
def test_stencil_kernel_variants(monkeypatch, kernel, N_c, pbc, dtype, k_c):
    """Both stencil kernels (LDS-ring DMA and register-queue) on the same
    inputs, with partial last tiles (odd rows per tile) and short x-runs;
    'persist' runs the ring kernel as 8 resident blocks that walk several
    tiles each."""
    monkeypatch.delenv('GPAW_AMD_STENCIL', raising=False)
    monkeypatch.delenv('GPAW_AMD_PERSIST', raising=False)
    if kernel == 'reg':
        monkeypatch.setenv('GPAW_AMD_STENCIL', 'reg')
    elif kernel == 'persist':
        monkeypatch.setenv('GPAW_AMD_PERSIST', '8')
    rng = np.random.default_rng(12)
    gd = GridDescriptor(N_c, np.array(N_c) * 0.3, pbc)
    ogd = og.Grid(N_c, np.array(N_c) * 0.3, pbc)
    a = rand(rng, (3,) + tuple(gd.n_c), dtype)
    vt = rng.standard_normal(tuple(gd.n_c)).astype(
        np.float32 if dtype == np.float32 else np.float64)
    ph = None if k_c is None else ogd.phase_cd(k_c)
    c, o = ost.laplace_stencil(ogd.h_cv, -0.5, 3)
    wide = np.complex128 if np.dtype(dtype).kind == 'c' else np.float64
    ref = ost.fd_apply(c, o, a.astype(wide), ogd.pbc_c, ph) + vt * a
    tol = 2e-5 if dtype == np.float32 else 1e-12
    for rows, planes in ((0, 0), (5, 7), (3, 4)):
        op = Laplace(gd, -0.5, 3, dtype)
        op.set_tile(rows, planes)
        out = torch.empty_like(dev(a))
        op.apply(dev(a), out, ph, vt_g=dev(vt))
        assert relerr(host(out), ref) < tol, (rows, planes)


@pytest.mark.parametrize('kernel', ['ring', 'reg'])
@pytest.mark.parametrize('dtype,k_c,cell', [
    (np.float64, None, 'ortho'), (np.complex128, (0.25, 0, 0), 'ortho'),
    (np.complex128, (0.1, -0.3, 0.45), 'ortho'), (np.float32, None, 'ortho'),
    (np.complex64, (0.25, 0, 0), 'ortho'), (np.complex128, (0.25, 0.1, 0), 'fcc')])
def test_apply_shifted_residual(monkeypatch, kernel, dtype, k_c, cell):
    """(T + v - eps_x) psi in one pass against the oracle, every kernel path
    (two-point real and complex, one-point complex, register-queue; the
    non-orthogonal stencil takes apply + axpy)."""
    monkeypatch.delenv('GPAW_AMD_STENCIL', raising=False)
    if kernel == 'reg':
        monkeypatch.setenv('GPAW_AMD_STENCIL', 'reg')
    rng = np.random.default_rng(16)
    N_c = (16, 20, 24)
    cell_cv = (4.0 * np.array([(0, 1, 1), (1, 0, 1), (1, 1, 0)]) if cell == 'fcc'
               else np.diag(np.array(N_c) * 0.3))
    gd = GridDescriptor(N_c, cell_cv)
    ogd = og.Grid(N_c, cell_cv)
    nb = 5
    a = rand(rng, (nb,) + N_c, dtype)
    single = dtype in (np.float32, np.complex64)
    vt = rng.standard_normal(N_c).astype(np.float32 if single else np.float64)
    eps = rng.uniform(-1, 1, nb)
    ph = None if k_c is None else ogd.phase_cd(k_c)
    op = Laplace(gd, -0.5, 2 if cell == 'fcc' else 3, dtype)
    out = torch.empty_like(dev(a))
    op.apply_shifted(dev(a), out, torch.from_numpy(eps).cuda(), ph,
                     vt_g=dev(vt))
    c, o = ost.laplace_stencil(ogd.h_cv, -0.5, 2 if cell == 'fcc' else 3)
    wide = np.complex128 if np.dtype(dtype).kind == 'c' else np.float64
    aw = a.astype(wide)
    ref = (ost.fd_apply(c, o, aw, ogd.pbc_c, ph) + vt * aw
           - eps[:, None, None, None] * aw)
    tol = 2e-5 if single else 1e-12
    assert relerr(host(out), ref) < tol


def test_laplace_fused_potential_and_source():
    rng = np.random.default_rng(2)
    N_c = (24, 20, 28)
    gd = GridDescriptor(N_c, np.array(N_c) * 0.3)
    ogd = og.Grid(N_c, np.array(N_c) * 0.3)
    a = rand(rng, (3,) + N_c, np.complex128)
    src = rand(rng, (3,) + N_c, np.complex128)
    vt = rng.standard_normal(N_c)
    ph = ogd.phase_cd((0.25, 0.0, 0.1))
    op = Laplace(gd, -0.5, 3, complex)
    out = torch.empty_like(dev(a))
    op.apply(dev(a), out, ph, vt_g=dev(vt), src_xg=dev(src), beta=-0.7)
    c, o = ost.laplace_stencil(ogd.h_cv, -0.5, 3)
    ref = ost.fd_apply(c, o, a, ogd.pbc_c, ph) + vt * a - 0.7 * src
    assert relerr(host(out), ref) < 1e-12


@pytest.mark.parametrize('cell', ['fcc', 'hex'])
def test_generic_stencil_nonorthogonal(cell):
    rng = np.random.default_rng(3)
    if cell == 'fcc':
        cell_cv = 4.0 * np.array([(0, 1, 1), (1, 0, 1), (1, 1, 0)])
    else:
        cell_cv = np.array([(6, 0, 0), (3, 3 * 3**0.5, 0), (0, 0, 7)])
    N_c = (16, 16, 18)
    gd = GridDescriptor(N_c, cell_cv)
    ogd = og.Grid(N_c, cell_cv)
    op = Laplace(gd, -0.5, 2, complex)
    assert not op.cfd
    a = rand(rng, (2,) + N_c, np.complex128)
    ph = ogd.phase_cd((0.25, 0.1, 0.0))
    out = torch.empty_like(dev(a))
    op.apply(dev(a), out, ph)
    c, o = ost.laplace_stencil(ogd.h_cv, -0.5, 2)
    assert relerr(host(out), ost.fd_apply(c, o, a, ogd.pbc_c, ph)) < 1e-12


def test_stencil_numpy_interface():
    """NumPy arrays in, NumPy arrays out (staged through the GPU)."""
    rng = np.random.default_rng(4)
    N_c = (12, 12, 12)
    gd = GridDescriptor(N_c, np.array(N_c) * 0.4)
    a = rng.standard_normal((2,) + N_c)
    out = np.zeros_like(a)
    Laplace(gd, -0.5, 3).apply(a, out)
    c, o = ost.laplace_stencil(og.Grid(N_c, np.array(N_c) * 0.4).h_cv,
                               -0.5, 3)
    assert relerr(out, ost.fd_apply(c, o, a, [True] * 3)) < 1e-12


def test_stencil_rejects_aliasing():
    gd = GridDescriptor((8, 8, 8), (3, 3, 3))
    a = torch.zeros((1, 8, 8, 8), dtype=torch.float64, device='cuda')
    with pytest.raises(ValueError):
        Laplace(gd, -0.5, 3).apply(a, a)


TRANSFORMER_CASES = [(nn, dt, pbc, k) for nn in (1, 2, 3, 4)
                     for dt in (np.float64, np.complex128)
                     for pbc in (True, (True, False, True))
                     for k in ((0.25, -0.1, 0.3), (0.25, 0, 0))]


@pytest.mark.parametrize('nn,dtype,pbc,k_c', TRANSFORMER_CASES)
def test_transformers_vs_oracle(nn, dtype, pbc, k_c):
    rng = np.random.default_rng(5)
    N_c = np.array([16, 20, 24])
    gd0 = GridDescriptor(N_c, N_c * 0.3, pbc)
    gd1 = gd0.coarsen()
    ogd0 = og.Grid(N_c, N_c * 0.3, pbc)
    ogd1 = ogd0.coarsen()
    ph = ogd0.phase_cd(k_c) if np.dtype(dtype).kind == 'c' else None
    for gin, gout, oin, oout in ((gd0, gd1, ogd0, ogd1),
                                 (gd1, gd0, ogd1, ogd0)):
        a = rand(rng, (3,) + tuple(gin.n_c), dtype)
        T = Transformer(gin, gout, nn, dtype)
        out = T.apply(dev(a), phases=ph)
        ref = otr.Transformer(oin, oout, nn).apply(a, ph)
        assert relerr(host(out), ref) < 1e-13


@pytest.mark.parametrize('dtype', [np.float32, np.complex64])
def test_transformers_single(dtype):
    rng = np.random.default_rng(6)
    N_c = np.array([16, 16, 20])
    gd0 = GridDescriptor(N_c, N_c * 0.3)
    ogd0 = og.Grid(N_c, N_c * 0.3)
    ph = ogd0.phase_cd((0.25, 0, 0)) if np.dtype(dtype).kind == 'c' else None
    a = rand(rng, (2,) + tuple(N_c), dtype)
    out = Transformer(gd0, gd0.coarsen(), 3, dtype).apply(dev(a), phases=ph)
    ref = otr.Transformer(ogd0, ogd0.coarsen(), 3).apply(
        a.astype(np.complex128 if ph is not None else np.float64), ph)
    assert relerr(host(out), ref) < 2e-6


@pytest.mark.parametrize('dtype,k_c', [(np.float64, None),
                                       (np.complex128, (0.25, 0.1, 0))])
def test_precond_single_precision_option(dtype, k_c):
    """precision='single': FP64 residuals in and out, the cycle in
    FP32 / complex64 (an extension); agrees with the FP64 oracle to single
    precision."""
    rng = np.random.default_rng(17)
    N_c = np.array([24, 28, 32])
    gd0 = GridDescriptor(N_c, N_c * 0.3)
    ogd0 = og.Grid(N_c, N_c * 0.3)
    kin0 = Laplace(gd0, -0.5, 3, dtype)
    P = Preconditioner(gd0, kin0, dtype, precision='single')
    R = rand(rng, (4,) + tuple(N_c), dtype)

    class K:
        phase_cd = None if k_c is None else ogd0.phase_cd(k_c)
    d = P(dev(R), K)
    assert d.dtype == dev(R).dtype
    c0, o0 = ost.laplace_stencil(ogd0.h_cv, -0.5, 3)
    ref = opc.Preconditioner(ogd0, c0, o0)(R, K.phase_cd)
    assert relerr(host(d), ref) < 2e-6


@pytest.mark.parametrize('dtype,k_c', [(np.float64, None),
                                       (np.complex128, (0.25, 0, 0)),
                                       (np.complex128, (0.1, 0.2, -0.3)),
                                       (np.float32, None),
                                       (np.complex64, (0.25, 0, 0))])
def test_precond_vs_oracle(dtype, k_c):
    rng = np.random.default_rng(7)
    N_c = np.array([24, 28, 32])
    h = 0.3
    gd0 = GridDescriptor(N_c, N_c * h)
    ogd0 = og.Grid(N_c, N_c * h)
    kin0 = Laplace(gd0, -0.5, 3, dtype)
    P = Preconditioner(gd0, kin0, dtype)
    R = rand(rng, (5,) + tuple(N_c), dtype)

    class K:
        phase_cd = None if k_c is None else ogd0.phase_cd(k_c)
    d = P(dev(R), K)
    c0, o0 = ost.laplace_stencil(ogd0.h_cv, -0.5, 3)
    wide = np.complex128 if np.dtype(dtype).kind == 'c' else np.float64
    ref = opc.Preconditioner(ogd0, c0, o0)(R.astype(wide), K.phase_cd)
    single = dtype in (np.float32, np.complex64)
    assert relerr(host(d), ref) < (2e-6 if single else 1e-12)


def _lfc_pair(rng, N_c, h, natoms, dtype, k_c):
    gd = GridDescriptor(N_c, np.array(N_c) * h)
    ogd = og.Grid(N_c, np.array(N_c) * h)
    r = np.linspace(0, 2.0, 101)
    spl, ospl = [], []
    for a in range(natoms):
        rc = rng.uniform(0.6, 1.0, 4)
        fs = [(l, np.exp(-r**2 / rc[j]**2)) for j, l in
              enumerate((0, 0, 1, 1))]
        spl.append([Spline.from_data(l, 2.0, f) for l, f in fs])
        ospl.append([osp.Spline(l, 2.0, f) for l, f in fs])
    spos = rng.uniform(0, 1, (natoms, 3))
    spos[0] = (0.01, 0.98, 0.5)  # crosses periodic faces
    kd = None if k_c is None else np.array([k_c])
    lfc = LFC(gd, spl, kd, dtype=dtype)
    lfc.set_positions(spos)
    olf = olfc.LFC(ogd, ospl, k_c, dtype)
    olf.set_positions(spos)
    return gd, lfc, olf


@pytest.mark.parametrize('dtype,k_c', [(np.float64, None),
                                       (np.complex128, (0.25, 0.1, 0.0))])
def test_lfc_vs_oracle(dtype, k_c):
    rng = np.random.default_rng(8)
    N_c = (20, 22, 24)
    gd, lfc, olf = _lfc_pair(rng, N_c, 0.38, 6, dtype, k_c)
    assert lfc.Mmax == olf.Mmax == 48
    a = rand(rng, (5,) + N_c, dtype)
    c = lfc.integrate_M(dev(a))
    ref = olf.integrate(a)
    assert relerr(host(c), ref) < 1e-13
    coef = rand(rng, ref.shape, dtype)
    b = dev(a.copy())
    lfc.add_M(b, dev(coef))
    refb = olf.add(a.copy(), coef)
    assert relerr(host(b), refb) < 1e-13
    # reference dict interface
    c_axi = lfc.dict(5)
    lfc.integrate(dev(a), c_axi, q=0 if k_c else -1)
    for at in range(6):
        M = olf.M_a[at]
        assert relerr(host(c_axi[at]), ref[:, M:M + 8]) < 1e-13


@pytest.mark.parametrize('dtype,k_c', [(np.float64, None),
                                       (np.complex128, (0.25, -0.3, 0.1))])
def test_lfc_self_overlap(dtype, k_c):
    """Cell (2.5 x 2.5 x 3 Bohr) shorter than the sphere diameter (4 Bohr):
    every sphere overlaps its own periodic images, so one atom has several
    entries on one grid point (the add must not race within the atom)."""
    rng = np.random.default_rng(12)
    N_c = (10, 10, 12)
    gd, lfc, olf = _lfc_pair(rng, N_c, 0.25, 3, dtype, k_c)
    # more entries than 3 atoms x grid points: some atom repeats a point
    assert lfc.nentries > 3 * np.prod(N_c)
    a = rand(rng, (6,) + N_c, dtype)
    c = lfc.integrate_M(dev(a))
    ref = olf.integrate(a)
    assert relerr(host(c), ref) < 1e-13
    coef = rand(rng, ref.shape, dtype)
    b = dev(a.copy())
    lfc.add_M(b, dev(coef))
    refb = olf.add(a.copy(), coef)
    assert relerr(host(b), refb) < 1e-13


def test_lfc_deterministic():
    """test_lf.py:8-25 -- add twice gives identical bits."""
    rng = np.random.default_rng(9)
    gd, lfc, olf = _lfc_pair(rng, (24, 24, 24), 0.38, 10, np.float64, None)
    coef = dev(rng.standard_normal((3, lfc.Mmax)))
    x1 = torch.zeros((3, 24, 24, 24), dtype=torch.float64, device='cuda')
    x2 = torch.zeros_like(x1)
    lfc.add_M(x1, coef)
    lfc.add_M(x2, coef)
    assert torch.equal(x1, x2)


def test_full_hpsi_vs_oracle():
    from gpaw_amd import synthetic as syn
    rng = np.random.default_rng(10)
    N = 28
    w = syn.full_hpsi_geometry(N, 6)
    gd = GridDescriptor((N,) * 3, (w.L,) * 3)
    ogd = og.Grid((N,) * 3, (w.L,) * 3)
    spl = [[Spline.from_data(l, rc, f) for l, rc, f in p] for p in w.proj]
    ospl = [[osp.Spline(l, rc, f) for l, rc, f in p] for p in w.proj]
    pt = LFC(gd, spl)
    pt.set_positions(w.spos_ac)
    opt = olfc.LFC(ogd, ospl)
    opt.set_positions(w.spos_ac)
    vt = syn.potential(gd, w.pos_av)
    kin = Laplace(gd, -0.5, 3)
    H = FDHamiltonian(gd, kin, vt, pt, w.dH)
    psi = rng.standard_normal((4, N, N, N))
    out = torch.empty_like(dev(psi))
    H.apply(dev(psi), out)
    c, o = ost.laplace_stencil(ogd.h_cv, -0.5, 3)
    ref = ost.fd_apply(c, o, psi, ogd.pbc_c) + vt * psi
    P = opt.integrate(psi)
    Q = np.concatenate([P[:, 8 * a:8 * a + 8] @ w.dH[a]
                        for a in range(6)], axis=1)
    ref = opt.add(ref, Q)
    assert relerr(host(out), ref) < 1e-12


def test_multi_dotc_and_axpy():
    from gpaw_amd import blas
    rng = np.random.default_rng(11)
    a = rand(rng, (7, 1000), np.complex128)
    b = rand(rng, (7, 1000), np.complex128)
    d = blas.multi_dotc(dev(a), dev(b), scale=0.5)
    ref = 0.5 * np.einsum('xg,xg->x', a.conj(), b).real
    np.testing.assert_allclose(host(d), ref, rtol=1e-13)
    lam = rng.standard_normal(7)
    y = dev(b.copy())
    blas.multi_axpy(dev(lam), dev(a), y)
    np.testing.assert_allclose(host(y), b + lam[:, None] * a, rtol=1e-14)


@pytest.mark.parametrize('dtype', [np.float64, np.complex128])
@pytest.mark.parametrize('k', [1, 3, 6])
def test_multi_gram(dtype, k):
    """Per-band Gram matrices of the DIIS step against numpy."""
    from gpaw_amd import blas
    rng = np.random.default_rng(13)
    vs = [rand(rng, (5, 7, 9, 11), dtype) for _ in range(k)]
    G = host(blas.multi_gram([dev(v) for v in vs], scale=0.3))
    ref = 0.3 * np.einsum('ixg,jxg->xij',
                          np.array([v.reshape(5, -1) for v in vs]).conj(),
                          np.array([v.reshape(5, -1) for v in vs]))
    np.testing.assert_allclose(G, ref, rtol=1e-13, atol=1e-13 * abs(ref).max())


@pytest.mark.parametrize('dtype', [np.float64, np.complex128, np.float32,
                                   np.complex64])
@pytest.mark.parametrize('k', [1, 2, 3, 8])
def test_multi_lincomb(dtype, k):
    """Per-band linear combinations (the DIIS update and x + lam dx)
    against numpy, terms added in the order given (the first term is a
    product, the rest are FMAs, so the error is a few ulps of the largest
    term); also written over one of its inputs."""
    from gpaw_amd import blas
    rng = np.random.default_rng(15)
    nb, shape = 5, (5, 7, 9, 11)
    vs = [rand(rng, shape, dtype) for _ in range(k)]
    cplx = np.iscomplexobj(vs[0])
    c = rng.standard_normal((nb, k))
    if cplx:
        c = c + 1j * rng.standard_normal((nb, k))
    ref = sum(c[:, i].reshape(-1, 1, 1, 1) * vs[i].astype(np.complex128)
              for i in range(k))
    if not cplx:
        ref = ref.real
    tol = 1e-6 if dtype in (np.float32, np.complex64) else 1e-14
    d = [dev(v) for v in vs]
    out = blas.multi_lincomb(d, torch.from_numpy(c).cuda())
    assert relerr(host(out), ref) < tol
    blas.multi_lincomb(d, torch.from_numpy(c).cuda(), out=d[-1])
    assert relerr(host(d[-1]), ref) < tol
    if not cplx:
        with pytest.raises(AssertionError):
            blas.multi_lincomb(d, torch.from_numpy(c + 1j).cuda())


@pytest.mark.parametrize('dtype', [np.float64, np.complex128, np.float32,
                                   np.complex64])
def test_add_to_density(dtype):
    """nt += sum_n f_n |psi_n|^2 against the oracle (c/utilities.c)."""
    from gpaw_amd.density import add_to_density, add_to_density_bands
    from oracle import density as ode
    rng = np.random.default_rng(14)
    psit = rand(rng, (7, 10, 12, 14), dtype)
    f = rng.uniform(0, 2, 7)
    nt0 = rng.standard_normal((10, 12, 14))
    ref = ode.add_to_density(f, psit.astype(np.complex128), nt0.copy())
    nt = dev(nt0)
    add_to_density_bands(f, dev(psit), nt)
    tol = 1e-6 if dtype in (np.float32, np.complex64) else 1e-14
    assert relerr(host(nt), ref) < tol
    # the per-band drop-in
    nt1 = dev(nt0)
    for fn, p in zip(f, psit):
        add_to_density(fn, dev(p), nt1)
    assert relerr(host(nt1), ref) < tol
