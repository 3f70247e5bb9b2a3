"""More LFC parity on the GPU (run on an MI355X): several k-points (the q
argument of integrate / add, lfc.py:552-631 with phase_qW of lfc.py:339-340),
atoms with different function counts and l up to 2 (the generic add kernel,
c/lfc.c:818-885), single-precision data, and the argument checks that keep
mis-shaped arrays away from the kernels."""
import numpy as np
import pytest

torch = pytest.importorskip('torch')

pytestmark = pytest.mark.gpu

if not torch.cuda.is_available():
    pytest.skip('no GPU', allow_module_level=True)

from gpaw_amd.grid_descriptor import GridDescriptor  # noqa: E402
from gpaw_amd.lfc import LFC  # noqa: E402
from gpaw_amd.spline import Spline  # noqa: E402
from oracle import grid as og  # noqa: E402
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


def build(rng, N_c, h, ls_a, dtype, kd, spos=None):
    """Product and oracle LFCs with atom a carrying radial functions of
    angular momenta ls_a[a]; kd: None or (nk, 3) k-points."""
    gd = GridDescriptor(N_c, np.array(N_c) * h)
    ogd = og.Grid(N_c, np.array(N_c) * h)
    r = np.linspace(0, 2.0, 101)
    spl, ospl = [], []
    for ls in ls_a:
        rc = rng.uniform(0.6, 1.0, len(ls))
        fs = [(l, np.exp(-r**2 / rc[j]**2)) for j, l in enumerate(ls)]
        spl.append([Spline.from_data(l, 2.0, f) for l, f in fs])
        ospl.append([osp.Spline(l, 2.0, f) for l, f in fs])
    if spos is None:
        spos = rng.uniform(0, 1, (len(ls_a), 3))
        spos[0] = (0.01, 0.98, 0.5)   # crosses periodic faces
    lfc = LFC(gd, spl, kd, dtype=dtype)
    lfc.set_positions(spos)

    def oracle(k_c):
        o = olfc.LFC(ogd, ospl, k_c, dtype)
        o.set_positions(spos)
        return o
    return lfc, oracle


def test_two_kpoints():
    """integrate / add with q = 0 and q = 1 use each k-point's phases."""
    rng = np.random.default_rng(21)
    N_c = (16, 18, 20)
    kpts = np.array([[0.25, 0.1, 0.0], [-0.1, 0.3, 0.2]])
    lfc, oracle = build(rng, N_c, 0.38, [(0, 0, 1, 1)] * 4, np.complex128,
                        kpts)
    a = rand(rng, (3,) + N_c, np.complex128)
    refs = []
    for q, k_c in enumerate(kpts):
        olf = oracle(k_c)
        ref = olf.integrate(a)
        refs.append(ref)
        c = lfc.integrate_M(dev(a), q=q)
        assert relerr(host(c), ref) < 1e-13
        coef = rand(rng, ref.shape, np.complex128)
        b = dev(a.copy())
        lfc.add_M(b, dev(coef), q=q)
        assert relerr(host(b), olf.add(a.copy(), coef)) < 1e-13
        # the dict interface of the reference
        c_axi = lfc.dict(3)
        lfc.integrate(dev(a), c_axi, q=q)
        for at in range(4):
            M = olf.M_a[at]
            assert relerr(host(c_axi[at]), ref[:, M:M + 8]) < 1e-13
    assert relerr(refs[0], refs[1]) > 1e-3     # the phases do matter
    with pytest.raises(ValueError):
        lfc.integrate_M(dev(a))                # q is required
    with pytest.raises(ValueError):
        lfc.integrate_M(dev(a), q=2)


@pytest.mark.parametrize('dtype,k_c', [(np.float64, None),
                                       (np.complex128, (0.25, -0.2, 0.1)),
                                       (np.float32, None),
                                       (np.complex64, (0.25, 0.0, 0.0))])
def test_mixed_function_counts(dtype, k_c):
    """Atoms with 9, 8, 5 and 1 functions (l = 0, 1, 2 mixed): the add
    takes the generic-count kernel."""
    rng = np.random.default_rng(22)
    N_c = (18, 20, 22)
    ls_a = [(0, 1, 2), (0, 0, 1, 1), (2,), (0,)]
    kd = None if k_c is None else np.array([k_c])
    lfc, oracle = build(rng, N_c, 0.38, ls_a, dtype, kd)
    olf = oracle(k_c)
    assert lfc.Mmax == olf.Mmax == 9 + 8 + 5 + 1
    assert lfc.Mmax_a == [9, 8, 5, 1]
    wide = np.complex128 if np.dtype(dtype).kind == 'c' else np.float64
    a = rand(rng, (7,) + N_c, dtype)
    ref = olf.integrate(a.astype(wide))
    c = lfc.integrate_M(dev(a))
    tol = 1e-5 if dtype in (np.float32, np.complex64) else 1e-13
    assert relerr(host(c), ref) < tol
    coef = rand(rng, ref.shape, dtype)
    b = dev(a.copy())
    lfc.add_M(b, dev(coef))
    refb = olf.add(a.astype(wide), coef.astype(wide))
    assert relerr(host(b), refb) < tol


@pytest.mark.parametrize('dtype', [np.float32, np.complex64])
def test_single_precision_uniform(dtype):
    """8 functions per atom in FP32 / complex64 against the FP64 oracle."""
    rng = np.random.default_rng(23)
    N_c = (20, 20, 24)
    k_c = None if dtype == np.float32 else (0.25, 0.0, 0.0)
    kd = None if k_c is None else np.array([k_c])
    lfc, oracle = build(rng, N_c, 0.38, [(0, 0, 1, 1)] * 6, dtype, kd)
    olf = oracle(k_c)
    wide = np.complex128 if np.dtype(dtype).kind == 'c' else np.float64
    a = rand(rng, (9,) + N_c, dtype)
    c = lfc.integrate_M(dev(a))
    ref = olf.integrate(a.astype(wide))
    assert relerr(host(c), ref) < 1e-5
    coef = rand(rng, ref.shape, dtype)
    b = dev(a.copy())
    lfc.add_M(b, dev(coef))
    assert relerr(host(b), olf.add(a.astype(wide), coef.astype(wide))) < 1e-5


def test_shape_checks():
    """Mis-shaped arrays raise before any pointer crosses the C ABI."""
    rng = np.random.default_rng(24)
    N_c = (12, 12, 12)
    lfc, _ = build(rng, N_c, 0.38, [(0, 0, 1, 1)] * 2, np.float64, None)
    a = torch.zeros((3,) + N_c, dtype=torch.float64, device='cuda')
    with pytest.raises(ValueError):
        lfc.integrate_M(torch.zeros((3, 12, 12, 13), dtype=torch.float64,
                                    device='cuda'))
    with pytest.raises(ValueError):
        lfc.integrate_M(a, torch.zeros((2, lfc.Mmax), dtype=torch.float64,
                                       device='cuda'))
    with pytest.raises(ValueError):
        lfc.add_M(a, torch.zeros((3, lfc.Mmax - 1), dtype=torch.float64,
                                 device='cuda'))
    from gpaw_amd.transformers import Transformer
    gd = GridDescriptor(N_c, np.array(N_c) * 0.38)
    t = Transformer(gd, gd.coarsen())
    with pytest.raises(ValueError):
        t.apply(a, torch.zeros((2, 6, 6, 6), dtype=torch.float64,
                               device='cuda'))


@pytest.mark.parametrize('dtype,k_c', [(np.float64, None),
                                       (np.complex128, (0.25, -0.2, 0.1)),
                                       (np.float32, None),
                                       (np.complex64, (0.25, 0.0, 0.0))])
@pytest.mark.parametrize('ls_a,h', [
    ([(0, 0, 1, 1)] * 7, 0.38),                 # 8 functions per atom
    ([(0, 1, 2, 2), (0, 0, 1, 1), (2,)], 0.38),  # up to 14: 16-wide rows
    ([(2, 2, 2, 2, 2), (0, 1)], 0.38),           # 25 functions: 32-wide rows
    ([(0, 0, 1, 1)] * 3, 0.25),                 # spheres overlap their images
])
def test_points_add_equals_colour_add(monkeypatch, dtype, k_c, ls_a, h):
    """The one-launch point-major add against the coloured per-atom
    launches (GPAW_AMD_LFC_ADD=colour) and the oracle, with 37 bands
    (ragged band tiles of both kernels)."""
    rng = np.random.default_rng(25)
    N_c = (16, 18, 20) if h > 0.3 else (10, 10, 12)
    kd = None if k_c is None else np.array([k_c])
    spos = rng.uniform(0, 1, (len(ls_a), 3))
    spos[0] = (0.01, 0.98, 0.5)
    lfc, oracle = build(np.random.default_rng(26), N_c, h, ls_a, dtype, kd,
                        spos)
    monkeypatch.setenv('GPAW_AMD_LFC_ADD', 'colour')
    lfc_c, _ = build(np.random.default_rng(26), N_c, h, ls_a, dtype, kd,
                     spos)
    monkeypatch.delenv('GPAW_AMD_LFC_ADD')
    a = rand(rng, (37,) + N_c, dtype)
    coef = dev(rand(rng, (37, lfc.Mmax), dtype))
    b_p = dev(a.copy())
    b_c = dev(a.copy())
    lfc.add_M(b_p, coef)
    lfc_c.add_M(b_c, coef)
    single = dtype in (np.float32, np.complex64)
    assert relerr(host(b_p), host(b_c)) < (2e-6 if single else 1e-14)
    wide = np.complex128 if np.dtype(dtype).kind == 'c' else np.float64
    ref = oracle(k_c).add(a.astype(wide), host(coef).astype(wide))
    assert relerr(host(b_p), ref) < (1e-5 if single else 1e-13)
    # deterministic: the same bits twice
    b_q = dev(a.copy())
    lfc.add_M(b_q, coef)
    assert torch.equal(b_p, b_q)
