"""GPU Jacobi relaxation and multigrid Poisson solver against the oracle
(oracle/poisson.py) and the exact FFT solution of the discrete equation.
Tolerances: 1e-12 relative for relaxation sweeps; 1e-10 for the converged
potential against the oracle's (same V-cycles), 1e-8 against FFT."""
import numpy as np
import pytest

torch = pytest.importorskip('torch')

pytestmark = pytest.mark.gpu

if not torch.cuda.is_available():
    pytest.skip('no GPU', allow_module_level=True)

from gpaw_amd.fd_operators import Laplace  # noqa: E402
from gpaw_amd.grid_descriptor import GridDescriptor  # noqa: E402
from gpaw_amd.poisson import FDPoissonSolver  # noqa: E402

from oracle import grid as og  # noqa: E402
from oracle import poisson as opo  # noqa: E402
from oracle.stencil import laplace_stencil  # noqa: E402
from test_oracle_poisson import charge  # noqa: E402


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


@pytest.mark.parametrize('nn,pbc,n', [(3, True, 3), (1, True, 4),
                                      (2, (True, False, True), 2)])
def test_relax_vs_oracle(nn, pbc, n):
    rng = np.random.default_rng(1)
    N_c = np.array([12, 14, 16])
    gd = GridDescriptor(N_c, N_c * 0.3, pbc)
    ogd = og.Grid(N_c, N_c * 0.3, pbc)
    f = rng.standard_normal(tuple(gd.n_c))
    s = rng.standard_normal(tuple(gd.n_c))
    op = Laplace(gd, -0.25 / np.pi, nn)
    fd = dev(f)
    op.relax(2, fd, dev(s), n, 2.0 / 3.0)
    c, o = laplace_stencil(ogd.h_cv, -0.25 / np.pi, nn)
    ref = opo.relax_jacobi(c, o, f.copy(), s, n, 2.0 / 3.0, ogd.pbc_c)
    assert abs(host(fd) - ref).max() < 1e-12 * abs(ref).max()


@pytest.mark.parametrize('nn,pbc,dtype', [(3, True, np.float64),
                                          (1, True, np.float64),
                                          (2, (True, False, True), np.float64),
                                          (3, False, np.float32)])
def test_relax_gauss_seidel_vs_oracle(nn, pbc, dtype):
    """Gauss-Seidel sweeps (relax_method 1, one launch per hyperplane)
    against the oracle's C-order sweep (itself pinned to the reference C)."""
    from oracle.poisson import relax_gauss_seidel
    rng = np.random.default_rng(6)
    N_c = np.array([10, 12, 14])
    gd = GridDescriptor(N_c, N_c * 0.3, pbc)
    ogd = og.Grid(N_c, N_c * 0.3, pbc)
    op = Laplace(gd, -0.25 / np.pi, nn, dtype)
    f = rng.standard_normal(tuple(gd.n_c))
    s = rng.standard_normal(tuple(gd.n_c))
    fd = dev(f.astype(dtype))
    op.relax(1, fd, dev(s.astype(dtype)), 3)
    c, o = laplace_stencil(ogd.h_cv, -0.25 / np.pi, nn)
    ref = relax_gauss_seidel(c, o, f.astype(dtype).astype(float), s.astype(
        dtype).astype(float), 3, ogd.pbc_c)
    tol = 2e-5 if dtype == np.float32 else 1e-13
    assert abs(host(fd) - ref).max() < tol * abs(ref).max()


def test_poisson_gauss_seidel_vs_oracle():
    """FDPoissonSolver(relax='GS'): same V-cycle count and potential as
    the oracle's solver with Gauss-Seidel smoothing."""
    rng = np.random.default_rng(7)
    N_c = np.array((16, 16, 16))
    ogd = og.Grid(N_c, N_c * 0.35)
    gd = GridDescriptor(N_c, N_c * 0.35)
    rho = charge(ogd, rng)
    ops = opo.FDPoissonSolver(nn=3, eps=1e-16, relax='GS')
    ops.set_grid_descriptor(ogd)
    ophi = ogd.zeros()
    oniter = ops.solve_neutral(ophi, rho)
    ps = FDPoissonSolver(nn=3, eps=1e-16, relax='GS')
    ps.set_grid_descriptor(gd)
    phi = gd.zeros()
    niter = ps.solve_neutral(phi, dev(rho))
    assert niter == oniter
    assert abs(host(phi) - ophi).max() < 1e-10 * abs(ophi).max()
    c, o = ops.operators[0]
    exact = opo.fft_solution(ogd, c, o, rho)
    assert abs(host(phi) - exact).max() < 1e-7 * abs(exact).max()


@pytest.mark.parametrize('N_c', [(24, 24, 24), (24, 28, 20), (40, 40, 40)])
def test_poisson_vs_oracle_and_fft(N_c):
    rng = np.random.default_rng(3)
    N_c = np.array(N_c)
    ogd = og.Grid(N_c, N_c * 0.35)
    gd = GridDescriptor(N_c, N_c * 0.35)
    rho = charge(ogd, rng)
    ops = opo.FDPoissonSolver(nn=3, eps=1e-16)
    ops.set_grid_descriptor(ogd)
    ophi = ogd.zeros()
    oniter = ops.solve_neutral(ophi, rho)
    ps = FDPoissonSolver(nn=3, eps=1e-16)
    ps.set_grid_descriptor(gd)
    assert ps.levels == ops.levels
    phi = gd.zeros()
    niter = ps.solve_neutral(phi, dev(rho))
    assert niter == oniter
    out = host(phi)
    assert abs(out - ophi).max() < 1e-10 * abs(ophi).max()
    c, o = ops.operators[0]
    exact = opo.fft_solution(ogd, c, o, rho)
    assert abs(out - exact).max() < 1e-7 * abs(exact).max()


@pytest.mark.parametrize('N_c', [(32, 32, 32), (24, 28, 20)])
def test_poisson_graph_replays_equal_eager(N_c):
    """The recorded V-cycle replayed (with the eager residual-norm read
    back and synchronisation between replays) gives, cycle for cycle, the
    residual norms and the potential of the eager solver, bit for bit; the
    one recording serves later solves with other densities."""
    rng = np.random.default_rng(4)
    N_c = np.array(N_c)
    ogd = og.Grid(N_c, N_c * 0.35)
    gd = GridDescriptor(N_c, N_c * 0.35)
    rhos = [charge(ogd, rng), charge(ogd, rng)]
    ps_g = FDPoissonSolver(nn=3, eps=1e-14)
    ps_g.set_grid_descriptor(gd)
    ps_e = FDPoissonSolver(nn=3, eps=1e-14, graph=False)
    ps_e.set_grid_descriptor(gd)
    for rho in (rhos[0], rhos[1], rhos[0]):
        phi_g = gd.zeros()
        phi_e = gd.zeros()
        n_g = ps_g.solve_neutral(phi_g, dev(rho))
        n_e = ps_e.solve_neutral(phi_e, dev(rho))
        assert n_g == n_e > 2
        assert ps_g.errors == ps_e.errors
        assert torch.equal(phi_g, phi_e)
    assert ps_g._graph is not None


def _nonperiodic(n=40, h=0.3):
    N_c = np.array((n, n, n))
    gd = GridDescriptor(N_c, N_c * h, False)
    ogd = og.Grid(N_c, N_c * h, False)
    return gd, ogd


def test_solve_charged_nonperiodic_monopole():
    """A Gaussian charge at the centre of a non-periodic box: the monopole
    Gaussian carries all of it, so phi is its free-space potential q v_0
    (gpaw/poisson.py:227-277)."""
    from gpaw_amd.poisson import Gaussian
    gd, ogd = _nonperiodic()
    g = Gaussian(gd)
    Q = -1.7
    rho = Q / np.sqrt(4 * np.pi) * g.get_gauss(0)
    ps = FDPoissonSolver(nn=3, eps=1e-18)
    ps.set_grid_descriptor(gd)
    phi = gd.zeros()
    niter = ps.solve(phi, dev(rho))
    assert niter > 1
    ref = Q / np.sqrt(4 * np.pi) * g.get_gauss_pot(0)
    assert abs(host(phi) - ref).max() < 1e-9 * abs(ref).max()


def test_solve_charged_nonperiodic_vs_oracle():
    """Charged density of two Gaussians off the centre: the product equals
    q v_0 + the oracle's neutral solve of rho - q g_0 from -q v_0."""
    from gpaw_amd.poisson import Gaussian
    gd, ogd = _nonperiodic(40, 0.3)
    r = gd.get_grid_point_coordinates()
    c0 = gd.cell_cv.sum(0) / 2
    rho = np.zeros(tuple(gd.n_c))
    for shift, w in (((1.0, 0.3, -0.5), 0.8), ((-0.7, -0.2, 0.9), 0.5)):
        d2 = sum((r[i] - c0[i] - shift[i])**2 for i in range(3))
        rho += w * np.exp(-d2 / 0.6)
    g = Gaussian(gd)
    Q = rho.sum() * gd.dv
    q = Q / np.sqrt(4 * np.pi)
    ps = FDPoissonSolver(nn=3, eps=1e-16)
    ps.set_grid_descriptor(gd)
    phi = gd.zeros()
    ps.solve(phi, dev(rho))
    ops = opo.FDPoissonSolver(nn=3, eps=1e-16)
    ops.set_grid_descriptor(ogd)
    ophi = -q * g.get_gauss_pot(0)
    ops.solve_neutral(ophi, rho - q * g.get_gauss(0))
    ref = ophi + q * g.get_gauss_pot(0)
    assert abs(host(phi) - ref).max() < 1e-9 * abs(ref).max()


def test_solve_remove_moment():
    """remove_moment=4: the monopole and the three dipoles of a density
    built from them are carried by Gaussians, leaving nothing for the
    multigrid solve: phi = sum_L q_L v_L."""
    from gpaw_amd.poisson import Gaussian
    gd, ogd = _nonperiodic()
    g = Gaussian(gd)
    qs = [0.9, -0.3, 0.25, 0.4]
    rho = sum(q * g.get_gauss(L) for L, q in enumerate(qs))
    ps = FDPoissonSolver(nn=3, eps=1e-18, remove_moment=4)
    ps.set_grid_descriptor(gd)
    phi = gd.zeros()
    ps.solve(phi, dev(rho))
    ref = sum(q * g.get_gauss_pot(L) for L, q in enumerate(qs))
    assert abs(host(phi) - ref).max() < 1e-8 * abs(ref).max()


def test_solve_charged_periodic_background():
    """A charged periodic cell: the uniform background of the net charge is
    removed; against the exact (FFT) solution of the neutralised density."""
    rng = np.random.default_rng(5)
    N_c = np.array((24, 24, 24))
    ogd = og.Grid(N_c, N_c * 0.35)
    gd = GridDescriptor(N_c, N_c * 0.35)
    rho = charge(ogd, rng) + 0.01          # net charge
    assert abs(rho.sum() * gd.dv) > 1e-3
    ps = FDPoissonSolver(nn=3, eps=1e-18)
    ps.set_grid_descriptor(gd)
    phi = gd.zeros()
    ps.solve(phi, dev(rho))
    ops = opo.FDPoissonSolver(nn=3)
    ops.set_grid_descriptor(ogd)
    c, o = ops.operators[0]
    exact = opo.fft_solution(ogd, c, o, rho - rho.mean())
    out = host(phi)
    assert abs(out - exact).max() < 1e-8 * abs(exact).max()


def test_solve_charge_center():
    """use_charge_center: the monopole Gaussian sits at the centre of the
    majority charge; for one off-centre Gaussian charge that is its own
    centre, and phi approaches its free-space potential inside the box."""
    from gpaw_amd.poisson import gaussian_potential
    gd, ogd = _nonperiodic(48, 0.25)
    r = gd.get_grid_point_coordinates()
    r0 = gd.cell_cv.sum(0) / 2 + np.array((0.8, -0.5, 0.3))
    xyz = [r[i] - r0[i] for i in range(3)]
    r2 = sum(v * v for v in xyz)
    a = 2.0
    Q = 1.3
    rho = Q * (a / np.pi)**1.5 * np.exp(-a * r2)
    ps = FDPoissonSolver(nn=3, eps=1e-16, use_charge_center=True)
    ps.set_grid_descriptor(gd)
    phi = gd.zeros()
    ps.solve(phi, dev(rho))
    exact = Q / np.sqrt(4 * np.pi) * gaussian_potential(a, 0, *xyz, r2)
    inner = r2 < 4.0
    out = host(phi)
    assert abs(out - exact)[inner].max() < 2e-4 * abs(exact).max()
