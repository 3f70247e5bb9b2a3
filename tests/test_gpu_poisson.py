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


def test_relax_gauss_seidel_not_provided():
    gd = GridDescriptor((8, 8, 8), (3.0, 3.0, 3.0))
    f = torch.zeros((8, 8, 8), dtype=torch.float64, device='cuda')
    with pytest.raises(NotImplementedError):
        Laplace(gd, 1.0, 1).relax(1, f, f.clone(), 1)


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
