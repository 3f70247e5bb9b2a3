"""Batched cells of uneven size (gh_stencil_batch + BatchedHamiltonian)
against the single-cell path and the oracle: the one persistent launch
gives bit-identical stencil outputs to one launch per cell (same per-point
arithmetic whatever the tiling), the full H psi of every cell matches the
FP64 oracle (gpaw/hamiltonian.py:445-479) to 2e-5 in FP32 and 1e-12 in
FP64, and a HIP-graph replay of the step equals the eager step on every
replay, with other work on the GPU between replays."""
import numpy as np
import pytest

torch = pytest.importorskip('torch')

pytestmark = pytest.mark.gpu

if not torch.cuda.is_available():
    pytest.skip('no GPU', allow_module_level=True)

from gpaw_amd import synthetic as syn  # noqa: E402
from gpaw_amd.batch import BatchedHamiltonian, StencilBatch  # noqa: E402

from oracle import grid as og  # noqa: E402
from oracle import lfc as olfc  # noqa: E402
from oracle import spline as osp  # noqa: E402
from oracle import stencil as ost  # noqa: E402

SMALL = ((16, 3, 2), (12, 2, 2), (24, 5, 4), (20, 4, 3), (32, 6, 6),
         (28, 2, 5))


def oracle_hpsi(w, vt, psi):
    N = w.N
    ogd = og.Grid((N,) * 3, (w.L,) * 3)
    opt = olfc.LFC(ogd, [[osp.Spline(l, rc, f) for l, rc, f in p]
                         for p in w.proj])
    opt.set_positions(w.spos_ac)
    c, o = ost.laplace_stencil(ogd.h_cv, -0.5, 3)
    ref = ost.fd_apply(c, o, psi, ogd.pbc_c) + vt * psi
    P = opt.integrate(psi)
    Q = np.concatenate([P[:, 8 * a:8 * a + 8] @ w.dH[a]
                        for a in range(w.natoms)], axis=1)
    return opt.add(ref, Q)


def inputs(cells, dtype, seed=1):
    g = torch.Generator(device='cuda')
    g.manual_seed(seed)
    psis = [torch.randn((nb,) + tuple(gd.n_c), dtype=torch.float64,
                        device='cuda', generator=g).to(dtype)
            for w, gd, vt, nb, s in cells]
    outs = [torch.empty_like(p) for p in psis]
    return psis, outs


@pytest.mark.parametrize('dtype', [np.float32, np.float64])
def test_batch_stencil_bitwise_and_hpsi(dtype):
    td = torch.float32 if dtype == np.float32 else torch.float64
    cells = syn.batch_cells(SMALL, dtype)
    psis, outs = inputs(cells, td)
    bh = BatchedHamiltonian([s.H for *_, s in cells],
                            [nb for w, gd, vt, nb, s in cells], tune=True)
    assert bh.batch.tiles >= len(cells)
    # stencil alone: batched == per-cell launch with the same tiles
    bh.batch.apply(psis, outs, [s.H.vt_G for *_, s in cells])
    for c, (w, gd, vt, nb, s) in enumerate(cells):
        ref = torch.empty_like(outs[c])
        s.kin.set_tile(bh.batch.rows[c], 0, bh.batch.depths[c])
        s.H.apply_local(psis[c], ref)
        assert torch.equal(outs[c], ref), c
    # full H psi against the FP64 oracle
    bh.apply(psis, outs)
    torch.cuda.synchronize()
    tol = 2e-5 if dtype == np.float32 else 1e-12
    for c, (w, gd, vt, nb, s) in enumerate(cells):
        x = psis[c].double().cpu().numpy()
        ref = oracle_hpsi(w, vt, x)
        got = outs[c].double().cpu().numpy()
        assert abs(got - ref).max() < tol * abs(ref).max(), c


def test_batch_graph_replays():
    cells = syn.batch_cells(SMALL, np.float32)
    psis, outs = inputs(cells, torch.float32, seed=3)
    bh = BatchedHamiltonian([s.H for *_, s in cells],
                            [nb for w, gd, vt, nb, s in cells], tune=False)
    bh.capture(psis, outs)
    g = torch.Generator(device='cuda')
    for rep in range(3):
        g.manual_seed(100 + rep)
        for p in psis:      # new inputs in place (eager work between replays)
            p.copy_(torch.randn(p.shape, dtype=torch.float64, device='cuda',
                                generator=g).to(torch.float32))
        bh.replay()
        torch.cuda.synchronize()
        got = [o.clone() for o in outs]
        eager = [torch.empty_like(o) for o in outs]
        bh.apply(psis, eager)
        torch.cuda.synchronize()
        for c in range(len(cells)):
            assert torch.equal(got[c], eager[c]), (rep, c)


def test_batch_refuses_bad_input():
    cells = syn.batch_cells(SMALL[:2], np.float32)
    sb = StencilBatch([s.kin for *_, s in cells],
                      [nb for w, gd, vt, nb, s in cells])
    psis, outs = inputs(cells, torch.float32)
    with pytest.raises(ValueError):
        sb.apply(psis, outs[::-1], [s.H.vt_G for *_, s in cells])
    with pytest.raises(ValueError):
        StencilBatch([s.kin for *_, s in cells], [3])


def test_batch_production_cells_sampled():
    """The six cells of the benchmark (40^3 .. 128^3, 16 .. 512 bands, 8 ..
    96 atoms, FP32): first and last band of every cell against the oracle."""
    cells = syn.batch_cells(syn.BATCH_CELLS, np.float32)
    psis, outs = inputs(cells, torch.float32, seed=5)
    bh = BatchedHamiltonian([s.H for *_, s in cells],
                            [nb for w, gd, vt, nb, s in cells], tune=False)
    bh.apply(psis, outs)
    torch.cuda.synchronize()
    for c, (w, gd, vt, nb, s) in enumerate(cells):
        for b in (0, nb - 1):
            x = psis[c][b:b + 1].double().cpu().numpy()
            ref = oracle_hpsi(w, vt, x)
            got = outs[c][b:b + 1].double().cpu().numpy()
            assert abs(got - ref).max() < 2e-5 * abs(ref).max(), (c, b)
