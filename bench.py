"""Benchmark: band-H-psi/s of the full FD Hamiltonian at 112^3 (FP64).

Workload (one step = one H psi over all bands): 256 real float64 bands on a
112^3 periodic grid (h = 0.2 A), 19-point kinetic stencil fused with the
local potential, projections P = <p|psi> of 64 atoms x 8 projectors,
P dH, and the projector add -- gpaw.hamiltonian.Hamiltonian.apply.  Inputs
are seeded synthetic data already resident in HBM.

Prints one JSON line.  `roofline` describes the dominant kernel (the fused
stencil + potential pass): algorithmic bytes per launch = nbands * 2 * N * 8
(read psi, write H psi) + N * 8 (potential), divided by its mean duration
from HIP events on the launch stream.  `cpu_baseline` times the reference's
own C kernels (oracle/_ref, built from the reference sources) on a bounded
sample of bands on the host cores.

    python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ('band-Hψ/s at 112³ (GB/s vs 8 TB/s); RMM-DIIS time to '
          '1e-8, NN vs MG precond')
HBM_PEAK_GBS = 8000.0
# HBM bytes per launch of the fused stencil at the default workload (256
# bands, 112^3, FP64), from rocprofv3 FETCH_SIZE / WRITE_SIZE runs of this
# script (--calibrate-copy) scaled by the plain copy of known size in the
# same runs (x2.0000 / x1.0000); the commands and raw values are in
# profiles/full_hpsi_fp64_112cubed_256bands.md.
STENCIL_TRAFFIC_BYTES = {(112, 256): 6.820e9}


def parse():
    p = argparse.ArgumentParser()
    p.add_argument('--gpus', type=int, default=1)
    p.add_argument('--steps', type=int, default=20)
    p.add_argument('--warmup', type=int, default=3)
    p.add_argument('--nbands', type=int, default=256)
    p.add_argument('--grid', type=int, default=112)
    p.add_argument('--natoms', type=int, default=64)
    p.add_argument('--cpu-bands', type=int, default=32)
    p.add_argument('--no-cpu', action='store_true')
    p.add_argument('--calibrate-copy', action='store_true',
                   help='also run 3 plain device copies of one band array '
                        '(known bytes) to calibrate HBM counters')
    p.add_argument('--rmm-sweeps', type=int, default=30,
                   help='RMM-DIIS sweeps (0: skip the eigensolver leg)')
    p.add_argument('--dav-iters', type=int, default=30,
                   help='Davidson iterates (0: skip the Davidson leg)')
    p.add_argument('--rmm-converge', type=int, default=64,
                   help='lowest bands whose residuals make up the error')
    p.add_argument('--no-scf-next', action='store_true',
                   help='skip the pseudo-density / Poisson leg')
    p.add_argument('--no-transformers', action='store_true',
                   help='skip the restrict / interpolate leg')
    p.add_argument('--no-learned', action='store_true',
                   help='skip the learned-vs-multigrid preconditioner leg')
    p.add_argument('--no-batched', action='store_true',
                   help='skip the batched uneven cells leg')
    p.add_argument('--no-complex64', action='store_true',
                   help='skip the complex64 stencil / transformer / V-cycle '
                        'leg')
    return p.parse_args()


def build_workload(N, natoms, nbands):
    import torch
    from gpaw_amd import synthetic as syn
    from gpaw_amd.fd_operators import Laplace
    from gpaw_amd.grid_descriptor import GridDescriptor
    from gpaw_amd.hamiltonian import FDHamiltonian
    from gpaw_amd.lfc import LFC
    from gpaw_amd.spline import Spline
    w = syn.full_hpsi_geometry(N, natoms)
    gd = GridDescriptor((N,) * 3, (w.L,) * 3)
    pt = LFC(gd, [[Spline.from_data(l, rc, f) for l, rc, f in p]
                  for p in w.proj])
    pt.set_positions(w.spos_ac)
    vt = syn.potential(gd, w.pos_av)
    kin = Laplace(gd, -0.5, 3)
    H = FDHamiltonian(gd, kin, vt, pt, w.dH)
    g = torch.Generator(device='cuda')
    g.manual_seed(1)
    psi = torch.randn((nbands, N, N, N), dtype=torch.float64, device='cuda',
                      generator=g)
    out = torch.empty_like(psi)
    return w, gd, pt, vt, kin, H, psi, out


def _cpu_model():
    try:
        with open('/proc/cpuinfo') as fh:
            for line in fh:
                if line.startswith('model name'):
                    return line.split(':', 1)[1].strip()
    except OSError:
        pass
    return None


def _set_omp_threads(n):
    """Thread count of the OpenMP runtime the reference C build uses."""
    import ctypes
    try:
        ctypes.CDLL('libgomp.so.1').omp_set_num_threads(int(n))
        return True
    except OSError:
        return False


def _median_time(fn, min_runs=5, budget=8.0):
    fn()   # warm-up
    times = []
    t_end = time.perf_counter() + budget
    while len(times) < min_runs or (time.perf_counter() < t_end and
                                    len(times) < 9):
        t0 = time.perf_counter()
        fn()
        times.append(time.perf_counter() - t0)
    return float(np.median(times)), len(times)


def cpu_baseline(w, gd_args, vt, nb_sample):
    """The reference's own C kernels (oracle/_ref, built from its sources)
    on the host cores, at one thread (the reference's default,
    gpaw/__init__.py:53-54) and at every core this process may use: full
    H psi (Operator.apply, the NumPy potential multiply, LFC.integrate,
    P dH, LFC.add) on a sample of the bands, and per band the V-cycle
    (fd_precond) and the nn = 3 restriction / interpolation 112^3 <-> 56^3.
    Median of >= 5 runs after a warm-up."""
    from oracle import refbuild, grid as og, spline as osp
    from oracle.stencil import laplace_stencil
    if not refbuild.available():
        return None
    N, L = gd_args
    ogd = og.Grid((N,) * 3, (L,) * 3)
    ogd1 = ogd.coarsen()
    c, o = laplace_stencil(ogd.h_cv, -0.5, 3)
    op = refbuild.Operator(c, o, ogd, True)
    spl = [[osp.Spline(l, rc, f) for l, rc, f in p] for p in w.proj]
    lfc = refbuild.LFC(ogd, spl, w.spos_ac)
    rst = refbuild.Transformer(ogd, ogd1, 3, True)
    itp = refbuild.Transformer(ogd1, ogd, 3, True)
    rng = np.random.default_rng(1)
    try:
        ncores = len(os.sched_getaffinity(0))
    except AttributeError:
        ncores = os.cpu_count()
    legs = {}
    hpsi = {}
    for threads in (1, ncores):
        if not _set_omp_threads(threads):
            return None
        nbs = max(2, nb_sample // 4) if threads == 1 else nb_sample
        psi = rng.standard_normal((nbs, N, N, N))
        out = np.empty_like(psi)

        def step():
            op.apply(psi, out, None)
            out.__iadd__(psi * vt)
            P = lfc.integrate(psi)
            Q = np.empty_like(P)
            for a in range(w.natoms):
                Q[:, 8 * a:8 * a + 8] = P[:, 8 * a:8 * a + 8] @ w.dH[a]
            lfc.add(out, Q)
        t, nrun = _median_time(step)
        hpsi[threads] = (nbs / t, nbs, nrun, t)
        nv = 4
        R = rng.standard_normal((nv, N, N, N))
        tv, _ = _median_time(lambda: refbuild.fd_precond(ogd, c, o, R))
        coarse = np.empty((nv,) + tuple(ogd1.n_c))
        tr, _ = _median_time(lambda: rst.apply(R, coarse, None))
        ti, _ = _median_time(lambda: itp.apply(coarse, R, None))
        legs['%d_threads' % threads] = {
            'vcycle_ms_per_band': round(1e3 * tv / nv, 3),
            'restrict_nn3_ms_per_band': round(1e3 * tr / nv, 3),
            'interpolate_nn3_ms_per_band': round(1e3 * ti / nv, 3)}
        del psi, out
    v_all, nb_all, nrun, t_all = hpsi[ncores]
    v_one, nb_one, nrun1, t_one = hpsi[1]
    return {'value': round(v_all, 2), 'unit': 'band-Hψ/s',
            'cores': ncores, 'kind': 'reference',
            'cpu_model': _cpu_model(),
            'single_thread_value': round(v_one, 2),
            'sample': ('full H psi on %d of the 256 bands (%d at 1 thread), '
                       'same 112^3 grid, 64 atoms; reference c/operators.c '
                       '+ c/lfc.c (OpenMP over bands) + NumPy potential / '
                       'dH; median of %d runs (%.2f s each at %d threads, '
                       '%.2f s at 1 thread)' % (nb_all, nb_one, nrun, t_all,
                                                ncores, t_one)),
            'legs': legs}


def fp64_ritz_values(sys64, psit, dv):
    """Ritz values of the span of psit under the FP64 Hamiltonian: psit
    promoted to FP64, the full H psi (projector terms included) in FP64, S
    and H from gh_gram (FP64), the generalized problem by Cholesky
    reduction."""
    import torch
    from gpaw_amd import blas
    tdt = torch.complex128 if psit.is_complex() else torch.float64
    x = psit.to(tdt)
    hx = torch.empty_like(x)
    sys64.H.apply(x, hx, sys64.phase_cd)      # T + v + the projector terms
    f, hf = x.reshape(len(x), -1), hx.reshape(len(x), -1)
    Hm = blas.gram(f, hf, alpha=dv, hermitian=True, conj=True)
    Sm = blas.gram(f, alpha=dv, hermitian=True, conj=True)
    del x, hx
    L = torch.linalg.cholesky(Sm)
    Li = torch.linalg.solve_triangular(
        L, torch.eye(len(Sm), dtype=Sm.dtype, device='cuda'), upper=False)
    A = Li @ Hm @ Li.conj().T
    return torch.linalg.eigvalsh(0.5 * (A + A.conj().T)).cpu().numpy()


def eigensolver_run(w, gd, vt, nb, nconv, psi0, sweeps, dtype=np.float64,
                    k_c=None, solver='rmm-diis', tol=1e-8, polish=0,
                    polish_tol=1e-13, final_fp64=None, make_pc=None,
                    sysm=None, warm=2):
    """RMM-DIIS (niter=3, all bands in one block) or Davidson (niter=2) with
    the multigrid preconditioner, wave functions and grid work in `dtype`
    (float64 / complex128: the plain FP64 form; float32 / complex64: the
    mixed form, whose subspace matrices are FP32 products with FP64 sums and
    whose eigh, Cholesky and reductions are FP64), from the start psi0
    (converted to dtype).  Error as the reference's (sum_n w_n int |R_n|^2
    dv) with w_n = 1 for the lowest `nconv` bands, divided by nconv; stops
    at err <= tol.  The timed region runs from the first orthonormalisation
    to the sweep that reaches tol.  polish > 0: up to that many further
    sweeps (untimed) while the error exceeds polish_tol or still falls, for
    the eigenvalues at the form's own floor.

    Before the timed run, `warm` sweeps on a copy of the start (a separate
    solver object) load the GPU code objects and library handles this
    precision uses for the first time in the process (FP32 / complex
    GEMMs, eigh) and fill the allocator's cache; the timed run then
    measures the solver's steady cost, as in an SCF loop.  The first,
    cold, pass is timed too (`seconds_cold`, the same iterations)."""
    import torch
    from gpaw_amd import synthetic as syn
    from gpaw_amd._arrays import torch_dtype
    from gpaw_amd.eigensolvers import Davidson, RMMDIIS, FDWaveFunctions
    if sysm is None:
        sysm = syn.System(w, gd, vt, dtype, k_c)
    pc = sysm.pc if make_pc is None else make_pc(sysm)
    weights = torch.zeros(nb, dtype=torch.float64, device='cuda')
    weights[:nconv] = 1.0
    t_cold = None
    if warm:
        wfs = FDWaveFunctions(gd, psi0.to(torch_dtype(dtype)).contiguous(),
                              sysm.pt, pc, sysm.phase_cd)
        es = RMMDIIS() if solver == 'rmm-diis' else Davidson()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        wfs.orthonormalize(sysm.H)
        for it in range(warm):
            es.iterate(sysm.H, wfs, weights)
        torch.cuda.synchronize()
        t_cold = time.perf_counter() - t0
        del wfs, es
        if hasattr(pc, 'reset'):
            pc.reset()
    wfs = FDWaveFunctions(gd, psi0.to(torch_dtype(dtype)).contiguous(),
                          sysm.pt, pc, sysm.phase_cd)
    es = RMMDIIS() if solver == 'rmm-diis' else Davidson()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    wfs.orthonormalize(sysm.H)
    errs, t_tol, t_first = [], None, None
    for it in range(sweeps):
        err = es.iterate(sysm.H, wfs, weights) / nconv
        errs.append(err)
        if it == warm - 1:
            t_first = time.perf_counter() - t0   # err is on the host: synced
        if err <= tol:
            torch.cuda.synchronize()
            t_tol = time.perf_counter() - t0
            break
    torch.cuda.synchronize()
    t = time.perf_counter() - t0
    eps_tol = wfs.eps[:nconv].cpu().numpy()
    eps_rr = None
    if final_fp64 is not None and t_tol is not None:
        # the mixed form's eigenvalues from one FP64 Rayleigh-Ritz step on
        # its converged wave functions (FP64 H psi, FP64 S and H), timed
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        eps_rr = fp64_ritz_values(final_fp64, wfs.psit, gd.dv)[:nconv]
        torch.cuda.synchronize()
        t_rr = time.perf_counter() - t1
    perrs = []
    for it in range(polish):
        err = es.iterate(sysm.H, wfs, weights) / nconv
        perrs.append(err)
        if err <= polish_tol or (len(perrs) > 3 and err > 0.5 * perrs[-4]):
            break
    res = {'solver': 'RMMDIIS(niter=3, all bands in one block)'
           if solver == 'rmm-diis' else repr(es),
           'dtype': str(np.dtype(dtype)),
           'precond': ('multigrid V-cycle (%s)' % np.dtype(dtype)
                       if make_pc is None else type(pc).__name__),
           'bands': nb, 'bands_converged': nconv,
           'sweeps': len(errs), 'seconds': round(t, 3),
           'ms_per_sweep': round(1e3 * t / max(len(errs), 1), 2),
           'errors': [float('%.3e' % e) for e in errs],
           'seconds_to_1e-8': None if t_tol is None else round(t_tol, 3),
           'eps_lowest': float(wfs.eps[0])}
    if t_cold is not None:
        # the same start and first sweeps in a process where this
        # precision's kernels had not run yet
        res['warm_sweeps'] = warm
        res['seconds_cold_first_sweeps'] = round(t_cold, 3)
        res['seconds_warm_first_sweeps'] = (None if t_first is None
                                            else round(t_first, 3))
    if polish:
        res['polish_errors'] = [float('%.3e' % e) for e in perrs]
    if eps_rr is not None:
        res['fp64_rayleigh_ritz_ms'] = round(1e3 * t_rr, 2)
        res['seconds_to_1e-8_with_fp64_ritz'] = round(t_tol + t_rr, 3)
    if hasattr(pc, 'fallback_active'):
        res['learned_calls'] = pc.calls
        res['fallback_to_vcycle'] = bool(pc.fallback_active)
    eps = (eps_tol, wfs.eps[:nconv].cpu().numpy(), eps_rr)
    del wfs, sysm, es, pc
    torch.cuda.empty_cache()
    return res, eps


def eigensolver_leg(w, gd, vt, H, nb, nconv, sweeps, k_c=None,
                    davidson=False):
    """The second metric's plain-FP64-multigrid and mixed-precision RMM-DIIS
    runs (and optionally Davidson) from one LCAO-like start (Rayleigh-Ritz
    in atom-centred functions, as an SCF starts): time to err <= 1e-8 and
    the mixed form's eigenvalue error against the FP64 run (the FP64 path
    is pinned to the oracle by the GPU parity tests)."""
    import torch
    from gpaw_amd import synthetic as syn
    cplx = k_c is not None
    d64, d32 = (np.complex128, np.complex64) if cplx else (np.float64,
                                                           np.float32)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    if cplx:
        Hk = syn.System(w, gd, vt, d64, k_c).H
        psi0 = syn.lcao_start(gd, w, Hk, nb, dtype=complex, k_c=k_c)
        del Hk
    else:
        psi0 = syn.lcao_start(gd, w, H, nb)
    torch.cuda.synchronize()
    t_start = time.perf_counter() - t0
    out = {'start': 'LCAO-like: lowest Rayleigh-Ritz vectors in 12 '
                    'atom-centred functions per atom (projector shapes + s/p '
                    'Gaussians), %.2f s, not in the timed region' % t_start,
           'k_c': list(k_c) if cplx else [0, 0, 0]}
    r64, e64 = eigensolver_run(w, gd, vt, nb, nconv, psi0, sweeps, d64, k_c,
                               polish=30)
    sys64 = syn.System(w, gd, vt, d64, k_c)
    r32, e32 = eigensolver_run(w, gd, vt, nb, nconv, psi0, sweeps, d32, k_c,
                               polish=20, final_fp64=sys64)
    del sys64
    ref = e64[1]   # FP64 polished to its floor: the eigenvalue reference
    r64['eigenvalue_max_abs_err_at_1e-8_Ha'] = float(abs(e64[0] - ref).max())
    r32['eigenvalue_max_abs_err_at_1e-8_Ha'] = float(abs(e32[0] - ref).max())
    r32['eigenvalue_max_abs_err_at_floor_Ha'] = float(abs(e32[1] - ref).max())
    d = abs(e32[0] - ref)
    r32['eigenvalue_err_band_of_max'] = int(d.argmax())
    r32['eigenvalue_err_median_Ha'] = float(np.median(d))
    if e32[2] is not None:
        r32['eigenvalue_max_abs_err_fp64_ritz_at_1e-8_Ha'] = float(
            abs(e32[2] - ref).max())
    if r64['seconds_to_1e-8'] and r32.get('seconds_to_1e-8_with_fp64_ritz'):
        out['speedup_mixed_with_fp64_ritz_vs_fp64'] = round(
            r64['seconds_to_1e-8'] / r32['seconds_to_1e-8_with_fp64_ritz'], 2)
    out['eigenvalue_reference'] = ('the FP64 run continued to its floor '
                                   '(GPU FP64 path, parity-pinned to the '
                                   'oracle)')
    out['fp64_multigrid'] = r64
    out['mixed_fp32'] = r32
    if r64['seconds_to_1e-8'] and r32['seconds_to_1e-8']:
        out['speedup_mixed_vs_fp64'] = round(
            r64['seconds_to_1e-8'] / r32['seconds_to_1e-8'], 2)
    if davidson:
        rd, ed = eigensolver_run(w, gd, vt, nb, nconv, psi0, sweeps, d64,
                                 k_c, solver='davidson')
        rd['eigenvalue_max_abs_err_at_1e-8_Ha'] = float(abs(ed[0] - ref).max())
        out['davidson_fp64'] = rd
    del psi0
    torch.cuda.empty_cache()
    return out


def learned_leg(weights, N=144, natoms=128, nb=512, nconv=128, sweeps=30):
    """The second metric at 144^3 with 512 bands and 128 atoms (generated
    as the 112^3 workload): FP64 RMM-DIIS to err <= 1e-8 from one LCAO-like
    start with the multigrid V-cycle and with the learned preconditioner
    (U-Net on the matrix cores, the V-cycle as fallback judged by the
    residual norm).  Error over the lowest `nconv` bands."""
    import torch
    from gpaw_amd import synthetic as syn
    from gpaw_amd.grid_descriptor import GridDescriptor
    from gpaw_amd.learned import LearnedPreconditioner
    w = syn.full_hpsi_geometry(N, natoms)
    gd = GridDescriptor((N,) * 3, (w.L,) * 3)
    vt = syn.potential(gd, w.pos_av)
    sysm = syn.System(w, gd, vt, np.float64)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    psi0 = syn.lcao_start(gd, w, sysm.H, nb)
    torch.cuda.synchronize()
    out = {'workload': '%d bands, %d^3 grid, %d atoms, FP64' % (nb, N,
                                                                 natoms),
           'start': 'LCAO-like, %.2f s, not in the timed region'
                    % (time.perf_counter() - t0),
           'weights': os.path.relpath(weights, ROOT)}
    torch.cuda.empty_cache()
    rmg, emg = eigensolver_run(w, gd, vt, nb, nconv, psi0, sweeps,
                               sysm=sysm)
    torch.cuda.empty_cache()

    def make(s):
        return LearnedPreconditioner(gd, s.kin, vt, weights, fallback=s.pc)
    rnn, enn = eigensolver_run(w, gd, vt, nb, nconv, psi0, sweeps,
                               sysm=sysm, make_pc=make)
    rnn['eigenvalue_max_abs_diff_vs_multigrid_Ha'] = float(
        abs(enn[0] - emg[0]).max())
    out['multigrid'] = rmg
    out['learned'] = rnn
    if rmg['seconds_to_1e-8'] and rnn['seconds_to_1e-8']:
        out['speedup_learned_vs_multigrid'] = round(
            rmg['seconds_to_1e-8'] / rnn['seconds_to_1e-8'], 2)
    # cost of one preconditioner application on all bands
    lp = make(sysm)
    R = torch.randn((nb,) + tuple(gd.n_c), dtype=torch.float64,
                    device='cuda')
    out['vcycle_ms_all_bands'] = round(_time_ms(lambda: sysm.pc(R), 3), 2)
    out['unet_ms_all_bands'] = round(_time_ms(lambda: lp(R), 3), 2)
    del psi0, R, lp, sysm
    torch.cuda.empty_cache()
    return out


def dense_leg(nb, N):
    """The subspace contractions of a sweep at the workload shape, timed
    alone: gh_gram (lower triangle on the matrix cores) against torch's
    full product, and the rotation U psi (rocBLAS), per precision."""
    import torch
    from gpaw_amd import blas
    out = {'peak_TFLOPs': {'fp64_mfma': 78.6, 'fp32_mfma': 157.3}}
    for name, dt in (('f64', torch.float64), ('f32', torch.float32),
                     ('c64', torch.complex64), ('c128', torch.complex128)):
        a = torch.randn((nb, N), dtype=dt, device='cuda')
        U = torch.randn((nb, nb), dtype=dt, device='cuda')
        cf = 8 if a.is_complex() else 2
        flop = cf * nb * nb * N
        t_g = _time_ms(lambda: blas.gram(a, hermitian=True))
        t_r = _time_ms(lambda: U @ a)
        rec = {'gram_lower_ms': round(t_g, 3),
               'gram_TFLOPs_executed': round(0.75 * flop / t_g / 1e9, 1),
               'rotation_ms': round(t_r, 3),
               'rotation_TFLOPs': round(flop / t_r / 1e9, 1)}
        if not a.is_complex():
            t_t = _time_ms(lambda: a @ a.T)
            rec['torch_full_product_ms'] = round(t_t, 3)
        out[name] = rec
        del a
        torch.cuda.empty_cache()
    return out


def batched_leg():
    """Six independent cells of uneven size (40^3 .. 128^3, 16 .. 512 real
    bands, 8 .. 96 atoms; FP32 with FP64 reductions): the kinetic stencils
    with the local potential of all cells in one persistent launch (tiles
    from the batch autotuner), then each cell's projections, P dH and
    projector add; the whole step captured in a HIP graph.  Compared with
    one launch per cell, each with its own fastest tile.  Algorithmic bytes of the
    batched stencil: sum over cells of nbands * 2 * N * 4 + N * 4."""
    import torch
    from gpaw_amd import synthetic as syn
    from gpaw_amd.batch import BatchedHamiltonian
    cells = syn.batch_cells()
    nbs = [nb for w, gd, vt, nb, s in cells]
    bh = BatchedHamiltonian([s.H for *_, s in cells], nbs, tune=True)
    g = torch.Generator(device='cuda')
    g.manual_seed(1)
    psis = [torch.randn((nb,) + tuple(gd.n_c), dtype=torch.float32,
                        device='cuda', generator=g)
            for w, gd, vt, nb, s in cells]
    outs = [torch.empty_like(p) for p in psis]
    vts = [s.H.vt_G for *_, s in cells]
    alg = sum(nb * 2 * int(np.prod(gd.n_c)) * 4 + int(np.prod(gd.n_c)) * 4
              for w, gd, vt, nb, s in cells)
    t_stencil = _time_ms(lambda: bh.batch.apply(psis, outs, vts), reps=10)

    def separate():
        for c, (w, gd, vt, nb, s) in enumerate(cells):
            s.H.apply_local(psis[c], outs[c])
    from gpaw_amd.autotune import tune_stencil
    own = [tune_stencil(s.kin, nb, with_v=True)
           for w, gd, vt, nb, s in cells]       # each cell's fastest tile
    t_sep = _time_ms(separate, reps=10)
    for c, (w, gd, vt, nb, s) in enumerate(cells):
        s.kin.set_tile(bh.batch.rows[c], 0, bh.batch.depths[c])
    t_sep_same = _time_ms(separate, reps=10)
    t_eager = _time_ms(lambda: bh.apply(psis, outs), reps=10)
    bh.capture(psis, outs)
    t_graph = _time_ms(bh.replay, reps=10)
    nbt = sum(nbs)
    return {'cells': [{'grid': [int(gd.n_c[0])] * 3, 'bands': nb,
                       'atoms': w.natoms, 'tile_rows': bh.batch.rows[c],
                       'tile_depth': bh.batch.depths[c]}
                      for c, (w, gd, vt, nb, s) in enumerate(cells)],
            'dtype': 'float32 (projections and P dH summed in FP64)',
            'tiles_from': 'gpaw_amd.autotune.tune_batch: %s' % (
                bh.tuned[3] if bh.tuned else 'default'),
            'own_tiles': [[o[0], o[1]] if o else None for o in own],
            'launch': {'blocks': bh.batch.blocks,
                       'threads': bh.batch.threads,
                       'lds_bytes': bh.batch.lds, 'tiles': bh.batch.tiles,
                       'rounds': bh.batch.rounds},
            'stencil_batched_ms': round(t_stencil, 4),
            'stencil_batched_GBps': round(alg / t_stencil / 1e6, 1),
            'stencil_batched_frac': round(alg / t_stencil / 1e6 /
                                          HBM_PEAK_GBS, 4),
            'stencil_one_launch_per_cell_ms': round(t_sep, 4),
            'stencil_one_launch_per_cell_batch_tiles_ms': round(t_sep_same,
                                                                4),
            'hpsi_step_eager_ms': round(t_eager, 4),
            'hpsi_step_graph_ms': round(t_graph, 4),
            'band_hpsi_per_s_graph': round(nbt / (t_graph * 1e-3), 1)}


def _time_ms(fn, reps=5):
    import torch
    fn()
    torch.cuda.synchronize()
    out = []
    for _ in range(reps):
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        fn()
        e1.record()
        out.append((e0, e1))
    torch.cuda.synchronize()
    return float(np.median([a.elapsed_time(b) for a, b in out]))


def complex64_leg(gd, nb):
    """Complex64 kernels on 256 complex residuals at 112^3 with the Bloch
    phase of k = (0.25, 0, 0): 19-point stencil, nn=3 restrict and
    interpolate, and the reference's V-cycle (nn=1 transformers, two coarse
    levels), with the complex128 V-cycle on the same input for the error
    of the single-precision form."""
    import torch
    from gpaw_amd.fd_operators import Laplace
    from gpaw_amd.preconditioner import Preconditioner
    from gpaw_amd.transformers import Transformer
    N = int(np.prod(gd.n_c))
    ph = gd.phase_cd((0.25, 0.0, 0.0))
    g = torch.Generator(device='cuda')
    g.manual_seed(1)
    R = torch.randn((nb,) + tuple(gd.n_c), dtype=torch.complex64,
                    device='cuda', generator=g)
    out = torch.empty_like(R)
    kin = Laplace(gd, -0.5, 3, np.complex64)
    gd1 = gd.coarsen()
    coarse = gd1.empty(nb, np.complex64)
    rst = Transformer(gd, gd1, 3, np.complex64)
    itp = Transformer(gd1, gd, 3, np.complex64)

    class K:
        phase_cd = ph
    pc = Preconditioner(gd, kin, np.complex64)
    res = {'dtype': 'c64 (FP32 arithmetic)', 'bands': nb, 'k_c': [0.25, 0, 0]}
    s = 8
    t = _time_ms(lambda: kin.apply(R, out, ph))
    res['stencil_ms'] = round(t, 4)
    res['stencil_GBps'] = round(nb * 2 * N * s / t / 1e6, 1)
    t = _time_ms(lambda: rst.apply(R, coarse, ph))
    res['restrict_nn3_ms'] = round(t, 4)
    res['restrict_nn3_GBps'] = round(nb * (N + N // 8) * s / t / 1e6, 1)
    t = _time_ms(lambda: itp.apply(coarse, out, ph))
    res['interpolate_nn3_ms'] = round(t, 4)
    res['interpolate_nn3_GBps'] = round(nb * (N + N // 8) * s / t / 1e6, 1)
    t = _time_ms(lambda: pc(R, K, out=out))
    res['vcycle_ms'] = round(t, 4)
    res['vcycle_GBps_compulsory'] = round(nb * 2 * N * s / t / 1e6, 1)
    del coarse
    # complex128 reference form on the same residuals
    nref = min(nb, 32)
    R2 = R[:nref].to(torch.complex128)
    kin2 = Laplace(gd, -0.5, 3, complex)
    pc2 = Preconditioner(gd, kin2, complex)
    out2 = torch.empty_like(R2)
    t = _time_ms(lambda: pc2(R2, K, out=out2))
    res['vcycle_c128_ms_per_band'] = round(t / nref, 5)
    res['vcycle_c64_ms_per_band'] = round(res['vcycle_ms'] / nb, 5)
    pc(R[:nref].contiguous(), K, out=out[:nref])
    err = (out[:nref].to(torch.complex128) - out2).abs().max() / \
        out2.abs().max()
    res['vcycle_c64_relerr_vs_c128'] = float(err)
    return res


def transformer_leg(gd, nb):
    """Restriction 112^3 -> 56^3 and interpolation back, 256 bands, for
    nn = 1 (the V-cycle's) and nn = 3 (density / potential) in float64,
    complex64 and complex128 (complex with the Bloch phase of
    k = (0.25, 0, 0)).  GB/s of the algorithmic bytes (N + N/8) * s per
    band."""
    import torch
    from gpaw_amd.transformers import Transformer
    N = int(np.prod(gd.n_c))
    gd1 = gd.coarsen()
    out = {}
    for name, dt, tdt in (('f64', np.float64, torch.float64),
                          ('c64', np.complex64, torch.complex64),
                          ('c128', np.complex128, torch.complex128)):
        ph = None if dt == np.float64 else gd.phase_cd((0.25, 0.0, 0.0))
        fine = torch.randn((nb,) + tuple(gd.n_c), dtype=tdt, device='cuda')
        coarse = gd1.empty(nb, dt)
        s = fine.element_size()
        for nn in (1, 3):
            rst = Transformer(gd, gd1, nn, dt)
            itp = Transformer(gd1, gd, nn, dt)
            tr = _time_ms(lambda: rst.apply(fine, coarse, ph))
            ti = _time_ms(lambda: itp.apply(coarse, fine, ph))
            byt = nb * (N + N // 8) * s
            out['%s_nn%d' % (name, nn)] = {
                'restrict_ms': round(tr, 4),
                'restrict_GBps': round(byt / tr / 1e6, 1),
                'interpolate_ms': round(ti, 4),
                'interpolate_GBps': round(byt / ti / 1e6, 1)}
        del fine, coarse
    return out


def scf_next_leg(w, gd, psi):
    """The steps after the eigensolver in an SCF iteration, on the same
    workload: the pseudo-density nt = sum_n f_n |psi_n|^2 over the 256
    bands (FP64), and the multigrid Poisson solver (nn = 3, Jacobi) on the
    fine grid (2N)^3 for a neutral charge: one Gaussian per atom
    (exp(-r^2 / (1 Bohr)^2), minimum image) minus its mean."""
    import torch
    from gpaw_amd.density import add_to_density_bands
    from gpaw_amd.grid_descriptor import GridDescriptor
    from gpaw_amd.poisson import FDPoissonSolver
    nb = psi.shape[0]
    N = int(np.prod(gd.n_c))
    res = {}
    nt = torch.zeros(tuple(gd.n_c), dtype=torch.float64, device='cuda')
    f = torch.full((nb,), 2.0, dtype=torch.float64, device='cuda')
    t = _time_ms(lambda: add_to_density_bands(f, psi, nt))
    res['density_ms'] = round(t, 4)
    res['density_GBps'] = round((nb + 2) * N * 8 / t / 1e6, 1)
    del nt
    nf = 2 * int(gd.n_c[0])
    gdf = GridDescriptor((nf,) * 3, (w.L,) * 3)
    h = w.L / nf
    ax = torch.arange(nf, dtype=torch.float64, device='cuda') * h
    rho = torch.zeros((nf,) * 3, dtype=torch.float64, device='cuda')
    for p in w.pos_av:
        d = []
        for c in range(3):
            x = ax - float(p[c])
            d.append(x - w.L * torch.round(x / w.L))
        rho += torch.exp(-(d[0][:, None, None]**2 + d[1][None, :, None]**2 +
                           d[2][None, None, :]**2))
    rho -= rho.mean()
    ps = FDPoissonSolver(nn=3)
    ps.set_grid_descriptor(gdf)
    phi = torch.zeros_like(rho)
    ts = []
    for rep in range(3):      # the first solve also records the V-cycle
        phi.zero_()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        niter = ps.solve_neutral(phi, rho)
        torch.cuda.synchronize()
        ts.append(time.perf_counter() - t0)
    ps_e = FDPoissonSolver(nn=3, graph=False)
    ps_e.set_grid_descriptor(gdf)
    ps_e.solve_neutral(phi.zero_(), rho)
    phi.zero_()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ps_e.solve_neutral(phi, rho)
    torch.cuda.synchronize()
    t_eager = time.perf_counter() - t0
    res['poisson_grid'] = [nf] * 3
    res['poisson_levels'] = ps.levels
    res['poisson_vcycles_to_2e-10'] = niter
    res['poisson_solve_ms'] = round(1e3 * min(ts[1:]), 2)
    res['poisson_solve_first_ms'] = round(1e3 * ts[0], 2)
    res['poisson_solve_eager_ms'] = round(1e3 * t_eager, 2)
    res['poisson_vcycle_ms'] = round(_time_ms(lambda: ps.iterate2(ps.step)), 3)
    op = ps.operators[0]
    t = _time_ms(lambda: op.relax(2, phi, rho, 1, 2.0 / 3.0))
    res['jacobi_sweep_ms'] = round(t, 4)
    res['jacobi_sweep_GBps'] = round(3 * nf**3 * 8 / t / 1e6, 1)
    return res


def main():
    args = parse()
    if args.gpus != 1:
        print('bench.py measures one GPU: --gpus %d is not supported (the '
              'Hamiltonian would shard by bands, one share per GPU)' % args.gpus,
              file=sys.stderr)
        sys.exit(2)
    import torch
    assert torch.cuda.is_available(), 'needs a GPU'
    N, natoms, nb = args.grid, args.natoms, args.nbands
    w, gd, pt, vt, kin, H, psi, out = build_workload(N, natoms, nb)
    ngpts = N ** 3
    stream = torch.cuda.current_stream()

    for _ in range(args.warmup):
        H.apply(psi, out)
    torch.cuda.synchronize()
    if args.calibrate_copy:
        # reads and writes nbands * N^3 * 8 bytes each: the reference
        # kernel for scaling FETCH_SIZE / WRITE_SIZE counters
        for _ in range(3):
            out.copy_(psi)
        torch.cuda.synchronize()

    ev = [(torch.cuda.Event(enable_timing=True),
           torch.cuda.Event(enable_timing=True)) for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for s in range(args.steps):
        ev[s][0].record(stream)
        H.apply_local(psi, out)          # fused stencil + potential
        ev[s][1].record(stream)
        P = pt.integrate_M(psi)
        pt.add_M(out, H.dH(P))
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    elapsed = t1 - t0
    ms_step = 1e3 * elapsed / args.steps
    value = nb * args.steps / elapsed
    stencil_ms = float(np.mean([a.elapsed_time(b) for a, b in ev]))
    alg_bytes = nb * 2 * ngpts * 8 + ngpts * 8
    achieved = alg_bytes / (stencil_ms * 1e-3) / 1e9
    full_bytes_per_band = 2 * ngpts * 8 + (ngpts * 8 + 8 * pt.nentries *
                                           8) / nb
    full_gbs = value * full_bytes_per_band / 1e9

    # the achievable HBM rate: a plain streaming copy of the same array
    # (16-byte vectors, non-temporal), best of a few grid sizes
    from gpaw_amd import _lib
    from gpaw_amd._arrays import stream_ptr
    L = _lib.lib()
    nbytes = psi.numel() * psi.element_size()
    copy_gbs = {}
    for nblk in (1024, 2048, 4096, 8192):
        t = _time_ms(lambda: _lib.check(L.gh_stream_copy(
            psi.data_ptr(), out.data_ptr(), nbytes, nblk, stream_ptr())),
            reps=5)
        copy_gbs[nblk] = 2 * nbytes / t / 1e6
    copy_best = max(copy_gbs.values())

    c3 = None
    if not args.no_complex64:
        del out
        out = None
        c3 = complex64_leg(gd, nb)

    trans = None
    if not args.no_transformers:
        out = None
        trans = transformer_leg(gd, nb)

    batched = None
    if not args.no_batched:
        batched = batched_leg()
        torch.cuda.empty_cache()

    scf_next = None
    if not args.no_scf_next:
        scf_next = scf_next_leg(w, gd, psi)

    rmm = None
    if args.rmm_sweeps > 0:
        out = None
        rmm = eigensolver_leg(w, gd, vt, H, nb, min(args.rmm_converge, nb),
                              args.rmm_sweeps, davidson=args.dav_iters > 0)
        if not args.no_complex64:
            rmm['complex_k'] = eigensolver_leg(
                w, gd, vt, H, nb, min(args.rmm_converge, nb),
                args.rmm_sweeps, k_c=(0.25, 0.0, 0.0))
        rmm['dense'] = dense_leg(nb, ngpts)

    learned = None
    weights = os.path.join(ROOT, 'gpaw_amd', 'data', 'unet_precond.npz')
    if not args.no_learned and os.path.exists(weights):
        out = None
        learned = learned_leg(weights)

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(w, (N, w.L), vt, args.cpu_bands)

    line = {
        'metric': METRIC,
        'value': round(value, 1),
        'unit': 'band-Hψ/s',
        'n_gpus': 1,
        'steps': args.steps,
        'warmup': args.warmup,
        'ms_per_step': round(ms_step, 4),
        'higher_is_better': True,
        'scaling': 'weak',
        'vs_baseline': None,
        'dtype': 'f64',
        'data': 'synthetic (seeded N(0,1) bands, random atoms/projectors/'
                'potential)',
        'config': {'workload': 'full FD H-psi FP64: %d bands, %d^3 periodic '
                               'grid, h=0.2 A, %d atoms x 8 projectors, '
                               '19-point stencil' % (nb, N, natoms),
                   'nbands': nb, 'grid': [N, N, N], 'natoms': natoms,
                   'parallelism': 'single GPU'},
        'full_hpsi_GBps': round(full_gbs, 1),
        'roofline': {'bound': 'hbm', 'achieved': round(achieved, 1),
                     'peak': HBM_PEAK_GBS, 'unit': 'GB/s',
                     'frac': round(achieved / HBM_PEAK_GBS, 4),
                     'traffic': STENCIL_TRAFFIC_BYTES.get((N, nb)),
                     'kernel': 'stencil_ring_kernel<double,R=3,V> '
                               '(T psi + V psi, one pass)',
                     'kernel_ms': round(stencil_ms, 4),
                     'traffic_GBps': (round(STENCIL_TRAFFIC_BYTES[(N, nb)] /
                                            stencil_ms / 1e6, 1)
                                      if (N, nb) in STENCIL_TRAFFIC_BYTES
                                      else None),
                     'copy_GBps': round(copy_best, 1)},
        'rmmdiis': rmm,
        'single_precision_complex_k': c3,
        'transformers': trans,
        'batched_cells': batched,
        'learned_vs_multigrid': learned,
        'density_and_poisson': scf_next,
        'cpu_baseline': cpu,
    }
    print(json.dumps(line))


if __name__ == '__main__':
    main()
