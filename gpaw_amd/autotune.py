"""Tile autotuner for the stencil kernels.

tune_stencil(op, nbands, ...) times the LDS-ring stencil of an FDOperator on
device buffers of the real shape for a set of (rows per tile, prefetch
depth) candidates, keeps the fastest and remembers it per (grid, dtype,
stencil, fused inputs, bands) for the rest of the process.  Timing uses HIP
events on the current stream.
"""
import torch

_CACHE = {}

ROWS = (4, 6, 7, 8, 10, 12, 14, 16, 20, 24, 28, 32)
DEPTHS = (2, 3)


def _key(op, nbands, with_v, with_src):
    return (op.shape, str(op.tdtype), tuple(op.coef_p.round(12)),
            tuple(map(tuple, op.offset_pc)), nbands, with_v, with_src)


def tune_stencil(op, nbands, with_v=False, with_src=False, reps=3,
                 verbose=False):
    """Pick the fastest tile for `op` on `nbands` bands; returns
    (rows, depth, ms)."""
    key = _key(op, nbands, with_v, with_src)
    if key in _CACHE:
        rows, depth, ms = _CACHE[key]
        op.set_tile(rows, 0, depth)
        return _CACHE[key]
    shape = (nbands,) + op.shape
    a = torch.randn(shape, dtype=torch.float64, device='cuda').to(op.tdtype)
    b = torch.empty_like(a)
    vt = None
    if with_v:
        vt = torch.randn(op.shape, dtype=torch.float64, device='cuda')
        vt = vt.to(torch.float32 if op.tdtype in (torch.float32,
                                                  torch.complex64)
                   else torch.float64)
    src = torch.empty_like(a) if with_src else None
    best = None
    tried = set()
    for depth in DEPTHS:
        for rows in ROWS:
            if rows > op.shape[1]:
                continue
            op.set_tile(rows, 0, depth)
            try:
                op.apply(a, b, vt_g=vt, src_xg=src)
            except Exception:
                continue
            ev0 = torch.cuda.Event(enable_timing=True)
            ev1 = torch.cuda.Event(enable_timing=True)
            ev0.record()
            for _ in range(reps):
                op.apply(a, b, vt_g=vt, src_xg=src)
            ev1.record()
            ev1.synchronize()
            ms = ev0.elapsed_time(ev1) / reps
            if verbose:
                print('tile rows=%d depth=%d: %.3f ms' % (rows, depth, ms))
            tried.add((rows, depth))
            if best is None or ms < best[2]:
                best = (rows, depth, ms)
    if best is None:
        op.set_tile(0, 0, 0)
        return None
    op.set_tile(best[0], 0, best[1])
    _CACHE[key] = best
    del a, b, src, vt
    return best


_BATCH_CACHE = {}


def tune_batch(ops, nbands, with_v=True, reps=3, verbose=False):
    """Tiles for one persistent launch over several cells
    (gpaw_amd.batch.StencilBatch): times the batched launch on device
    buffers of the real shapes for (a) every cell's own fastest tile
    (tune_stencil), (b) the same rows at prefetch depth 2, and (c, d) the
    largest tile each cell's ring allows at depth 2 and 3 (rows = 0), and
    returns (rows, depths, ms, label) of the fastest.  One block shape
    serves every tile of a batch, so the tiles that are fastest alone are
    not necessarily the fastest together."""
    from .batch import StencilBatch
    key = tuple(_key(op, nb, with_v, False) for op, nb in zip(ops, nbands))
    if key in _BATCH_CACHE:
        return _BATCH_CACHE[key]
    own = [tune_stencil(op, nb, with_v=with_v) for op, nb in zip(ops, nbands)]
    n = len(ops)
    cands = []
    if all(o is not None for o in own):
        cands.append(('own tiles', [o[0] for o in own], [o[1] for o in own]))
        cands.append(('own rows, depth 2', [o[0] for o in own], [2] * n))
    cands.append(('largest tiles, depth 2', [0] * n, [2] * n))
    cands.append(('largest tiles, depth 3', [0] * n, [3] * n))
    ins, outs, vts = [], [], []
    for op, nb in zip(ops, nbands):
        a = torch.randn((nb,) + op.shape, dtype=torch.float64,
                        device='cuda').to(op.tdtype)
        ins.append(a)
        outs.append(torch.empty_like(a))
        vts.append(torch.randn(op.shape, dtype=torch.float64, device='cuda')
                   .to(torch.float32 if op.tdtype == torch.float32
                       else torch.float64) if with_v else None)
    best = None
    for label, rows, depths in cands:
        try:
            sb = StencilBatch(ops, nbands, rows, depths, with_potential=with_v)
            sb.apply(ins, outs, vts)
        except Exception:
            continue
        ev0 = torch.cuda.Event(enable_timing=True)
        ev1 = torch.cuda.Event(enable_timing=True)
        ev0.record()
        for _ in range(reps):
            sb.apply(ins, outs, vts)
        ev1.record()
        ev1.synchronize()
        ms = ev0.elapsed_time(ev1) / reps
        if verbose:
            print('batch %s rows=%s depths=%s: %.3f ms'
                  % (label, sb.rows, sb.depths, ms))
        if best is None or ms < best[2]:
            best = (list(sb.rows), list(sb.depths), ms, label)
        del sb
    del ins, outs, vts
    _BATCH_CACHE[key] = best
    return best
