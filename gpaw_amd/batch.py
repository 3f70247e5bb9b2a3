"""Several independent cells of uneven size in one persistent launch.

StencilBatch binds gh_stencil_batch: the kinetic stencils (with the local
potential fused) of up to 8 cells -- each its own grid, band count and
FDOperator -- run as one persistent kernel whose blocks walk a global tile
list.  The tiles of the cells (rows per tile, prefetch depth) come from the
batch autotuner (gpaw_amd.autotune.tune_batch: the batched launch timed for
the cells' own fastest tiles and for the largest tiles each ring allows)
unless given.

BatchedHamiltonian applies H psi = T psi + v psi + sum_a p_a dH_a <p_a|psi>
(gpaw/hamiltonian.py:445-479 per cell) to every cell: the one batched
stencil launch, then each cell's projections, P dH and projector add.  With
single-precision cells the projections and P dH are summed in FP64.
capture() records the whole step in a HIP graph (torch.cuda.CUDAGraph);
replay() then issues it with one call.
"""
import ctypes

import numpy as np
import torch

from . import _lib
from ._arrays import stream_ptr


class StencilBatch:
    def __init__(self, ops, nbands, rows=None, depths=None,
                 with_potential=True):
        if not 1 <= len(ops) <= 8:
            raise ValueError('1..8 cells per batch')
        if len(nbands) != len(ops):
            raise ValueError('one band count per cell')
        self.ops = list(ops)
        self.nbands = [int(n) for n in nbands]
        self.with_potential = bool(with_potential)
        self._lib = _lib.lib()
        n = len(ops)
        hp = (_lib.P * n)(*[op.handle for op in ops])
        nb = np.ascontiguousarray(self.nbands, np.int64)
        r = np.ascontiguousarray(rows if rows is not None else [0] * n,
                                 np.int32)
        d = np.ascontiguousarray(depths if depths is not None else [0] * n,
                                 np.int32)
        h = ctypes.c_void_p()
        _lib.check(self._lib.gh_stencil_batch_create(
            ctypes.byref(h), n, hp, nb.ctypes.data_as(_lib.PI64),
            r.ctypes.data_as(_lib.PI32), d.ctypes.data_as(_lib.PI32),
            int(self.with_potential)))
        self.handle = h
        info = np.zeros(5, np.int64)
        rr = np.zeros(n, np.int32)
        dd = np.zeros(n, np.int32)
        _lib.check(self._lib.gh_stencil_batch_info(
            h, info.ctypes.data_as(_lib.PI64), rr.ctypes.data_as(_lib.PI32),
            dd.ctypes.data_as(_lib.PI32)))
        self.blocks, self.threads, self.lds, self.tiles, self.rounds = (
            int(x) for x in info)
        self.rows = rr.tolist()
        self.depths = dd.tolist()

    def __del__(self):
        h = getattr(self, 'handle', None)
        if h:
            try:
                self._lib.gh_stencil_batch_destroy(h)
            except Exception:
                pass
            self.handle = None

    def apply(self, ins, outs, vts=None):
        """outs[c] = T_c ins[c] (+ vts[c] * ins[c]) for every cell c."""
        n = len(self.ops)
        assert len(ins) == n and len(outs) == n
        for c, op in enumerate(self.ops):
            shape = (self.nbands[c],) + op.shape
            for t in (ins[c], outs[c]):
                if tuple(t.shape) != shape or t.dtype != op.tdtype:
                    raise ValueError('cell %d: array %s %s, expected %s %s'
                                     % (c, tuple(t.shape), t.dtype, shape,
                                        op.tdtype))
                if not (t.is_cuda and t.is_contiguous()):
                    raise ValueError('cell %d: contiguous device arrays only'
                                     % c)
            if self.with_potential and tuple(vts[c].shape) != op.shape:
                raise ValueError('cell %d: potential of shape %s'
                                 % (c, tuple(vts[c].shape)))
        pin = (_lib.P * n)(*[t.data_ptr() for t in ins])
        pout = (_lib.P * n)(*[t.data_ptr() for t in outs])
        pvt = ((_lib.P * n)(*[t.data_ptr() for t in vts])
               if self.with_potential else None)
        _lib.check(self._lib.gh_stencil_batch_apply(
            self.handle, pin, pout, pvt, stream_ptr()))


class BatchedHamiltonian:
    """H psi of several independent cells (each an FDHamiltonian with its own
    grid, potential and projectors) with one stencil launch for all."""

    def __init__(self, hams, nbands, tune=True, tiles=None):
        from .autotune import tune_batch
        self.hams = list(hams)
        self.nbands = [int(n) for n in nbands]
        n = len(self.hams)
        self.tuned = None
        if tiles is not None:
            rows = [r for r, d in tiles]
            depths = [d for r, d in tiles]
        elif tune:
            best = tune_batch([H.kin for H in self.hams], self.nbands)
            self.tuned = best
            rows, depths = ((best[0], best[1]) if best
                            else ([0] * n, [0] * n))
        else:
            rows, depths = [0] * n, [0] * n
        self.batch = StencilBatch([H.kin for H in self.hams], self.nbands,
                                  rows, depths, with_potential=True)
        self.graph = None

    def apply(self, psis, outs):
        """outs[c] = H_c psis[c]; returns the projections per cell."""
        self.batch.apply(psis, outs, [H.vt_G for H in self.hams])
        Ps = []
        for H, psi, out in zip(self.hams, psis, outs):
            if H.pt is None:
                Ps.append(None)
                continue
            P = H.pt.integrate_M(psi)
            H.pt.add_M(out, H.dH(P))
            Ps.append(P)
        return Ps

    def capture(self, psis, outs):
        """Record apply(psis, outs) in a HIP graph (these arrays, fixed)."""
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):
            self.apply(psis, outs)          # warm-up outside the graph
        torch.cuda.current_stream().wait_stream(s)
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            self.apply(psis, outs)
        self.graph = g
        return g

    def replay(self):
        self.graph.replay()
