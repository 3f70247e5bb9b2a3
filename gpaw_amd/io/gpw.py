"""The wave functions of a .gpw file (FD mode): write
(gpaw/new/ibzwfs.py:370-408, under the ``wave_functions`` child of
gpaw/new/gpw.py:29-71) and read (gpaw/new/fd/builder.py:89-124).

Layout, as the reference's: ``wave_functions.coefficients`` of shape
(nspins, nibzkpts, nbands, n0, n1, n2) in the wave functions' dtype, stored
in units of Bohr^-1.5 (the values times Bohr^-1.5); an optional
``wave_functions.projections`` (nspins, nibzkpts, nbands, nproj).  Reading
multiplies by ``bohr**1.5`` (1 for files of version < 0) and accepts the
older name ``values``.  Device arrays stream through host memory one band
block at a time."""
import numpy as np
import torch

from .ulm import Reader, Writer

# ASE's atomic units (CODATA 2014), as written into .gpw files
BOHR = 0.5291772105638411
HARTREE = 27.211386024367243


def _host_blocks(t, block):
    """Host copies of t[b0:b0 + block] in turn."""
    for b0 in range(0, t.shape[0], block):
        yield t[b0:b0 + block].detach().to('cpu').numpy()


def write_wave_functions(target, psit_sk, P_sk=None, block=16, bohr=BOHR):
    """Write wave functions psit_sk[s][k] (arrays (nbands, n0, n1, n2), on
    the GPU or the host) and optionally projections P_sk[s][k]
    (nbands, nproj).  target: a file name (a new .gpw-style ULM file with
    version, ha and bohr at the top) or a ulm.Writer whose
    ``wave_functions`` child receives them."""
    own = not isinstance(target, Writer)
    writer = Writer(target, tag='gpaw') if own else target
    if own:
        writer.write(version=4, ha=HARTREE, bohr=bohr)
    nspins = len(psit_sk)
    nk = len(psit_sk[0])
    first = psit_sk[0][0]
    nbands = int(first.shape[0])
    grid = tuple(int(n) for n in first.shape[1:])
    dtype = np.dtype(str(first.dtype).replace('torch.', '')) \
        if isinstance(first, torch.Tensor) else first.dtype
    wf = writer.child('wave_functions')
    if P_sk is not None:
        P0 = P_sk[0][0]
        nproj = int(P0.shape[-1])
        wf.add_array('projections', (nspins, nk, nbands, nproj), dtype)
        for s in range(nspins):
            for k in range(nk):
                P = P_sk[s][k]
                P = P.detach().cpu().numpy() if isinstance(
                    P, torch.Tensor) else np.asarray(P)
                wf.fill(P.reshape(nbands, nproj))
    c = bohr ** -1.5
    wf.add_array('coefficients', (nspins, nk, nbands) + grid, dtype)
    for s in range(nspins):
        for k in range(nk):
            psit = psit_sk[s][k]
            if not isinstance(psit, torch.Tensor):
                psit = torch.from_numpy(np.ascontiguousarray(psit))
            assert tuple(psit.shape) == (nbands,) + grid
            for a in _host_blocks(psit, block):
                wf.fill(a * c)
    if own:
        writer.close()
    else:
        writer.sync()


def read_wave_functions(source, spin=0, k=0, device='cuda', block=16):
    """The wave functions of (spin, k) from a .gpw / ULM file name or a
    ulm.Reader, as a tensor (nbands, n0, n1, n2) on `device`, in the units
    of the grid (times bohr**1.5)."""
    reader = source if isinstance(source, Reader) else Reader(source)
    wfs = reader.wave_functions
    if 'coefficients' in wfs:
        name = 'coefficients'
    elif 'values' in wfs:
        name = 'values'
    else:
        raise ValueError('the file holds no wave functions')
    c = reader.bohr ** 1.5
    if reader.get('version', 0) < 0:
        c = 1.0                             # old gpw file
    data = getattr(wfs, name).proxy(spin, k)
    nbands = data.shape[0]
    out = None
    for b0 in range(0, nbands, block):
        n = min(block, nbands - b0)
        a = data._read(b0 * int(np.prod(data.shape[1:])),
                       n * int(np.prod(data.shape[1:])),
                       (n,) + data.shape[1:]) * c
        t = torch.from_numpy(a)
        if out is None:
            out = torch.empty(data.shape, dtype=t.dtype, device=device)
        out[b0:b0 + n].copy_(t)
    return out
