"""Train the learned preconditioner.

    python -m gpaw_amd.learned.train --cells 2000 --out weights.npz

Generates residual -> correction pairs on seeded cells (datagen), fits the
U-Net (torch, FP32) to them with a scale-free loss |f(x) - y|^2 / |y|^2
per sample, and writes the weights (npz) plus a JSON report: the loss
curve and, on held-out cells, the fraction of the residual norm^2 left
after one line-search step along the network's correction and along the
multigrid V-cycle's (lower is better)."""
import argparse
import json
import time

import numpy as np
import torch

from . import datagen
from .unet import UNet3, save_weights


def rho_of(es, ham, wfs, R, d):
    """1 - <R|A d>^2 / (|R|^2 |A d|^2) per band, A = H - eps S."""
    from .. import blas
    Ad = torch.empty_like(d)
    es.residual_of(ham, wfs, d, wfs.eps, Ad)
    G = blas.multi_gram([Ad, R])
    return 1.0 - G[:, 0, 1] ** 2 / (G[:, 0, 0] * G[:, 1, 1])


def validate(model, ncells, seed0, N=24, nbands=8):
    """Median rho of the U-Net (HIP path) and of the V-cycle on held-out
    cells."""
    from .. import synthetic as syn
    from ..eigensolvers import RMMDIIS, FDWaveFunctions
    from ..grid_descriptor import GridDescriptor
    from .preconditioner import LearnedPreconditioner
    rn, rm = [], []
    for c in range(ncells):
        w = datagen.cell_geometry(seed0 + c, N)
        gd = GridDescriptor((N,) * 3, (w.L,) * 3)
        vt = datagen.cell_potential(gd, w)
        sysm = syn.System(w, gd, vt, np.float64)
        psi0 = syn.lcao_start(gd, w, sysm.H, nbands)
        wfs = FDWaveFunctions(gd, psi0, sysm.pt, sysm.pc)
        es = RMMDIIS()
        wfs.orthonormalize(sysm.H)
        for s in range(w.sweeps):
            es.iterate(sysm.H, wfs, None)
        R = torch.empty_like(wfs.psit)
        es.residual_of(sysm.H, wfs, wfs.psit, wfs.eps, R)
        lp = LearnedPreconditioner(gd, sysm.kin, vt, model)
        rn.append(rho_of(es, sysm.H, wfs, R, lp(R)))
        rm.append(rho_of(es, sysm.H, wfs, R, sysm.pc(R)))
    rn, rm = torch.cat(rn), torch.cat(rm)
    return {'rho_median_unet': float(rn.median()),
            'rho_median_vcycle': float(rm.median()),
            'rho_mean_unet': float(rn.mean()),
            'rho_mean_vcycle': float(rm.mean())}


def main(argv=None):
    p = argparse.ArgumentParser()
    p.add_argument('--cells', type=int, default=2000)
    p.add_argument('--val-cells', type=int, default=50)
    p.add_argument('--grid', type=int, default=24)
    p.add_argument('--steps', type=int, default=3000)
    p.add_argument('--batch', type=int, default=32)
    p.add_argument('--lr', type=float, default=2e-3)
    p.add_argument('--rho-weight', type=float, default=1.0,
                   help='weight of the line-search residual fraction in the '
                        'loss (0: the CG target alone)')
    p.add_argument('--out', required=True)
    p.add_argument('--report')
    a = p.parse_args(argv)
    torch.manual_seed(0)
    t0 = time.time()
    data = datagen.generate(a.cells, N=a.grid)
    X, Y = data['x'], data['y']
    t_data = time.time() - t0
    print('data: %d samples from %d cells in %.1f s' % (len(X), a.cells,
                                                       t_data), flush=True)
    torch.backends.cudnn.benchmark = True
    model = UNet3().cuda()
    opt = torch.optim.Adam(model.parameters(), lr=a.lr)
    sched = torch.optim.lr_scheduler.CosineAnnealingLR(opt, a.steps)
    g = torch.Generator(device='cuda')
    g.manual_seed(0)
    losses = []
    for step in range(a.steps):
        idx = torch.randint(0, len(X), (a.batch,), device='cuda', generator=g)
        x, y = X[idx], Y[idx]
        f = model(x)
        num = ((f - y) ** 2).flatten(1).sum(1)
        den = (y ** 2).flatten(1).sum(1)
        loss = (num / den).mean()
        if a.rho_weight > 0:
            # the eigensolver's own judge: the fraction of |R|^2 left after
            # the line search along f, with A = H - eps of the sample's cell
            Af = datagen.apply_shifted(data, idx, f)
            r = x[:, :1]
            ra = (r * Af).flatten(1).sum(1)
            rho = 1 - ra ** 2 / ((r * r).flatten(1).sum(1) *
                                 (Af * Af).flatten(1).sum(1))
            loss = loss + a.rho_weight * rho.mean()
        opt.zero_grad()
        loss.backward()
        opt.step()
        sched.step()
        if step % 100 == 0 or step == a.steps - 1:
            losses.append((step, float(loss.detach())))
            print('step %d loss %.4f (%.1f s)' % (step, float(loss.detach()),
                                                  time.time() - t0),
                  flush=True)
    model.eval()
    save_weights(model.cpu(), a.out)
    model.cuda()
    rep = {'cells': a.cells, 'grid': a.grid, 'samples': len(X),
           'steps': a.steps, 'batch': a.batch, 'lr': a.lr,
           'data_seconds': round(t_data, 1),
           'train_seconds': round(time.time() - t0 - t_data, 1),
           'loss_curve': losses}
    with torch.no_grad():
        rep['validation'] = validate(model, a.val_cells, seed0=10**6,
                                     N=a.grid)
    print(json.dumps(rep['validation']), flush=True)
    if a.report:
        with open(a.report, 'w') as fh:
            json.dump(rep, fh, indent=1)


if __name__ == '__main__':
    main()
