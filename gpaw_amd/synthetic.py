"""Seeded synthetic FD workloads (no files, no network).

Units: Bohr and Hartree.  Grid spacing h = 0.2 Angstrom.  Seeds: psi 1,
atoms 2, potential 3, dH 4, projector widths 5.

full_hpsi(N, nbands, natoms): periodic N^3 grid; atoms uniform in the cell
with a minimum-image separation of at least 1.5 A; per atom four radial
functions exp(-r^2/rc^2) with l = 0, 0, 1, 1 (rc ~ U[0.6, 1.0] A, cutoff
2.0 A, 201 spline points), i.e. 8 projectors r^l exp(-r^2/rc^2) Y_lm;
dH_a random symmetric 8x8 with entries N(0, 0.1); v(r) a sum of one
Gaussian per atom, depth U[-2, 0] Ha, exp(-r^2 / w^2) with w = 0.8 A
(minimum image).
"""
import numpy as np

BOHR = 0.52917721067  # Angstrom
H_GRID = 0.2 / BOHR


class Workload:
    pass


def atoms(natoms, L, seed=2, dmin=1.5 / BOHR, max_tries=100000):
    rng = np.random.default_rng(seed)
    pos = []
    tries = 0
    while len(pos) < natoms:
        tries += 1
        if tries > max_tries:
            raise RuntimeError('cannot place %d atoms' % natoms)
        p = rng.uniform(0, L, 3)
        ok = True
        for q in pos:
            d = p - q
            d -= L * np.round(d / L)
            if np.dot(d, d) < dmin**2:
                ok = False
                break
        if ok:
            pos.append(p)
    return np.array(pos)


def potential(gd, pos_av, seed=3, width=0.8 / BOHR):
    rng = np.random.default_rng(seed)
    depth = rng.uniform(-2.0, 0.0, len(pos_av))
    r_vG = gd.get_grid_point_coordinates()
    L_c = np.diag(gd.cell_cv)
    v = np.zeros(tuple(gd.n_c))
    for d, p in zip(depth, pos_av):
        x = [(r_vG[c] - p[c]) - L_c[c] * np.round((r_vG[c] - p[c]) / L_c[c])
             for c in range(3)]
        v += d * np.exp(-(x[0]**2 + x[1]**2 + x[2]**2) / width**2)
    return v


def projector_data(natoms, seed=5, rcut=2.0 / BOHR, npts=201):
    """Per atom: list of (l, rcut, f_g) for l = 0, 0, 1, 1."""
    rng = np.random.default_rng(seed)
    r = np.linspace(0, rcut, npts)
    out = []
    for a in range(natoms):
        rc = rng.uniform(0.6, 1.0, 4) / BOHR
        out.append([(l, rcut, np.exp(-r**2 / rc[j]**2))
                    for j, l in enumerate((0, 0, 1, 1))])
    return out


def dH_matrices(natoms, ni=8, seed=4, sigma=0.1):
    rng = np.random.default_rng(seed)
    out = []
    for a in range(natoms):
        m = rng.normal(0, sigma, (ni, ni))
        out.append(np.triu(m) + np.triu(m, 1).T)
    return out


def full_hpsi_geometry(N, natoms):
    """Host description of the full H-psi workload on an N^3 grid."""
    w = Workload()
    w.N = N
    w.L = N * H_GRID
    w.natoms = natoms
    w.pos_av = atoms(natoms, w.L)
    w.spos_ac = w.pos_av / w.L
    w.proj = projector_data(natoms)
    w.dH = dH_matrices(natoms)
    return w


def initial_guess(gd, pos_av, nbands, sigma=1.2, seed=1):
    """Atom-centred start for RMM-DIIS (an LCAO-like guess): per atom an s
    and three p Gaussians, exp(-r^2/sigma^2) and (x, y, z) exp(-r^2/sigma^2)
    (minimum image), atom after atom; bands beyond 4 * natoms are the
    lowest plane waves cos(G.r + phi) with seeded random phi.  Returns a float64 tensor
    (nbands, n0, n1, n2) on the GPU (not orthonormalized)."""
    import torch
    n_c = tuple(int(n) for n in gd.n_c)
    L_c = torch.tensor(np.diag(gd.cell_cv), dtype=torch.float64,
                       device='cuda')
    axes = [torch.arange(n, dtype=torch.float64, device='cuda') * h
            for n, h in zip(n_c, np.diag(gd.h_cv))]
    psi = torch.empty((nbands,) + n_c, dtype=torch.float64, device='cuda')
    b = 0
    for p in pos_av:
        if b >= nbands:
            break
        d = []
        for c in range(3):
            x = axes[c] - float(p[c])
            d.append(x - L_c[c] * torch.round(x / L_c[c]))
        x = d[0][:, None, None]
        y = d[1][None, :, None]
        z = d[2][None, None, :]
        g = torch.exp(-(x * x + y * y + z * z) / sigma**2)
        for f in (g, x * g, y * g, z * g):
            if b < nbands:
                psi[b] = f
                b += 1
    if b < nbands:
        rng = np.random.default_rng(seed)
        kmax = 1
        while (2 * kmax + 1)**3 < 2 * (nbands - b) + 1:
            kmax += 1
        ks = [(i, j, k) for i in range(-kmax, kmax + 1)
              for j in range(-kmax, kmax + 1) for k in range(-kmax, kmax + 1)
              if (i, j, k) > (0, 0, 0)]
        ks.sort(key=lambda t: (t[0]**2 + t[1]**2 + t[2]**2, t))
        for i, j, k in ks:
            if b >= nbands:
                break
            phase = (2 * np.pi) * (
                i * axes[0][:, None, None] / L_c[0] +
                j * axes[1][None, :, None] / L_c[1] +
                k * axes[2][None, None, :] / L_c[2])
            shift = float(rng.uniform(0, 2 * np.pi))
            psi[b] = torch.cos(phase + shift)
            b += 1
    return psi


def basis_splines(w, sigma=1.2, rcut_gauss=5.0, npts=201):
    """Atom-centred basis radial functions per atom (an LCAO-like basis):
    the atom's four projector shapes (l = 0, 0, 1, 1) and an s and a p
    Gaussian exp(-r^2 / sigma^2) cut at rcut_gauss Bohr -- 12 functions per
    atom.  Returns [[(l, rcut, f_g)]] in the format of Workload.proj."""
    r = np.linspace(0, rcut_gauss, npts)
    g = np.exp(-r**2 / sigma**2)
    g[-1] = 0.0
    return [list(p) + [(0, rcut_gauss, g), (1, rcut_gauss, g)]
            for p in w.proj]


def lcao_start(gd, w, H, nbands, dtype=float, k_c=None, basis=None):
    """LCAO-like start for the iterative eigensolvers, as an SCF starts from
    the LCAO eigenvectors: the lowest `nbands` Rayleigh-Ritz vectors of the
    Hamiltonian H (an FDHamiltonian) in the span of atom-centred basis
    functions (basis_splines by default), evaluated on the grid with
    periodic images and Bloch phases through an LFC.  The overlap is
    orthogonalised canonically (eigenvectors with eigenvalues below 1e-9 of
    the largest are dropped), so near-dependent basis functions are
    harmless.  Returns device wave functions (nbands, n0, n1, n2),
    orthonormal (dv sum psi_n^* psi_m = delta_nm)."""
    import torch
    from . import blas
    from .lfc import LFC
    from .spline import Spline
    from ._arrays import torch_dtype
    basis = basis_splines(w) if basis is None else basis
    kd = None if k_c is None else np.array([k_c], float)
    lfc = LFC(gd, [[Spline.from_data(l, rc, f) for l, rc, f in p]
                   for p in basis], kd, dtype=dtype)
    lfc.set_positions(w.spos_ac)
    nbasis = lfc.Mmax
    if nbasis < nbands:
        raise ValueError('%d basis functions for %d bands' % (nbasis, nbands))
    td = torch_dtype(dtype)
    phi = gd.zeros(nbasis, dtype)
    lfc.add_M(phi, torch.eye(nbasis, dtype=td, device='cuda'),
              q=-1 if k_c is None else 0)
    hphi = torch.empty_like(phi)
    ph = None if k_c is None else gd.phase_cd(k_c)
    H.apply(phi, hphi, ph)
    f = phi.reshape(nbasis, -1)
    S = blas.gram(f, alpha=gd.dv, hermitian=True, conj=True)
    if H.pt is not None and H.dS_flat is not None:
        P = H.pt.integrate_M(phi)
        S = S + P.conj() @ H.dS(P).T
    Hm = blas.gram(f, hphi.reshape(nbasis, -1), alpha=gd.dv, hermitian=True,
                   conj=True)
    del hphi
    s, U = torch.linalg.eigh(0.5 * (S + S.conj().T))
    keep = s > 1e-9 * s.max()
    X = U[:, keep] / torch.sqrt(s[keep])          # X^H S X = 1
    e, V = torch.linalg.eigh(X.conj().T @ Hm @ X)
    C = X @ V[:, :nbands]                          # basis -> band
    return (C.T @ f).reshape((nbands,) + tuple(gd.n_c)).contiguous()


class System:
    """The FD Hamiltonian of a workload in one precision: LFC of the
    projectors, 19-point kinetic stencil, FDHamiltonian (potential and dH in
    the data's real precision / FP64) and the multigrid preconditioner."""

    def __init__(self, w, gd, vt, dtype=float, k_c=None):
        from .fd_operators import Laplace
        from .hamiltonian import FDHamiltonian
        from .lfc import LFC
        from .preconditioner import Preconditioner
        from .spline import Spline
        self.dtype = dtype
        self.k_c = k_c
        kd = None if k_c is None else np.array([k_c], float)
        self.pt = LFC(gd, [[Spline.from_data(l, rc, f) for l, rc, f in p]
                           for p in w.proj], kd, dtype=dtype)
        self.pt.set_positions(w.spos_ac)
        self.kin = Laplace(gd, -0.5, 3, dtype)
        self.H = FDHamiltonian(gd, self.kin, vt, self.pt, w.dH, dtype=dtype)
        self.pc = Preconditioner(gd, self.kin, dtype)
        self.phase_cd = None if k_c is None else gd.phase_cd(k_c)


# Six independent cells of uneven size: grid points per axis, bands, atoms
# (h = 0.2 A each; potentials and projectors as in full_hpsi_geometry).
BATCH_CELLS = ((40, 16, 8), (56, 48, 16), (72, 96, 32), (96, 192, 48),
               (120, 384, 80), (128, 512, 96))


def batch_cells(cells=BATCH_CELLS, dtype=np.float32):
    """Geometry, grid descriptor and single-precision System per cell."""
    from .grid_descriptor import GridDescriptor
    out = []
    for N, nb, natoms in cells:
        w = full_hpsi_geometry(N, natoms)
        gd = GridDescriptor((N,) * 3, (w.L,) * 3)
        vt = potential(gd, w.pos_av)
        out.append((w, gd, vt, nb, System(w, gd, vt, dtype)))
    return out
