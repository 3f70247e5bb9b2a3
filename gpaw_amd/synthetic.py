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
