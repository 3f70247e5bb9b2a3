"""Sphere boxes (GridDescriptor.get_boxes, grid_descriptor.py:344-411) --
CPU only.  The oracle's boxes are pinned by brute force: every (grid point,
periodic image) pair within rcut of the sphere centre lies in exactly one
box with that image, and no box holds a point of an image it does not name
beyond the box's corners; the product's boxes equal the oracle's."""
import itertools

import numpy as np
import pytest

from gpaw_amd.grid_descriptor import GridDescriptor
from oracle.grid import Grid

CASES = [
    ((12, 14, 16), np.diag([5.0, 6.0, 7.0]), (True, True, True), 2.6),
    ((12, 12, 12), np.diag([4.0, 4.0, 4.0]), (True, False, True), 2.2),
    ((10, 12, 14), [(4, 0, 0), (1.5, 4.5, 0), (0.5, 0.5, 5)],
     (True, True, True), 2.0),
    ((9, 10, 11), np.diag([3.0, 3.5, 4.0]), (False, True, False), 1.9),
    ((8, 8, 8), np.diag([2.0, 2.0, 2.0]), (True, True, True), 2.5),
]


def brute(gd, spos_c, rcut):
    """{(g, image)} for every domain point within rcut of an image."""
    r_vg = gd.coordinates()
    out = set()
    for disp in itertools.product(range(-3, 4), repeat=3):
        if any(d and not p for d, p in zip(disp, gd.pbc_c)):
            continue
        c = (np.asarray(spos_c) + disp) @ gd.cell_cv
        d2 = ((r_vg - c[:, None, None, None]) ** 2).sum(0)
        for g in zip(*np.nonzero(d2 < rcut ** 2)):
            out.add((tuple(int(i) + int(b) for i, b in zip(g, gd.beg_c)),
                     tuple(-x for x in disp)))
    return out


@pytest.mark.parametrize('N_c,cell,pbc,rcut', CASES)
def test_boxes_cover_sphere(N_c, cell, pbc, rcut):
    rng = np.random.default_rng(7)
    gd = Grid(N_c, cell, pbc)
    for _ in range(6):
        spos = rng.uniform(0, 1, 3)
        boxes = gd.get_boxes(spos, rcut)
        covered = {}
        for beg, end, disp in boxes:
            assert (beg >= gd.beg_c).all() and (end <= gd.end_c).all()
            for g in itertools.product(*[range(b, e)
                                         for b, e in zip(beg, end)]):
                key = (g, tuple(int(x) for x in disp))
                assert key not in covered, 'point covered twice'
                covered[key] = True
        for key in brute(gd, spos, rcut):
            assert key in covered


@pytest.mark.parametrize('N_c,cell,pbc,rcut', CASES)
def test_product_boxes_equal_oracle(N_c, cell, pbc, rcut):
    rng = np.random.default_rng(11)
    og = Grid(N_c, cell, pbc)
    pg = GridDescriptor(N_c, cell, pbc)
    for _ in range(20):
        spos = rng.uniform(-0.2, 1.2, 3)
        if not all(pbc):
            spos = np.where(pbc, spos, rng.uniform(0.3, 0.7, 3))
        a = og.get_boxes(spos, rcut)
        b = pg.get_boxes(spos, rcut)
        assert len(a) == len(b)
        for (b0, e0, d0), (b1, e1, d1) in zip(a, b):
            np.testing.assert_array_equal(b0, b1)
            np.testing.assert_array_equal(e0, e1)
            np.testing.assert_array_equal(d0, d1)


def test_uncut_box_raises_at_open_boundary():
    from gpaw_amd.grid_descriptor import GridBoundsError
    gd = GridDescriptor((10, 10, 10), np.diag([4.0] * 3), (True, False, True))
    with pytest.raises(GridBoundsError):
        gd.get_boxes(np.array([0.5, 0.02, 0.5]), 1.5, cut=False)
    assert gd.get_boxes(np.array([0.5, 0.5, 0.5]), 1.5, cut=False)
