"""ULM container and .gpw wave-function I/O (CPU).  The ULM layout is
restated from ASE's published format (ase.io.ulm, version 3; ASE is not
in this image and the reference holds no .gpw fixture): the tests check
the layout as restated (header, offset table, 8-byte aligned arrays,
JSON items, growth of the offset table) and the round trips; agreement
with files written by ASE itself is not pinned."""
import json

import numpy as np
import pytest

torch = pytest.importorskip('torch')

from gpaw_amd.io import Reader, Writer, read_wave_functions, \
    write_wave_functions  # noqa: E402
from gpaw_amd.io.gpw import BOHR  # noqa: E402


def test_ulm_layout_and_round_trip(tmp_path):
    f = tmp_path / 'a.ulm'
    rng = np.random.default_rng(1)
    x = rng.standard_normal((3, 5))
    z = (rng.standard_normal(7) + 1j * rng.standard_normal(7)).astype(
        np.complex64)
    with Writer(f, tag='gpaw') as w:
        w.write(version=4, name='test', flag=True, value=1.5,
                small=np.arange(3))
        c = w.child('sub')
        c.write(k=2)
        c.add_array('x', x.shape)
        c.fill(x[:2])
        c.fill(x[2:])
        w.add_array('z', z.shape, z.dtype)
        w.fill(z)
    raw = f.read_bytes()
    assert raw[:8] == b'- of Ulm'
    assert raw[8:24] == b'gpaw            '
    version, nitems, pos0 = np.frombuffer(raw[24:48], '<i8')
    assert (version, nitems, pos0) == (3, 1, 48)
    off = np.frombuffer(raw[48:56], '<i8')[0]
    n = np.frombuffer(raw[off:off + 8], '<i8')[0]
    item = json.loads(raw[off + 8:off + 8 + n])
    assert set(item) == {'version', 'name', 'flag', 'value', 'small',
                         'sub.', 'z.'}
    shape, dtype, pos = item['sub.']['x.']['ndarray']
    assert shape == [3, 5] and dtype == 'float64' and pos % 8 == 0
    assert np.array_equal(np.frombuffer(raw[pos:pos + 120], '<f8'),
                          x.ravel())
    r = Reader(f)
    assert r.tag == 'gpaw' and r.version == 4 and r.name == 'test'
    assert r.flag is True and r.value == 1.5
    assert np.array_equal(r.small, np.arange(3))
    assert r.sub.k == 2
    assert np.array_equal(r.sub.x.read(), x)
    assert np.array_equal(r.sub.x[1], x[1])
    assert np.array_equal(r.sub.x[2, 1:4], x[2, 1:4])
    assert np.array_equal(r.z.read(), z) and r.z.dtype == np.complex64
    assert r.asdict()['sub']['k'] == 2


def test_ulm_many_items_grow_the_offset_table(tmp_path):
    f = tmp_path / 'b.ulm'
    w = Writer(f)
    for i in range(50):                     # beyond the first table
        w.write(i=i)
        w.sync()
    w.close()
    raw = f.read_bytes()
    version, nitems, pos0 = np.frombuffer(raw[24:48], '<i8')
    assert nitems == 50 and pos0 != 48 and pos0 % 8 == 0
    for i in (0, 1, 41, 42, 49):
        assert Reader(f, index=i).i == i
    assert Reader(f).i == 49


@pytest.mark.parametrize('dtype', [np.float64, np.complex128, np.float32])
def test_gpw_wave_functions_round_trip(tmp_path, dtype):
    rng = np.random.default_rng(2)
    shape = (5, 6, 7, 8)
    psit = [[rng.standard_normal(shape).astype(dtype) for k in range(2)]
            for s in range(2)]
    if np.dtype(dtype).kind == 'c':
        psit = [[p + 1j * rng.standard_normal(shape) for p in ps]
                for ps in psit]
    P = [[rng.standard_normal((5, 16)).astype(dtype) for k in range(2)]
         for s in range(2)]
    f = tmp_path / 'wf.gpw'
    write_wave_functions(f, [[torch.from_numpy(p) for p in ps]
                             for ps in psit], P, block=2)
    r = Reader(f)
    assert r.bohr == BOHR and r.version == 4
    assert r.wave_functions.coefficients.shape == (2, 2) + shape
    raw = r.wave_functions.coefficients[1, 0]
    assert np.allclose(raw, psit[1][0] * BOHR**-1.5, rtol=1e-6, atol=0)
    assert np.array_equal(r.wave_functions.projections[0, 1], P[0][1])
    for s in range(2):
        for k in range(2):
            t = read_wave_functions(f, s, k, device='cpu', block=3)
            tol = 1e-6 if dtype == np.float32 else 1e-14
            assert np.allclose(t.numpy(), psit[s][k], rtol=tol, atol=0)
