"""ULM files: the container of GPAW's .gpw files.

The format is ASE's ULM (``ase.io.ulm``, format version 3; GPAW calls it
through gpaw/io/__init__.py:1-15 and gpaw/new/gpw.py).  ASE is a separate
package, not part of this image, so the layout below is restated from its
published description:

* bytes 0-7 ``b'- of Ulm'``; 8-23 the tag, ASCII, left-justified and
  space-padded to 16; then three little-endian int64: the format version
  (3), the number of items written, and the file offset of the item-offset
  table (48 for a new file, whose table of one entry, -1, follows);
* arrays: raw little-endian data, each starting at an 8-byte boundary (the
  gap filled with ``b'#'``), described in their item as
  ``"name.": {"ndarray": [shape, dtype-name, offset]}``;
* an item (one ``sync``): an int64 byte count and the item's dictionary as
  JSON; child writers are nested dictionaries under ``"name."``.  When the
  offset table is full a table 42 times longer is written at the end of the
  file and the header's table offset updated.

Reader reads the last item (or `index`); its keys become attributes,
arrays come as NDArrayReader (shape, dtype, slicing read from the file,
``proxy``)."""
import json
from pathlib import Path

import numpy as np

MAGIC = b'- of Ulm'
VERSION = 3
N1 = 42


def _align(fd):
    pos = fd.tell()
    r = pos % 8
    if r:
        fd.write(b'#' * (8 - r))
        pos += 8 - r
    return pos


def _writeint(fd, n, pos=None):
    if pos is not None:
        fd.seek(pos)
    fd.write(np.array(n, '<i8').tobytes())


class _Encoder(json.JSONEncoder):
    def default(self, obj):
        if isinstance(obj, np.integer):
            return int(obj)
        if isinstance(obj, np.floating):
            return float(obj)
        if isinstance(obj, np.bool_):
            return bool(obj)
        if isinstance(obj, (complex, np.complexfloating)):
            return {'__complex__': [obj.real, obj.imag]}
        if isinstance(obj, np.ndarray):
            a = np.ascontiguousarray(obj)
            flat = a.ravel()
            if a.dtype.kind == 'c':
                flat = np.stack([flat.real, flat.imag], 1).ravel()
            return {'__ndarray__': [list(a.shape), a.dtype.name,
                                    flat.tolist()]}
        return json.JSONEncoder.default(self, obj)


def _object_hook(d):
    if '__complex__' in d:
        re, im = d['__complex__']
        return complex(re, im)
    if '__ndarray__' in d:
        shape, dtype, flat = d['__ndarray__']
        dt = np.dtype(dtype)
        if dt.kind == 'c':
            f = np.array(flat, float).reshape(-1, 2)
            return (f[:, 0] + 1j * f[:, 1]).astype(dt).reshape(shape)
        return np.array(flat, dt).reshape(shape)
    return d


class Writer:
    """Write a ULM file (mode 'w'); `child(name)` returns a writer whose
    data nest under name."""

    def __init__(self, filename, tag='', mode='w', _parent=None, _data=None):
        if _parent is not None:
            self._root = _parent._root
            self.data = _data
            return
        assert mode == 'w'
        self._root = self
        self.fd = Path(filename).open('wb')
        self.nitems = 0
        self.pos0 = 48
        self.offsets = np.array([-1], '<i8')
        self.fd.write(MAGIC + ('%-16s' % tag).encode('ascii')[:16])
        self.fd.write(np.array([VERSION, self.nitems, self.pos0],
                               '<i8').tobytes())
        self.fd.write(self.offsets.tobytes())
        self.data = {}
        self._missing = 0

    def __enter__(self):
        return self

    def __exit__(self, *args):
        self.close()

    def write(self, **kwargs):
        for name, value in kwargs.items():
            self.data[name] = value

    def child(self, name):
        dct = self.data[name + '.'] = {}
        return Writer(None, _parent=self, _data=dct)

    def add_array(self, name, shape, dtype=float):
        root = self._root
        assert root._missing == 0, 'previous array not filled'
        if isinstance(shape, (int, np.integer)):
            shape = (shape,)
        shape = tuple(int(s) for s in shape)
        dtype = np.dtype(dtype)
        pos = _align(root.fd)
        self.data[name + '.'] = {'ndarray': [list(shape), dtype.name, pos]}
        root._dtype = dtype
        root._missing = int(np.prod(shape))

    def fill(self, a):
        root = self._root
        a = np.ascontiguousarray(a, dtype=root._dtype)
        root._missing -= a.size
        assert root._missing >= 0, 'more data than the array holds'
        root.fd.write(a.astype(a.dtype.newbyteorder('<'), copy=False)
                      .tobytes())

    def sync(self):
        root = self._root
        assert root._missing == 0, 'last array not filled'
        fd = root.fd
        fd.seek(0, 2)
        pos = fd.tell()
        text = json.dumps(root.data, cls=_Encoder).encode()
        _writeint(fd, len(text))
        fd.write(text)
        n = len(root.offsets)
        if root.nitems >= n:
            offsets = np.zeros(n * N1, '<i8')
            offsets[:n] = root.offsets
            root.pos0 = _align(fd)
            fd.write(offsets.tobytes())
            _writeint(fd, root.pos0, 40)
            root.offsets = offsets
        root.offsets[root.nitems] = pos
        _writeint(fd, pos, root.pos0 + root.nitems * 8)
        root.nitems += 1
        _writeint(fd, root.nitems, 32)
        fd.flush()
        fd.seek(0, 2)
        root.data = {}

    def close(self):
        root = self._root
        if root.fd is None:
            return
        if root.data:
            root.sync()
        root.fd.close()
        root.fd = None


class NDArrayReader:
    """An array of a ULM file, read on indexing: a[i], a[i, j], a[:]."""

    def __init__(self, path, shape, dtype, offset):
        self.path = path
        self.shape = tuple(shape)
        self.dtype = np.dtype(dtype)
        self.offset = int(offset)
        self.scale = 1.0
        self.size = int(np.prod(self.shape))
        self.ndim = len(self.shape)

    def __len__(self):
        return self.shape[0]

    def _read(self, start, count, shape):
        a = np.fromfile(self.path, dtype=self.dtype.newbyteorder('<'),
                        count=count,
                        offset=self.offset + start * self.dtype.itemsize)
        a = a.astype(self.dtype).reshape(shape)
        if self.scale != 1.0:
            a = a * self.scale
        return a

    def read(self):
        return self._read(0, self.size, self.shape)

    def __getitem__(self, index):
        if not isinstance(index, tuple):
            index = (index,)
        lead = []
        for i in index:
            if isinstance(i, (int, np.integer)):
                lead.append(int(i))
            else:
                break
        rest = index[len(lead):]
        stride = int(np.prod(self.shape[len(lead):]))
        start = 0
        for d, i in enumerate(lead):
            n = self.shape[d]
            if not -n <= i < n:
                raise IndexError(i)
            start += (i % n) * int(np.prod(self.shape[d + 1:]))
        a = self._read(start, stride, self.shape[len(lead):])
        return a[rest] if rest else a

    def proxy(self, *indices):
        """The sub-array at the leading indices, still on disk."""
        start = 0
        for d, i in enumerate(indices):
            start += int(i) * int(np.prod(self.shape[d + 1:]))
        p = NDArrayReader(self.path, self.shape[len(indices):], self.dtype,
                          self.offset + start * self.dtype.itemsize)
        p.scale = self.scale
        return p


class Reader:
    """Read item `index` (default: the last) of a ULM file."""

    def __init__(self, filename, index=None, _data=None, _header=None):
        self._path = str(filename)
        if _data is None:
            with open(self._path, 'rb') as fd:
                head = fd.read(48)
                if len(head) < 48 or head[:8] != MAGIC:
                    raise ValueError('%s is not a ULM file' % filename)
                self.tag = head[8:24].decode('ascii').rstrip()
                version, nitems, pos0 = np.frombuffer(head[24:48], '<i8')
                if version > VERSION:
                    raise ValueError('ULM format version %d is newer than %d'
                                     % (version, VERSION))
                self.format_version = int(version)
                self.nitems = int(nitems)
                fd.seek(int(pos0))
                offsets = np.frombuffer(fd.read(8 * self.nitems), '<i8')
                if index is None:
                    index = self.nitems - 1
                fd.seek(int(offsets[index]))
                n = int(np.frombuffer(fd.read(8), '<i8')[0])
                _data = json.loads(fd.read(n).decode(),
                                   object_hook=_object_hook)
        self._data = {}
        for key, value in _data.items():
            if key.endswith('.'):
                name = key[:-1]
                if isinstance(value, dict) and 'ndarray' in value:
                    shape, dtype, offset = value['ndarray']
                    self._data[name] = NDArrayReader(self._path, shape,
                                                     dtype, offset)
                else:
                    self._data[name] = Reader(self._path, _data=value)
            else:
                self._data[key] = value

    def __getattr__(self, name):
        try:
            return self.__dict__['_data'][name]
        except KeyError:
            raise AttributeError(name)

    def __contains__(self, name):
        return name in self._data

    def get(self, name, default=None):
        return self._data.get(name, default)

    def keys(self):
        return self._data.keys()

    def asdict(self):
        return {k: (v.asdict() if isinstance(v, Reader) else v)
                for k, v in self._data.items()}
