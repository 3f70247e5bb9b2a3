"""File I/O of the hot path's data: ULM files (the container format of
GPAW's .gpw, gpaw/io/__init__.py:1-15) and the wave-function part of a
.gpw (gpaw/new/ibzwfs.py:370-408 writing, gpaw/new/fd/builder.py:89-124
reading)."""
from .ulm import Reader, Writer  # noqa: F401
from .gpw import read_wave_functions, write_wave_functions  # noqa: F401
