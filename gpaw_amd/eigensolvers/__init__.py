"""Eigensolver drivers over device tensors (FD mode)."""
from .rmmdiis import RMMDIIS, FDWaveFunctions  # noqa: F401
from .davidson import Davidson  # noqa: F401
