"""Learned preconditioner: a small 3D U-Net that maps residuals R (and the
local potential v) to corrections, trained on pairs the FD Hamiltonian
generates, run on the matrix cores, with the multigrid V-cycle as fallback
and the residual norm as judge."""
from .preconditioner import LearnedPreconditioner  # noqa: F401
