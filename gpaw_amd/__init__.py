"""MI355X-native finite-difference H-psi path (GPAW FD mode drop-ins).

Public classes keep the reference's module layout:

    gpaw_amd.grid_descriptor.GridDescriptor
    gpaw_amd.fd_operators.FDOperator / Laplace / GUCLaplace
    gpaw_amd.transformers.Transformer
    gpaw_amd.preconditioner.Preconditioner
    gpaw_amd.spline.Spline
    gpaw_amd.lfc.LFC / LocalizedFunctionsCollection
    gpaw_amd.hamiltonian.FDHamiltonian

All compute runs in libgpawhip.so (HIP for gfx950) through its C ABI.
"""
__version__ = '0.1.0'
