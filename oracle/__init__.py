"""CPU oracle for the finite-difference H-psi path -- TEST INFRASTRUCTURE ONLY.

This package is a plain NumPy restatement of the reference algorithms
(GPAW's ``c/bmgs``, ``c/bc.c``, ``c/lfc.c``, ``c/fd_preconditioner.c`` and the
Python layers above them).  It is the *checker*: only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import it.  The product package ``gpaw_amd`` never imports it and has no CPU
fallback.

Modules
-------
grid        serial GridDescriptor geometry (grid_descriptor.py:77-160, 344-411)
stencil     Laplace coefficients and the padded-halo stencil apply
            (fd_operators.py:171-232, bc.c:146-258, bmgs/fd.c:9-33)
transformers restrict / interpolate (transformers.py:18-91,
            bmgs/restrict.c, bmgs/interpolate.c)
spline      radial cubic spline (bmgs/spline.c:11-58)
lfc         sphere boxes, spline_to_grid, LFC integrate / add
            (lfc.py:63-119, 273-374, 392-631; lfc.c:432-505, 818-977)
precond     the 14-step multigrid V-cycle (preconditioner.py:35-93)
rmmdiis     Hamiltonian with dH/dS, subspace diagonalisation, inverse-Cholesky
            orthonormalisation and the blocked RMM-DIIS sweep
            (eigensolver.py:153-236, fdpw.py:394-424, rmmdiis.py:63-254)
davidson    one Davidson iterate (eigensolvers/davidson.py:80-222)
poisson     Jacobi and Gauss-Seidel relaxation, the multigrid Poisson solver
            (c/bmgs/relax.c, c/operators.c:55-95, poisson.py:318-546)
density     pseudo-density accumulation (c/utilities.c:313-356)
refbuild    ctypes loader for ``oracle/_ref/libgpawref.so``, the reference C
            kernels compiled from their own sources by ``oracle/Makefile``

Pinning (what each function is checked against) is recorded in each module's
docstring; see ``tests/test_oracle_*.py``.
"""
