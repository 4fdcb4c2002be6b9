"""CPU oracle for the pressure-projection Poisson solve -- TEST INFRASTRUCTURE ONLY.

This package restates, in numpy, the algorithm of the reference's multigrid
pressure solver (``mgMK/``) and of the projection step around it
(``dns/NSsolver.F90`` ``Psource``/``Pgrad``, ``dns/Routines.F90`` ``divergence``).
Every function cites the reference file:line it follows.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline``
leg may import it, and only as the checker (or as the timed CPU baseline).
The product path (``anelastic_solve_amd``) never imports it and has no CPU
fallback.

Pinning.  The reference is Fortran 90 that only builds with an MPI library
(its serial build is broken: ``mgMK/mg_solver.F90:42`` uses module ``dd`` that
exists only under ``-DPARALLEL``), so it is not built here.  Each oracle
function is pinned instead by mathematics, in ``tests/test_oracle_*.py``:

* ``grid``      -- metric identities (cell widths sum to L, centre spacing).
* ``ops``       -- Fourier eigenvalues of the periodic stencil; ``D.G == A``
                  (``psource(pgrad(p)) == A p``, which ties three independent
                  restatements together); closed-form quadratic/linear
                  solutions of ``mgMK/test/test.F90``.
* ``direct``    -- two independent exact solvers (FFT(x,y)+Thomas(z) and a
                  scipy sparse LU) agree and reproduce the closed forms.
* ``vcycle``    -- red-black sweep: fixed point of the exact solution, zero
                  residual on the just-updated colour, equality with a literal
                  lexicographic loop; z-line sweep: zero residual on the
                  updated columns; restriction: sum/constant preservation and
                  a delta-function index map; prolongation: constants and
                  linears; the V-cycle converges to the direct solution.
"""
