"""CPU oracle for the shear-periodic RK3 time advance (TEST INFRASTRUCTURE ONLY).

This package is a numpy restatement of the reference Fortran's algorithm
(hiiragiyua/shear_nnf: cross.f90, diffy_shear_hepta.f90, fou3D_thread.f,
operators.f90, initcr.f90).  It exists only to check the HIP product path:
only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import it.  The product library never calls into it.

The reference itself is not buildable in this image (its FFTW include is an
HTML error page and it needs an HDF5 Fortran module built by another
compiler), so each function here is pinned instead by closed forms and
invariants in ``tests/test_oracle_*.py``:

* ``compact``  -- coefficient design points and Taylor order, the
  twisted-Fourier eigenvalue identity, dense brute-force solves;
* ``fftxz``    -- textbook DFT sums and the 3/2-rule index map;
* ``cross``    -- divergence/curl identities of the KMM reconstruction,
  energy neutrality of the dealiased u x omega term, the RDT Kelvin-mode
  closed form (operators.f90:1516-1562) for the whole RK3 step, and
  Parseval for the statistics.
"""
