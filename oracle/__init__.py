"""CPU oracle for the shear-periodic RK3 time advance (TEST INFRASTRUCTURE ONLY).

This package is a numpy restatement of the reference Fortran's algorithm
(hiiragiyua/shear_nnf: cross.f90, diffy_shear_hepta.f90, fou3D_thread.f,
operators.f90, initcr.f90).  It exists only to check the HIP product path:
only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import it.  The product library never calls into it.

The reference as a whole is not buildable in this image (its FFTW include is
an HTML error page, it needs MPI and an HDF5 Fortran module built by another
compiler).  Its compact y-operators are: ``oracle/Makefile`` compiles
mod.f90 + diffy_shear_hepta.f90 in place with ROCm flang into
``oracle/_ref/libshear_ref.so`` and ``oracle/ref.py`` calls its derivyc and
lapcdy (``tests/test_ref_compact.py``).  Every function here is pinned by
the tests:

* ``compact``  -- the reference's own derivyc/lapcdy (above), coefficient
  design points and Taylor order, the twisted-Fourier eigenvalue identity,
  dense brute-force solves (``tests/test_oracle_compact.py``);
* ``fftxz``    -- textbook DFT sums and the 3/2-rule index map;
* ``cross``    -- curl/divergence identities of the KMM reconstruction,
  orthogonality of the dealiased u x omega term to u, Parseval for the 'f'
  second moments and means of ``_hist_f`` (``tests/test_oracle_xz.py``);
  the Kelvin-mode closed form (``rdt``, rdtfunc) for the composition of
  ``Cross.run`` (convergence at the scheme's order) and for every UNER
  record and per-plane statistic, the lift-up closed form of a kx = 0 mode
  and diffusing mean profiles for the kx = 0 plane and the U, W path
  (``tests/test_oracle_rdt.py``); getv round
  trips, twisted eigenmodes and Parseval for ``getv``/``energy``
  (``tests/test_oracle_getv.py``);
* ``rdt``      -- a high-accuracy integration of the Kelvin-mode equations.

Also pinned, by ``tests/test_oracle_pins.py`` and ``tests/test_ref_compact.py``:
the 00-mode forcing rf0u/rf0w (plane means of u x omega formed by direct
double DFT sums), every branch of ``_new_dt`` (states with exact physical
maxima: the umax(0) x term, the y and z terms, the explicit viscous cap, the
dt < 1e-10 stop, the dump and etime clips, the umax >= 1 first step), the
``ias=0`` factor reuse (through the compiled derivyc) and the variant-2
dealiasing windows (mode by mode against hand-worked cut sets).  The
192x193x192 fixture ``tests/golden/config2_192x193x192.npz`` is this oracle's
10-step run, made by ``tests/golden/make_config2.py``.
"""
