"""FP64 CPU oracle for the convergent Born series (CBS) path of BMOLKAIST/Inverse_solver.

TEST INFRASTRUCTURE ONLY.  Nothing in ``inverse_solver_amd`` imports this package;
only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may use it, and only as the checker (or, for ``cpu_baseline``, as the
timed CPU port of the reference loop).

It is a NumPy/SciPy restatement, function by function, of the reference's MATLAB
code.  Every function cites the reference file:line it follows.  The reference
itself (MATLAB R2021a + a CUDA MEX) cannot run here, so the oracle is pinned by
the checks listed below instead of by reference-produced vectors:

  optics.py   DERIVE_OPTICAL_TOOL, RI2potential, fft_flip, centred FFTs
              pinned by closed forms (tests/test_oracle_optics.py)
  hankel.py   JnRoots, discrete_hankel_croped_fourier, truncated_green_plus
              JnRoots pinned by the reference's own dht.mat table and SciPy;
              the Hankel pair by the disk -> 2*pi*J1(k)/k known answer that
              test_hankel.m uses; truncated_green_plus by the cyclic closed form
              limit and by symmetry properties.  MATLAB's single-precision
              rounding of this setup is NOT reproduced (parity pinned to the
              formula, not to bits).
  cbs.py      FORWARD_SOLVER_CONVERGENT_BORN (scalar) set_RI/init/solve/solve_raw
              pinned by dense-matrix brute force of the two-chain recurrence on
              tiny grids, the DFT shift theorem for the +-1/4-bin Green's
              function, linearity in the source, and the first-Born limit.
  backward.py BACKWARD_SOLVER_MULTI gradient + TV_FISTA_inner_v2
              pinned by the adjoint dot-product test, a finite-difference
              gradient check of the oracle's own cost, and TV-prox special cases.
  inputs.py   FIELD_GENERATOR 'circle', PHANTOM 'bead', seeded sphere phantoms
              pinned by closed forms (plane-wave positions, bead volume).

Every function above has at least one such check; none is "parity unpinned"
except where noted in the function's docstring.
"""
