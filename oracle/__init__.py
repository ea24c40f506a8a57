"""FP64 CPU oracle for the convergent Born series (CBS) path of BMOLKAIST/Inverse_solver.

TEST INFRASTRUCTURE ONLY.  Nothing in ``inverse_solver_amd`` imports this package;
only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may use it, and only as the checker (or, for ``cpu_baseline``, as the
timed CPU port of the reference loop).

It is a NumPy/SciPy restatement, function by function, of the reference's MATLAB
code.  Every function cites the reference file:line it follows.  The reference
itself (MATLAB R2021a + a CUDA MEX) cannot run here, so the oracle is pinned by
the checks listed below instead of by reference-produced vectors:

  optics.py   DERIVE_OPTICAL_TOOL, RI2potential, fft_flip, centred FFTs,
              polarisation_utility
              pinned by closed forms (tests/test_oracle_optics.py); the
              polarisation basis by the Ewald-tangent closed form, transversality,
              energy preservation and the 2D->3D->2D round trip
              (tests/test_oracle_vector.py)
  hankel.py   JnRoots, discrete_hankel_croped_fourier, truncated_green_plus
              JnRoots pinned by the reference's own dht.mat table and SciPy;
              the Hankel pair by the disk -> 2*pi*J1(k)/k known answer that
              test_hankel.m uses and by a Gaussian pair; truncated_green_plus by
              its Hankel pieces and bin_search_nearest = searchsorted.  MATLAB's
              single-precision rounding of this setup is NOT reproduced (parity
              pinned to the formula, not to bits).
  cbs.py      FORWARD_SOLVER_CONVERGENT_BORN set_RI/init/solve/solve_raw, scalar
              and vector (dyadic Green's function) branches
              pinned by dense-matrix brute force of the two-chain recurrence on
              tiny grids (scalar and dyadic), G- = G+^T for the +-1/4-bin Green's
              function, linearity in the source, the mask/ramp/eps/Bornmax closed
              forms, and (vector) the scalar limit for normal incidence on a
              z-stratified potential.
  backward.py BACKWARD_SOLVER_MULTI gradient + TV_FISTA_inner_v2
              the gradient by the sign and bounded spread of finite-difference
              directional derivatives of the oracle's own misfit (partially
              pinned: the reference gradient is preconditioned, so its magnitude
              is pinned to the formula only); the TV prox by its special cases
              (lambda -> 0, constants are fixed points, objective decrease).

  rytov.py    BACKWARD_SOLVER_RYTOV, vector2scalarfield, the goldstein_GPU residue and
              raster-unwrap kernels
              pinned by a phase vortex (one residue, winding sign), smooth-map
              unwrapping to the true phase, step-map closed forms, a weak bead
              reconstructed from CBS-oracle fields, and the no-scatterer case;
              the raster update's single rounding against exact rationals.
  goldstein.py the goldstein_GPU branch-cut placement (dipoles, sorted residue
              lookup, 8-nearest box search, the CPU branch search, rastering)
              pinned by closed-form vortex pairs and border vortices (signs and
              cut geometry), edge-jump dipoles, brute-force nearest lists where
              the box search is exhaustive (and a hand case of its ring bound),
              and charge balance of every branch network.

The seeded input builders (phantoms, illumination) live in the product package
(inverse_solver_amd/phantoms.py) and are pinned by closed forms in
tests/test_phantoms.py.

Every function above has at least one such check; none is "parity unpinned"
except where noted in the function's docstring.
"""
