"""CPU oracle for the CCD operator and the variable-density FVM and FCCD PPEs.

TEST INFRASTRUCTURE ONLY.  This package is a NumPy/SciPy restatement of the
reference algorithms, used as the checker by ``tests/``, by
``__graft_entry__.smoke()`` and by the ``cpu_baseline`` leg of ``bench.py``.
The product path (``twophaseflow_amd``) never imports it and fails loudly when
its HIP library is missing.

Each function cites the reference file:line it restates (paths relative to
``src/twophase`` of t0m0h1r0/TwoPhaseFlow).  Pinning (what checks each
function against something other than its own formula):

* ``ccd.differentiate_periodic``  -- closed-form Fourier symbol
  (``ccd.periodic_symbol``), paper table U1-a, ``test_ccd.py:65-81``.
* ``ccd.differentiate_wall_extended`` -- the same wall system in long
  double: cubic exactness, agreement with the dense LU on random data.
* ``ccd.differentiate_wall``      -- polynomial exactness (x^0..x^3),
  ``test_ccd.py:201-221``, D2 null space (``test_ccd.py:257-299``),
  convergence slopes, block-Thomas vs dense LU.
* ``ppe.fvm_apply``               -- assembled CSR twin, row sums, symmetry of
  the CG system, manufactured 1-D exact answers.
* ``ppe.solve_cg`` / ``ppe.solve_direct`` -- each other, paper table U2-b,
  and the FP64 true residual of the reference's CG system.
* ``fields``                       -- closed forms (sphere volumes, ψ bounds).
* ``fccd.OracleFCCDPPE``          -- the closed-form periodic symbol
  (``fccd.fccd_symbol``), cubic exactness of the wall face gradient and rows,
  the discrete divergence theorem, and a dense direct solve for its GMRES;
  its phase-separated forms by the reference's own cut-face examples, the
  closed-form affine drive of a straight interface, gauge fixed points with
  zero weighted phase means, exact recovery of J, and planted mistakes.
* ``rhie_chow.OracleRhieChow``    -- closed forms: linear velocity, cubic
  pressure bracket (interior and wall one-sided values), and the
  balanced-force equilibrium p = kappa psi / We.
* ``dccd.OracleDCCD``             -- the TVD-RK3 amplification polynomial of a
  Fourier mode, the filter's transfer factor, unfiltered wall ends, and
  mass restoration by the correction.
* ``curvature.OracleCurvaturePsi`` -- exact -1/r and -2/r on circle / sphere
  level sets of quadratic psi, the rotated-ellipse closed form (mixed term),
  the filter's action on linear fields and wall planes, and its transfer
  factor (1 - 4 eps (2 psi - 1)^2 sin^2(theta/2)) with a pointwise
  coefficient.
* ``reinit.OracleSplitReinit``    -- the CN-ADI closed-form factor on cosine
  modes, the pseudo-time step formula, mass restoration and the tanh profile
  as a near fixed point; the compression step by the filter's transfer
  factor, exact linear fluxes with each ghost rule and the monotone tanh
  divergence (three planted mistakes fail them).
* ``legacy_step`` (the config-1 step) -- the reference's own outputs on
  seeded inputs for every building block and its 64 x 64 trajectory
  (tests/golden/, written by make_legacy_step_fixtures.py from the
  reference source), and closed forms (the extension of a quadratic field,
  the viscous body of linear fields, the UCCD6 skew weights and
  hyperviscosity symbol).

No function here is "parity unpinned" except the GMRES/CG iteration counts,
which the reference does not fix (SciPy internals).
"""
