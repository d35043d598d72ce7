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
  the discrete divergence theorem, and a dense direct solve for its GMRES.
* ``rhie_chow.OracleRhieChow``    -- closed forms: linear velocity, cubic
  pressure bracket (interior and wall one-sided values), and the
  balanced-force equilibrium p = kappa psi / We.
* ``dccd.OracleDCCD``             -- the TVD-RK3 amplification polynomial of a
  Fourier mode, the filter's transfer factor, unfiltered wall ends, and
  mass restoration by the correction.
* ``curvature.OracleCurvaturePsi`` -- exact -1/r and -2/r on circle / sphere
  level sets of quadratic psi, the rotated-ellipse closed form (mixed term),
  and the filter's action on linear fields and wall planes.
* ``reinit.OracleSplitReinit``    -- the CN-ADI closed-form factor on cosine
  modes, the pseudo-time step formula, mass restoration and the tanh profile
  as a near fixed point.

No function here is "parity unpinned" except the GMRES/CG iteration counts,
which the reference does not fix (SciPy internals).
"""
