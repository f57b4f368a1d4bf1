"""CPU oracle for the even-odd Wilson-clover operator and its CG solve.

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and
the ``cpu_baseline`` leg of ``bench.py`` may import this package, and only as
the reference the GPU results are compared with (or as the timed CPU stand-in
for the reference's Grid CPU path).
The product (``gpt_amd``) never imports it and has no CPU fallback.

Contents, each restating a reference file (paths relative to the reference
repository root):

* ``conventions``  -- gamma matrices, sigma_{mu nu}, checkerboards, shifts
  (lib/gpt/core/gamma.py:28-53, lib/gpt/core/checkerboard.py:23-35,
  lib/gpt/core/covariant.py:28-59)
* ``rng``          -- gpt's lattice RNG (C restatement in ``gpt_rng.c``)
  (lib/gpt/core/random.py:23-148, lib/cgpt/lib/random/*.h)
* ``gauge``        -- SU(3) ``element`` / ``gauge.random``, matrix exp,
  plaquette, clover-leaf field strength and the seeded smooth-link generator
  (lib/gpt/qcd/gauge/create.py:66-71, lib/gpt/qcd/gauge/loops.py:162-174,
  lib/gpt/core/foundation/lattice/matrix/exp.py:167-220,
  lib/gpt/core/object_type/su_n.py:208-240)
* ``wilson_clover`` -- the Python reference operator
  (lib/gpt/qcd/fermion/reference/wilson_clover.py:27-281)
* ``solvers``      -- CG, Schur complement two + normal equation (eo2_ne),
  preconditioned, defect correcting, block CG
  (lib/gpt/algorithms/inverter/cg.py:24-117, preconditioned.py:24-54,
  defect_correcting.py:74-142, block_cg.py:36-149,
  lib/gpt/algorithms/preconditioner/schur_complement_two.py:58-152,
  normal_equation.py:35-65)
* ``wc_cpu.c``     -- C/OpenMP restatement of the EO hop, clover blocks and
  CG used as the timed CPU baseline (a stand-in for Grid's CPU path, which
  cannot be built here).

How each piece is pinned (known-answer values from the reference's own
tests, closed forms, brute force) is listed in the header of each module.
"""
