"""CPU oracle for the closed-form alpha solve and the multilevel foreground.

TEST INFRASTRUCTURE ONLY. This package restates pymatting 1.1.15's algorithms
for the hot path on the CPU (NumPy/SciPy plus two small C files) so that the
HIP implementation in ``pymatting_amd`` can be checked against it. Only
``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import it, and only as the checker or as the timed CPU
baseline; the product package never imports it.

Modules and the reference code they restate:

- ``oracle.cf``: ``trimap_split``/``sanity_check_image``
  (pymatting/util/util.py:439-578), ``cf_laplacian``
  (pymatting/laplacian/cf_laplacian.py:5-177), ``make_linear_system``
  (pymatting/laplacian/laplacian.py:5-60), ``estimate_alpha_cf``
  (pymatting/alpha/estimate_alpha_cf.py:8-78).
- ``oracle.solver``: ``cg`` (pymatting/solver/cg.py:4-117), callbacks
  (pymatting/solver/callback.py:1-23), ``jacobi``
  (pymatting/preconditioner/jacobi.py:1-31), ``make_P``/``vcycle``
  (pymatting/preconditioner/vcycle.py:6-153).
- ``oracle.ichol``: thresholded incomplete Cholesky
  (pymatting/preconditioner/ichol.py:6-301), C restatement in ``ichol.c``.
- ``oracle.foreground``: ``estimate_foreground_ml``
  (pymatting/foreground/estimate_foreground_ml.py:5-253), C restatement in
  ``fb_ml.c``.
- ``oracle.foreground_cf``: ``estimate_foreground_cf``
  (pymatting/foreground/estimate_foreground_cf.py:8-130).

How each function is pinned (tests/test_oracle_*.py): known-answer values
from the reference's own tests and docstrings, brute-force dense assembly on
tiny images, closed-form invariants (symmetry, L*1 = 0, PSD, exact solves),
and the lemur fixtures with the tolerances the reference's tests use.
"""
