"""CPU oracle for the Wilson-fermion hot path -- TEST INFRASTRUCTURE ONLY.

This package is a clean-room NumPy (complex128) restatement of the reference
TfLeonardYM algorithms on the fermion-matrix path.  It exists to *check* the
HIP implementation in ``tfleonardym_amd``; it is never part of the product
path.  Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline``
leg of ``bench.py`` may import it.

Modules
-------
``oracle.lattice``   neighbour tables, translate, antiperiodic boundary sign
                     (lattice.py:86-128, boundary_conditions.py:10-46)
``oracle.wilson``    gamma matrices, Wilson D in the four (mode, representation)
                     variants, gamma5, adjoint links, dense-matrix brute force
                     and the free-field closed form (dirac.py:253-523,649-662)
``oracle.solvers``   multishift MR and BiCGStab exactly as the reference runs
                     them (dirac_solver.py:13-110), plus textbook CG / CG-M

How each function is pinned (the reference ships no golden vectors and its
TensorFlow dependency is not installed, so nothing can be imported from it):

* tables / translate       brute force over coordinates (tests/test_oracle_lattice.py)
* gamma matrices           Clifford algebra, hermiticity, gamma5 product
* Wilson D (all variants)  explicit dense-matrix assembly by loops on tiny
                           lattices, free-field plane-wave closed form, gauge
                           covariance, gamma5-hermiticity, projected == simple,
                           adjoint == fundamental with adjoint links
* solvers                  numpy.linalg.solve on tiny lattices, true residuals,
                           MR and BiCGStab agreement (tests.py:99-119 invariants)
"""
