"""The CPU oracle of the hip backend — TEST INFRASTRUCTURE, not product code.

A restatement of the reference's sequential backend (pyop2/sequential.py,
pyop2/host.py) and of its native helpers (pyop2/sparsity_utils.cxx,
pyop2/mat_utils.cxx, PETSc KSPCG + PCJACOBI), compiled with gcc.  Only
``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg
import it, as the checker; nothing under ``pyop2_amd/`` does.

How each part is pinned (known-answer tests in tests/test_oracle_*.py):

* loop and argument conventions: the reference unit tests' closed-form
  assertions (test_direct_loop, test_indirect_loop, test_global_reduction,
  test_vector_map, test_iteration_space_dats, test_constants), re-run on
  this oracle;
* sparsity: the reference fixture (test_matrices.py:54-62), brute force with
  numpy unique on random meshes, closed-form nnz of structured meshes;
* matrix values: the reference's expected matrices and right-hand sides
  (test_matrices.py:579-608) and closed forms (P1 mass |T|/12 (1+d_ij),
  Laplace rows summing to zero, total mass = area);
* CG: x = f for M x = M f (test_matrices.py:647-650, 780-783), scipy's
  direct solve on random SPD systems; PETSc iteration counts and residual
  histories are parity unpinned (PETSc cannot run here).
"""
