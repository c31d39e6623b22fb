"""CPU oracle for the indirect par_loop path -- TEST INFRASTRUCTURE ONLY.

This package restates, on the host, what the reference PyOP2 does for the
indirect par_loop hot path, so that the HIP implementation in ``pyop2_amd``
can be checked against it:

* :mod:`oracle.seqwrap` -- the sequential per-element wrapper the reference
  generates (``pyop2/codegen/builder.py`` pack / kernel call / unpack order,
  Mat insertion as ``MatSetValuesBlockedLocal``), emitted as C and compiled
  with gcc around the user's C-string kernel.
* :mod:`oracle.sparsity` -- the sparsity pattern closed form of
  ``pyop2/sparsity.pyx`` (sorted unique union of element pairs plus the
  diagonal) and the block/dof ``nnz`` unwinding.
* :mod:`oracle.colouring` -- the greedy bit-mask colouring documented in
  ``doc/sphinx/source/plan.rst:46-80``, in C (``oracle/colour.c``).

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU baseline
may import this package, and only as the checker or the CPU baseline.  The
product (``pyop2_amd``) never imports it and has no CPU fallback.

Pinning: every function here is checked by the tests in
``tests/test_oracle_*.py`` against the reference's own known-answer tests
(``test/unit/test_indirect_loop.py``, ``test_matrices.py``,
``test_vector_map.py``, ``test_iteration_space_dats.py``), against closed
forms (lumped mass, P1 stiffness, rigid-body modes, patch tests, brute-force
quadrature of the P2 stiffness) and against brute force on tiny inputs.
"""
