"""CPU float64 oracle for the SIPG operator and its CG solve.

TEST INFRASTRUCTURE ONLY.  Nothing in the product package (``ddg_amd``)
imports, calls or links this package; only ``tests/``, the smoke check in
``__graft_entry__.py`` and the ``cpu_baseline`` leg of ``bench.py`` use it,
and there only as the checker (or as the timed CPU restatement).

It is a NumPy/SciPy restatement of the reference's algorithm (JAX is not
available, so the reference itself cannot be run).  Each function cites the
reference ``file:line`` it follows.  Modules:

* ``reference_element`` -- 1D/2D nodal reference element
  (``src/ddg/dg1d/reference.py``, ``src/ddg/dg2d/reference.py``).
* ``mesh`` -- ``Mesh2D`` and its builders, literal loop transcription of
  ``src/ddg/dg2d/mesh.py`` (small meshes only).
* ``sipg`` -- the bilinear form (literal), its hand-written reverse-mode
  gradient (the reference computes ``jax.grad`` of the form), loads, dense
  assembly, LU and CG solves, block-Jacobi (``src/ddg/dg2d/sipg.py``).
* ``cg`` -- the semantics of ``jax.scipy.sparse.linalg.cg`` (third-party,
  unpinned JAX version; restated from its published source).
* ``multigrid`` -- p- and hp-multigrid V-cycles (``src/ddg/dg2d/multigrid.py``,
  ``src/ddg/dg2d/refinement.py`` uniform refinement only).

Parity pins (tests under ``tests/test_oracle_*.py``): closed-form known
answers of the reference's own tests, the literal bilinear form against the
hand-written gradient on unit vectors, symmetry, SPD-ness, manufactured
solutions, CG-vs-LU and preconditioner equivalences.  Unpinned: exact CG
iteration counts of JAX's ``cg`` (XLA reduction order) -- see the header of
``cg.py``.
"""
