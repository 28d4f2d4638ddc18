"""CPU oracle for the LaPy FEM hot path -- TEST INFRASTRUCTURE ONLY.

This package restates, in numpy + scipy, the reference algorithms the GPU
library replaces:

* ``fem.fem_tria``        -- lapy/solver.py:106-194   (``Solver._fem_tria``)
* ``fem.fem_tria_aniso``  -- lapy/solver.py:197-308   (``Solver._fem_tria_aniso``)
* ``fem.fem_tria_mass``   -- lapy/solver.py:311-377   (``Solver.fem_tria_mass``)
* ``fem.fem_tetra``       -- lapy/solver.py:380-533   (``Solver._fem_tetra``)
* ``fem.eigs``            -- lapy/solver.py:667-716   (``Solver.eigs``: splu + eigsh)
* ``fem.poisson``         -- lapy/solver.py:718-889   (``Solver.poisson``)
* ``fem.diffusion``       -- lapy/heat.py:114-232     (``heat.diffusion``)
* ``fem.avg_edge_length`` -- lapy/tria_mesh.py:735-748, lapy/tet_mesh.py:182-195

and, in ``metrics``, the comparison measures used to compare GPU results with
it (eigenvalue relative error, per-cluster principal angles).

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this package, and only to check results or as the
CPU baseline.  The product package ``lapy_amd`` never imports it.

Pinning (how each function is checked independently of the GPU path):

* assembly (``fem_tria*``, ``fem_tetra``): closed-form element matrices of the
  unit right triangle / unit right tetrahedron, A*1 = 0, exact symmetry,
  sum(B) = area / volume, the linear-patch test, and the reference's
  ``cubeTria.ev`` / ``cubeTetra.ev`` / ``expected_outcomes.json`` eigenvalues
  through ``eigs`` (tests/test_oracle.py);
* ``eigs``: dense ``scipy.linalg.eigh`` on small meshes, the reference's
  ``.ev`` eigenvalues and eigenvectors, the continuum value pi^2 on the cube;
* ``poisson`` / ``diffusion``: dense direct solves on small meshes, Dirichlet
  re-insertion exactness, batch == single-column solves;
* ``avg_edge_length``: the reference's known answers (1.1380711874576983 for
  the 8-vertex cube surface, 1.0543647924813107 for the 9-vertex tet cube).
"""
