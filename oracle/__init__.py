"""CPU oracle for the piezo-FEM frequency sweep.

TEST INFRASTRUCTURE ONLY.  This package restates the reference algorithm
(SawSim ``saw2d/piezo_fem.py``, ``saw2d/sweep.py``, ``saw2d/materials.py``,
``crystal.py``, ``metrics.py``) in NumPy/SciPy so that the HIP product path can
be checked against it.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import it; the product package
``sawsim_amd`` never does.

Pinning (what checks each function, besides retyping it):

* ``fem.q9_shapes`` / ``fem.gauss_3x3``: partition of unity, Kronecker delta at
  the nodes, exact integration of degree-5 polynomials (tests/test_oracle_pins.py).
* ``fem.element_geometry``: closed-form Gram matrices of rectangles
  (∫∇N·∇N and ∫N·N of the biquadratic Lagrange basis), sum(GM) = area, signed
  detJ for clockwise elements.
* ``fem.assemble`` / ``fem.assemble_me0``: brute-force element-loop B-matrix
  assembly ∫BᵀCB (tests/test_oracle_pins.py), rigid-translation null space of
  K_uu, constant-potential null space of K_φφ, K_φu = SCALE·K_uφᵀ, total mass
  = SCALE·Σρ·area, and the golden fixtures written from the reference
  (tests/golden/).
* ``materials.rotate``: brute-force 4th-rank tensor rotation, identity angles,
  invariance of the permittivity eigenvalues; crystal builders against the
  library records (linbo3_bouchy_2022, sio2).
* ``sweep.bc_periodic`` / ``sweep.merge_transform``: closed-form pairings on the
  structured meshes and the golden DOF maps.
* ``sweep.FrequencySolver``: SciPy SuperLU (library routine) with a residual
  check; end-to-end |Y| against the reference-held curves
  ``tests/data/sp_double_layer.npz``/``sp_quad_layer.npz`` (copied fixtures).
* ``metrics.resonances``: the reference's synthetic-curve tests restated.
"""
