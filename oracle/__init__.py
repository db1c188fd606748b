"""CPU oracle for the JAXTrace particle-advection and density hot path.

TEST INFRASTRUCTURE ONLY.  This package restates, op by op in NumPy float32
(with float64 twins where noted), the algorithm the reference JAXTrace runs
under ``jax.jit``.  Only ``tests/``, ``__graft_entry__.smoke()`` and, in ``bench.py``, the
``cpu_baseline`` leg and the float64 reference tracks of the trajectory-error
check may import it, and only as the checker or as the timed CPU baseline; the product package ``jaxtrace_amd`` never imports
it and has no CPU fallback.

Every function cites the reference file:line it follows (paths relative to the
reference repository root).  Conventions shared by all modules:

* float32 arithmetic is IEEE round-to-nearest, one operation at a time, in the
  reference's evaluation order (left to right for ``a + b + c``);
* ``jmod`` mirrors ``jnp.mod`` on floats: ``fmod`` plus a sign fix-up
  (jax/_src/numpy/ufuncs.py ``remainder``), not ``x - floor(x/y)*y``;
* gathers clamp out-of-range indices to the array bounds, which is XLA's
  gather semantics for ``x[idx]`` retrieval.

Parity status per module is stated in each module header; the ones the
reference's own tests cannot pin (the reference has no golden vectors) are
pinned by closed forms, invariants and brute force in ``tests/``.
"""
