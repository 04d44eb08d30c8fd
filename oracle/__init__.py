"""Oracle: CPU restatement of the reference's algorithms for the MD hot path.

TEST INFRASTRUCTURE ONLY. Nothing in the product (``timemachine_amd``) may
import or call this package; it is the checker the HIP path is compared
against, used by ``tests/``, ``__graft_entry__.smoke()`` and the CPU-baseline
leg of ``bench.py``.

Modules (each function cites the reference file:line it restates):
  nonbonded   FP64 nonbonded_v3 (dense) and a pair-list FP64 evaluator, with
              Torch FP64 autograd standing in for jax.grad
  bonded      FP64 harmonic bond / cosine angle / periodic torsion
  nblist      NumPy per-block neighborlist (f64) and the exact FP32 predicate
              (C restatement in nblist_ref.c, built by oracle/build.py)
  integrator  Langevin coefficients and the BAOAB step
  barostat    MC barostat move and acceptance
  fire        FIRE minimisation with the reference's numpy types

Each function is pinned by golden data, closed forms, brute force, finite
differences or invariants in tests/test_oracle.py (and tests/test_fire.py).
"""
