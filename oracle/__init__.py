"""CPU oracle for the 2D KIFMM hot path -- TEST INFRASTRUCTURE ONLY.

This package restates, in plain NumPy (FP64), the algorithm of the reference
flexmm ``kifmm2d`` path: the adaptive quadtree of ``flexmm/tree.py`` and the
kernel-independent FMM of ``flexmm/kifmm2d/scalar/{fmm,helpers}.py``, plus exact
FP64 direct sums.  It exists only to check the HIP implementation:

* only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
  ``bench.py`` may import it;
* nothing in the product package ``flexmm_amd`` imports it, and the product
  path never falls back to it.

How each function is pinned (independently of the restatement itself) is
listed in each module's header; see ``tests/test_oracle_*.py`` for the checks.
"""
