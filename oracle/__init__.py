"""CPU oracle for the graph-state Born sampler (test infrastructure only).

This package restates, in NumPy complex128 / float64 / uint64 arithmetic, the
algorithms of the reference's CPU and GPU paths so that the HIP kernels can be
checked against them.  It is the checker, never the product: only the tests,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` import
it.  The product path (``rgs70_classical_challenge_amd``) never imports this
package and fails loudly when its HIP library is missing.  Seeded inputs come
from the product's pure-Python ``rgs70_classical_challenge_amd.instances``
(input generation, not computation under test).

Modules
  graph_state quadratic phase, Walsh transform, 2x2 gates, full states,
              GF(2) algebra, the conditional / Fourier-covariant / deferred
              conditional states and the chunked lookup tables.
  draw        the sampling semantics: the reference's searchsorted draw and
              the pairwise FP64 mass pyramid descent used on the device.
  scan        the linear rank-width scan: plan compiler and executor.
  made        float64 MADE network, gradients and Adam for gradient checks.
  stats       linear XEB, bootstrap and Hoeffding statistics.
"""
