"""MI355X-native exact conditional Born sampler for graph-state circuits.

Modules
  engine     GraphStateSampler (full state, plain method) and
             ConditionalSampler (balanced full-rank cut, the reference's
             production algorithm) over the fused HIP slab pipeline.
  instances  seeded random 3-regular / Erdos-Renyi graphs and QR unitaries.
  build      hipcc recipe for librgs_amd.so (gfx950).
  _lib       ctypes binding of the C ABI in include/rgs_amd.h.

Importing the package loads nothing; the HIP library is loaded on first use
and its absence raises RuntimeError (there is no CPU fallback).
"""
__all__ = ['engine', 'instances', 'build']
