// Multi-state launches of the pyramid and draw kernels: one launch serves
// many states, each workgroup finding its job by binary search over the
// jobs' first workgroup indices (first[count] = total workgroups).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rgs {

struct PyrJob {          // one k-level reduction of one state's pyramid
  double* pyr;
  uint64_t level_off;    // start of the input level
  uint64_t out_off;      // start of the first output level
  int level_bits;        // log2 entries of the input level
  int k;                 // levels reduced (<= 9)
};

struct DrawJob {         // all draws of one state
  const float2* state;
  const double* pyr;
  const double* uniforms;
  int64_t* out_index;
  double* out_p;
  int nbits;
  int mass_bits;
  uint64_t flip;         // outcome x is held at state index x ^ flip
};

hipError_t launch_pyramid_multi(const PyrJob* jobs, const uint32_t* first, int count, uint32_t groups,
                                hipStream_t st);
hipError_t launch_draw_multi(const DrawJob* jobs, const uint32_t* first, int count, uint32_t draws, hipStream_t st);

}  // namespace rgs
