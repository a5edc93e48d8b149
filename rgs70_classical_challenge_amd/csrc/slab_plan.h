// Pass planning of the slab pipeline, shared by the single-state entry
// points (slab.hip) and the multi-state batch (batch.hip).
#pragma once
#include <string>
#include <vector>

#include "slab_kernel.h"

namespace rgs {

struct PassPlan {
  int flags;
  int low;     // 1: contiguous tile
  int qstart;
  int count;   // target bits
  int tb;      // tile bits
  int rbits;   // log2 registers per thread
  int c;       // lane bits
  int rb;      // log2 LDS rounds
  int first;   // 1: the schedule's first pass
};

// The pass schedule of a descriptor (RGS_EINVAL with a message if its tiles
// do not fit), the LDS bytes of one pass, the descriptor's table checks, and
// the kernel arguments of one pass.
int plan_passes(const rgs_slab_desc* d, std::vector<PassPlan>& out, std::string& err);
size_t lds_bytes(const PassPlan& p, int nbits, const rgs_slab_desc* d);
int check_desc(const rgs_slab_desc* d, std::string& err);
void build_args(const rgs_slab_desc* d, const PassPlan& p, PassArgs& a);
hipError_t launch_pass(const PassArgs& a, const PassPlan& p, size_t shmem, hipStream_t st);
// The deferred last pass's values of listed tiles (rgs_slab_leaf_tiles) and
// its tile bits.
int leaf_tiles(const rgs_slab_desc* d, const int64_t* tiles, int count, void* out, hipStream_t st);
int leaf_bits(const rgs_slab_desc* d, int* bits);
hipError_t launch_pass_multi(const PassArgs* args, int shift, int count, const PassPlan& p, size_t shmem,
                             hipStream_t st);

}  // namespace rgs
