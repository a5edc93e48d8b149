// FP64 mass pyramid, device draws and probability gathers on a complex64 state.
#include "rgs_common.h"
#include "rgs_amd.h"
#include "multi.h"
#include "slab_plan.h"

#include <algorithm>

namespace rgs {

// ------------------------------------------------------------------ pyramid
// One workgroup reduces 2^k consecutive entries of level L into levels L+1..L+k.
__device__ __forceinline__ void pyramid_group(double* pyr, uint64_t level_off, int level_bits, int k,
                                              uint64_t out_off_first, uint64_t group) {
  __shared__ double s[512];
  const uint32_t tid = threadIdx.x;
  const uint32_t width = 1u << k;  // entries per group, <= 512
  const double* in = pyr + level_off + (group << k);
  for (uint32_t i = tid; i < width; i += 256) s[i] = in[i];
  __syncthreads();
  uint64_t out_off = out_off_first;
  int bits = level_bits;
  for (int l = 1; l <= k; ++l) {
    uint32_t cnt = width >> l;
    double val = 0.0;
    if (tid < cnt) val = __dadd_rn(s[2 * tid], s[2 * tid + 1]);
    __syncthreads();
    if (tid < cnt) {
      s[tid] = val;
      pyr[out_off + (group << (k - l)) + tid] = val;
    }
    __syncthreads();
    bits -= 1;
    out_off += 1ull << bits;
  }
}

__global__ void __launch_bounds__(256) k_pyramid(double* pyr, uint64_t level_off, int level_bits, int k,
                                                 uint64_t out_off_first) {
  pyramid_group(pyr, level_off, level_bits, k, out_off_first, blockIdx.x);
}

// The index of the job owning workgroup b: the last i with first[i] <= b.
__device__ __forceinline__ int job_of(const uint32_t* first, int count, uint32_t b) {
  int lo = 0, hi = count;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (first[mid] <= b) lo = mid;
    else hi = mid;
  }
  return lo;
}

__global__ void __launch_bounds__(256) k_pyramid_multi(const PyrJob* __restrict__ jobs,
                                                       const uint32_t* __restrict__ first, int count) {
  const int i = job_of(first, count, blockIdx.x);
  const PyrJob& j = jobs[i];
  pyramid_group(j.pyr, j.level_off, j.level_bits, j.k, j.out_off, blockIdx.x - first[i]);
}

// ------------------------------------------------------------------ draw
// One wavefront per draw.  Lanes hold 4 consecutive outcomes of the leaf
// block (mass_bits <= 8 -> 256 amplitudes); outcome x is read from state
// index x ^ flip, and the pyramid is in outcome order.  Level l of the
// pyramid (0 = base with 2^B entries) starts at 2^(B+1) - 2^(B+1-l).
__device__ __forceinline__ uint64_t level_off(int B, int l) {
  return (1ull << (B + 1)) - (1ull << (B + 1 - l));
}

// Pyramid descent from the root to a leaf block: t = u * root, then at each
// level go to child 1 iff t >= mass(child 0) and mass(child 1) > 0,
// subtracting mass(child 0).  Returns the block; t is what is left of it.
__device__ __forceinline__ uint64_t descend(const double* pyr, int B, double u, double& t) {
  t = __dmul_rn(u, pyr[level_off(B, B)]);
  uint64_t idx = 0;
  for (int l = B - 1; l >= 0; --l) {
    const double* lev = pyr + level_off(B, l);
    double left = lev[2 * idx], right = lev[2 * idx + 1];
    if (t >= left && right > 0.0) {
      t = __dsub_rn(t, left);
      idx = 2 * idx + 1;
    } else {
      idx = 2 * idx;
    }
  }
  return idx;
}

// The draw inside leaf block idx (outcomes bbase .. bbase + 2^mass_bits - 1,
// bbase = idx * 2^mass_bits) with the remainder t, one wavefront: outcome x
// is read from src[(x ^ flip) & smask] (the state itself, smask = ~0, or a
// recomputed tile holding it).
__device__ __forceinline__ void draw_leaf(const float2* src, uint64_t smask, int mass_bits, uint64_t idx, double t,
                                          uint64_t flip, int64_t* out_index, double* out_p) {
  const int lane = threadIdx.x;
  // Leaf block: recompute the pairwise tree with the same association.
  const int bsize = 1 << mass_bits;
  const uint64_t bbase = idx << mass_bits;
  double p[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    int l = lane * 4 + j;
    p[j] = l < bsize ? prob64(src[((bbase + l) ^ flip) & smask]) : 0.0;
  }
  double lv[7];
  lv[0] = __dadd_rn(__dadd_rn(p[0], p[1]), __dadd_rn(p[2], p[3]));
#pragma unroll
  for (int b = 0; b < 6; ++b) lv[b + 1] = __dadd_rn(lv[b], __shfl_xor(lv[b], 1 << b));
  // Descend over local bits mass_bits-1 .. 2 (lane bit b = bit - 2).
  int node = 0;
#pragma unroll
  for (int b = 5; b >= 0; --b) {
    if (b + 2 <= mass_bits - 1) {
      int left_lane = node << (b + 1);
      double left = __shfl(lv[b], left_lane);
      double right = __shfl(lv[b], left_lane | (1 << b));
      if (t >= left && right > 0.0) {
        t = __dsub_rn(t, left);
        node = 2 * node + 1;
      } else {
        node = 2 * node;
      }
    }
  }
  double q0 = __shfl(p[0], node), q1 = __shfl(p[1], node), q2 = __shfl(p[2], node), q3 = __shfl(p[3], node);
  double a01 = __dadd_rn(q0, q1), a23 = __dadd_rn(q2, q3);
  int sub;
  double ps;
  if (t >= a01 && a23 > 0.0) {
    t = __dsub_rn(t, a01);
    if (t >= q2 && q3 > 0.0) { sub = 3; ps = q3; } else { sub = 2; ps = q2; }
  } else {
    if (t >= q0 && q1 > 0.0) { sub = 1; ps = q1; } else { sub = 0; ps = q0; }
  }
  if (lane == 0) {
    *out_index = (int64_t)(bbase + (uint64_t)(node * 4 + sub));
    *out_p = ps;
  }
}

__device__ __forceinline__ void draw_one(const float2* state, int nbits, const double* pyr, int mass_bits, double u,
                                         uint64_t flip, int64_t* out_index, double* out_p) {
  double t;
  const uint64_t idx = descend(pyr, nbits - mass_bits, u, t);
  draw_leaf(state, ~0ull, mass_bits, idx, t, flip, out_index, out_p);
}

__global__ void __launch_bounds__(64) k_draw(const float2* state, int nbits, const double* pyr, int mass_bits,
                                             const double* uniforms, int count, uint64_t flip, int64_t* out_index,
                                             double* out_p) {
  const int d = blockIdx.x;
  if (d >= count) return;
  draw_one(state, nbits, pyr, mass_bits, uniforms[d], flip, out_index + d, out_p + d);
}

__global__ void __launch_bounds__(64) k_draw_multi(const DrawJob* __restrict__ jobs, const uint32_t* __restrict__ first,
                                                   int count) {
  const int i = job_of(first, count, blockIdx.x);
  const DrawJob& j = jobs[i];
  const uint32_t d = blockIdx.x - first[i];
  draw_one(j.state, j.nbits, j.pyr, j.mass_bits, j.uniforms[d], j.flip, j.out_index + d, j.out_p + d);
}

hipError_t launch_pyramid_multi(const PyrJob* jobs, const uint32_t* first, int count, uint32_t groups,
                                hipStream_t st) {
  hipLaunchKernelGGL(k_pyramid_multi, dim3(groups), dim3(256), 0, st, jobs, first, count);
  return hipGetLastError();
}

hipError_t launch_draw_multi(const DrawJob* jobs, const uint32_t* first, int count, uint32_t draws, hipStream_t st) {
  hipLaunchKernelGGL(k_draw_multi, dim3(draws), dim3(64), 0, st, jobs, first, count);
  return hipGetLastError();
}

// Deferred leaves: the descent of every draw (one thread each), recording
// the block, the remainder and the leaf-bits tile holding the block...
__global__ void __launch_bounds__(256) k_draw_descend(const double* pyr, int nbits, int mass_bits, int leaf_bits,
                                                      const double* uniforms, int count, uint64_t flip,
                                                      int64_t* tiles, int64_t* blocks, double* rest) {
  const int d = blockIdx.x * 256 + threadIdx.x;
  if (d >= count) return;
  double t;
  const uint64_t idx = descend(pyr, nbits - mass_bits, uniforms[d], t);
  blocks[d] = (int64_t)idx;
  rest[d] = t;
  tiles[d] = (int64_t)(((idx << mass_bits) ^ flip) >> leaf_bits);
}

// ... then, after the tiles are recomputed (draw d's into slot d), the
// leaf-block draws.
__global__ void __launch_bounds__(64) k_draw_leaf(const float2* leaves, int leaf_bits, int mass_bits,
                                                  const int64_t* blocks, const double* rest, int count,
                                                  uint64_t flip, int64_t* out_index, double* out_p) {
  const int d = blockIdx.x;
  if (d >= count) return;
  draw_leaf(leaves + ((uint64_t)d << leaf_bits), (1ull << leaf_bits) - 1ull, mass_bits, (uint64_t)blocks[d],
            rest[d], flip, out_index + d, out_p + d);
}

__global__ void k_gather_prob(const float2* state, const int64_t* index, int count, double* out) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) out[i] = prob64(state[index[i]]);
}

}  // namespace rgs

using namespace rgs;

extern "C" int rgs_mass_pyramid(double* pyr, int base_bits, void* stream) {
  if (!pyr || base_bits < 0 || base_bits > 40) return fail(kInvalid, "bad pyramid");
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  uint64_t off = 0;
  int bits = base_bits;
  while (bits > 0) {
    int k = std::min(bits, 9);
    uint64_t groups = 1ull << (bits - k);
    uint64_t out_off = off + (1ull << bits);
    hipLaunchKernelGGL(k_pyramid, dim3((unsigned)groups), dim3(256), 0, st, pyr, off, bits, k, out_off);
    RGS_LAUNCH_CHECK("k_pyramid");
    // advance past k levels
    for (int l = 0; l < k; ++l) {
      off += 1ull << bits;
      bits -= 1;
    }
  }
  return kOk;
}

extern "C" int rgs_draw(const void* state, int nbits, const double* pyr, int mass_bits, const double* uniforms,
                        int count, uint64_t flip, int64_t* out_index, double* out_p, void* stream) {
  if (!state || !pyr || !uniforms || !out_index || !out_p) return fail(kInvalid, "null pointer");
  if (mass_bits < 2 || mass_bits > 8 || nbits - mass_bits < 0 || nbits - mass_bits > 47)
    return fail(kInvalid, "draw needs 2 <= mass_bits <= 8");
  if (count <= 0) return kOk;
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  hipLaunchKernelGGL(k_draw, dim3(count), dim3(64), 0, st, reinterpret_cast<const float2*>(state), nbits, pyr,
                     mass_bits, uniforms, count, flip, out_index, out_p);
  RGS_LAUNCH_CHECK("k_draw");
  return kOk;
}

extern "C" int rgs_gather_prob(const void* state, const int64_t* index, int count, double* out_p, void* stream) {
  if (!state || !index || !out_p) return fail(kInvalid, "null pointer");
  if (count <= 0) return kOk;
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  hipLaunchKernelGGL(k_gather_prob, dim3((count + 255) / 256), dim3(256), 0, st,
                     reinterpret_cast<const float2*>(state), index, count, out_p);
  RGS_LAUNCH_CHECK("k_gather_prob");
  return kOk;
}

// Workspace of rgs_draw_leaves: tile, block and remainder per draw, then the
// recomputed tiles from a 256-byte boundary.
static size_t leaves_head(int count) { return (((size_t)count * 24) + 255) & ~(size_t)255; }

extern "C" int rgs_draw_leaves_workspace(const rgs_slab_desc* d, int count, size_t* bytes) {
  if (!bytes || count < 0) return fail(kInvalid, "bad workspace query");
  int lb = 0;
  int rc = leaf_bits(d, &lb);
  if (rc) return rc;
  *bytes = leaves_head(count) + ((size_t)count << lb) * sizeof(float2);
  return kOk;
}

extern "C" int rgs_draw_leaves(const rgs_slab_desc* d, const double* pyr, const double* uniforms, int count,
                               void* work, size_t work_bytes, int64_t* out_index, double* out_p, void* stream) {
  if (!pyr || !uniforms || !work || !out_index || !out_p) return fail(kInvalid, "null pointer");
  if (count < 0 || count > (1 << 24)) return fail(kInvalid, "draw count must be in 0..2^24");
  int lb = 0;
  int rc = leaf_bits(d, &lb);
  if (rc) return rc;
  const int mb = d->mass_bits;
  if (mb < 2 || mb > 8 || mb > lb) return fail(kInvalid, "draw needs 2 <= mass_bits <= min(8, leaf bits)");
  if (count == 0) return kOk;
  if ((reinterpret_cast<uintptr_t>(work) & 255) != 0) return fail(kInvalid, "workspace must be 256-byte aligned");
  if (work_bytes < leaves_head(count) + ((size_t)count << lb) * sizeof(float2))
    return fail(kInvalid, "workspace smaller than rgs_draw_leaves_workspace");
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  char* w = reinterpret_cast<char*>(work);
  int64_t* tiles = reinterpret_cast<int64_t*>(w);
  int64_t* blocks = tiles + count;
  double* rest = reinterpret_cast<double*>(blocks + count);
  float2* leaves = reinterpret_cast<float2*>(w + leaves_head(count));
  hipLaunchKernelGGL(k_draw_descend, dim3((count + 255) / 256), dim3(256), 0, st, pyr, d->nbits, mb, lb, uniforms,
                     count, d->flip, tiles, blocks, rest);
  RGS_LAUNCH_CHECK("k_draw_descend");
  rc = leaf_tiles(d, tiles, count, leaves, st);
  if (rc) return rc;
  hipLaunchKernelGGL(k_draw_leaf, dim3(count), dim3(64), 0, st, leaves, lb, mb, blocks, rest, count, d->flip,
                     out_index, out_p);
  RGS_LAUNCH_CHECK("k_draw_leaf");
  return kOk;
}
