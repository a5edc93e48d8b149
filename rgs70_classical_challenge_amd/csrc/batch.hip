// Many states in few launches: the slab pipeline, mass pyramids and draws of
// a batch of independent states, grouped by pass plan.
//
// States whose descriptors give the same pass schedule (same size and tile
// plan) form a group; each pass of a group is one k_tile_multi launch over
// the tiles of all its states (equal counts per state, so a tile's state is
// its index shifted).  The pyramids of all states are built level
// chunk by level chunk (one k_pyramid_multi launch per chunk round), and all
// draws of all states are one k_draw_multi launch.  The per-state kernel
// arguments and the job tables are built on the host once, at creation, and
// uploaded to one device allocation; running the batch only enqueues kernel
// launches, so it can be captured in a graph.
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "multi.h"
#include "rgs_amd.h"
#include "rgs_status.h"
#include "slab_plan.h"

namespace rgs {

// States of at least 2^kSingleShift tiles per pass fill the device on their
// own: their passes are launched one state at a time with the arguments in
// the kernel's argument segment (k_tile), which measured faster for large
// states than reading them from global memory in a shared launch.
constexpr int kSingleShift = 11;

struct BatchPass {
  PassPlan plan;
  size_t shmem;
  int count;                    // states in the group
  int shift;                    // log2 tiles per state
  size_t args_off;              // byte offset in the device allocation
  std::vector<PassArgs> host;   // per-state launches (shift >= kSingleShift)
};

struct BatchRound {
  int count;
  uint32_t groups;
  size_t jobs_off, first_off;
};

struct SlabBatch {
  void* dev = nullptr;
  std::vector<BatchPass> passes;
  std::vector<BatchRound> rounds;
  int ngroups = 0;
  int draw_count = 0;
  uint32_t draws = 0;
  size_t djobs_off = 0, dfirst_off = 0;
};

// A host image of the device allocation, filled at aligned offsets.
struct Image {
  std::vector<unsigned char> bytes;
  size_t put(const void* p, size_t n) {
    size_t off = (bytes.size() + 255) & ~(size_t)255;
    bytes.resize(off + n);
    std::memcpy(bytes.data() + off, p, n);
    return off;
  }
};

static std::string plan_key(const std::vector<PassPlan>& plan, const rgs_slab_desc* d) {
  std::string k = std::to_string(d->nbits) + "|";
  for (const PassPlan& p : plan) {
    k += std::to_string(p.flags) + "," + std::to_string(p.low) + "," + std::to_string(p.qstart) + "," +
         std::to_string(p.count) + "," + std::to_string(p.tb) + "," + std::to_string(p.c) + "," +
         std::to_string(p.rb) + "," + std::to_string(lds_bytes(p, d->nbits, d)) + ";";
  }
  return k;
}

}  // namespace rgs

using namespace rgs;

extern "C" int rgs_slab_batch_create(const rgs_slab_batch_item* items, int count, void** batch) {
  if (!items || count <= 0 || !batch) return fail(kInvalid, "slab_batch: need items and an output handle");
  *batch = nullptr;
  std::map<std::string, std::vector<int>> groups;
  std::vector<std::vector<PassPlan>> plans(count);
  for (int i = 0; i < count; ++i) {
    const rgs_slab_desc* d = items[i].desc;
    if (!d || !d->state || !d->mass) return fail(kInvalid, "slab_batch: every item needs a state and a mass level");
    if (!items[i].uniforms || !items[i].out_index || !items[i].out_p || items[i].samples < 0)
      return fail(kInvalid, "slab_batch: draw buffers missing");
    if (d->mass_bits < 2 || d->mass_bits > 8) return fail(kInvalid, "slab_batch: draws need 2 <= mass_bits <= 8");
    if (d->defer_leaves) return fail(kInvalid, "slab_batch: items draw from finished states (defer_leaves = 0)");
    std::string err;
    int rc = plan_passes(d, plans[i], err);
    if (!rc) rc = check_desc(d, err);
    if (rc) return fail(rc, "slab_batch item " + std::to_string(i) + ": " + err);
    groups[plan_key(plans[i], d)].push_back(i);
  }
  auto* b = new SlabBatch();
  b->ngroups = (int)groups.size();
  Image img;
  for (auto& kv : groups) {
    const std::vector<int>& idx = kv.second;
    const rgs_slab_desc* d0 = items[idx[0]].desc;
    const std::vector<PassPlan>& plan = plans[idx[0]];
    for (const PassPlan& p : plan) {
      BatchPass bp;
      bp.plan = p;
      bp.shmem = lds_bytes(p, d0->nbits, d0);
      bp.count = (int)idx.size();
      // every state of the group has the same size (nbits is in the key)
      bp.shift = d0->nbits - p.tb;
      std::vector<PassArgs> args(idx.size());
      for (size_t k = 0; k < idx.size(); ++k) build_args(items[idx[k]].desc, p, args[k]);
      // one launch's work-items (tiles x 2^(tb-4) threads) must stay below 2^31
      if ((((uint64_t)idx.size() << bp.shift) << (p.tb - 4)) >= (1ull << 31)) {
        delete b;
        return fail(kInvalid, "slab_batch: a group's pass needs 2^31 or more work-items; split the batch");
      }
      if (bp.shift >= kSingleShift) {
        bp.host = args;
        bp.args_off = 0;
      } else {
        bp.args_off = img.put(args.data(), args.size() * sizeof(PassArgs));
      }
      b->passes.push_back(bp);
    }
  }
  // pyramid rounds: round r reduces the r-th chunk of levels of every state
  std::vector<std::vector<PyrJob>> rjobs;
  std::vector<std::vector<uint32_t>> rgroups;
  for (int i = 0; i < count; ++i) {
    const rgs_slab_desc* d = items[i].desc;
    uint64_t off = 0;
    int bits = d->nbits - d->mass_bits;
    for (int r = 0; bits > 0; ++r) {
      const int k = std::min(bits, 9);
      PyrJob j{d->mass, off, off + (1ull << bits), bits, k};
      if ((int)rjobs.size() <= r) {
        rjobs.emplace_back();
        rgroups.emplace_back();
      }
      rjobs[r].push_back(j);
      rgroups[r].push_back((uint32_t)(1ull << (bits - k)));
      for (int l = 0; l < k; ++l) {
        off += 1ull << bits;
        bits -= 1;
      }
    }
  }
  for (size_t r = 0; r < rjobs.size(); ++r) {
    std::vector<uint32_t> first(rjobs[r].size() + 1, 0);
    uint64_t total = 0;
    for (size_t k = 0; k < rjobs[r].size(); ++k) {
      first[k] = (uint32_t)total;
      total += rgroups[r][k];
    }
    if (total >= (1ull << 23)) {      // 256 threads each: below 2^31 work-items
      delete b;
      return fail(kInvalid, "slab_batch: a pyramid round needs 2^23 or more workgroups; split the batch");
    }
    first[rjobs[r].size()] = (uint32_t)total;
    BatchRound br{(int)rjobs[r].size(), (uint32_t)total, 0, 0};
    br.jobs_off = img.put(rjobs[r].data(), rjobs[r].size() * sizeof(PyrJob));
    br.first_off = img.put(first.data(), first.size() * sizeof(uint32_t));
    b->rounds.push_back(br);
  }
  // draws
  std::vector<DrawJob> djobs;
  std::vector<uint32_t> dfirst;
  uint64_t draws = 0;
  for (int i = 0; i < count; ++i) {
    const rgs_slab_desc* d = items[i].desc;
    if (items[i].samples == 0) continue;
    djobs.push_back(DrawJob{reinterpret_cast<const float2*>(d->state), d->mass, items[i].uniforms, items[i].out_index,
                            items[i].out_p, d->nbits, d->mass_bits, items[i].flip});
    dfirst.push_back((uint32_t)draws);
    draws += (uint64_t)items[i].samples;
  }
  if (draws >= (1ull << 25)) {       // one 64-lane workgroup each: below 2^31 work-items
    delete b;
    return fail(kInvalid, "slab_batch: 2^25 or more draws in one batch; split the batch");
  }
  dfirst.push_back((uint32_t)draws);
  b->draw_count = (int)djobs.size();
  b->draws = (uint32_t)draws;
  if (!djobs.empty()) {
    b->djobs_off = img.put(djobs.data(), djobs.size() * sizeof(DrawJob));
    b->dfirst_off = img.put(dfirst.data(), dfirst.size() * sizeof(uint32_t));
  }
  hipError_t e = hipMalloc(&b->dev, std::max<size_t>(img.bytes.size(), 256));
  if (e != hipSuccess) {
    delete b;
    return fail(kNoMemory, std::string("slab_batch: ") + hipGetErrorString(e));
  }
  e = hipMemcpy(b->dev, img.bytes.data(), img.bytes.size(), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    (void)hipFree(b->dev);
    delete b;
    return fail(kHip, std::string("slab_batch upload: ") + hipGetErrorString(e));
  }
  *batch = b;
  return kOk;
}

extern "C" int rgs_slab_batch_run(void* batch, void* stream) {
  if (!batch) return fail(kInvalid, "slab_batch: null handle");
  const SlabBatch* b = static_cast<const SlabBatch*>(batch);
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  unsigned char* base = static_cast<unsigned char*>(b->dev);
  for (const BatchPass& p : b->passes) {
    if (!p.host.empty()) {
      for (const PassArgs& a : p.host) {
        hipError_t e = launch_pass(a, p.plan, p.shmem, st);
        if (e != hipSuccess) return fail(kHip, std::string("k_tile launch: ") + hipGetErrorString(e));
      }
      continue;
    }
    hipError_t e = launch_pass_multi(reinterpret_cast<const PassArgs*>(base + p.args_off), p.shift, p.count, p.plan,
                                     p.shmem, st);
    if (e != hipSuccess) return fail(kHip, std::string("k_tile_multi launch: ") + hipGetErrorString(e));
  }
  for (const BatchRound& r : b->rounds) {
    hipError_t e = launch_pyramid_multi(reinterpret_cast<const PyrJob*>(base + r.jobs_off),
                                        reinterpret_cast<const uint32_t*>(base + r.first_off), r.count, r.groups, st);
    if (e != hipSuccess) return fail(kHip, std::string("k_pyramid_multi launch: ") + hipGetErrorString(e));
  }
  if (b->draw_count) {
    hipError_t e = launch_draw_multi(reinterpret_cast<const DrawJob*>(base + b->djobs_off),
                                     reinterpret_cast<const uint32_t*>(base + b->dfirst_off), b->draw_count, b->draws,
                                     st);
    if (e != hipSuccess) return fail(kHip, std::string("k_draw_multi launch: ") + hipGetErrorString(e));
  }
  return kOk;
}

extern "C" int rgs_slab_batch_info(void* batch, int* launches, int* groups) {
  if (!batch) return fail(kInvalid, "slab_batch: null handle");
  const SlabBatch* b = static_cast<const SlabBatch*>(batch);
  if (launches) {
    size_t n = b->rounds.size() + (b->draw_count ? 1 : 0);
    for (const BatchPass& p : b->passes) n += p.host.empty() ? 1 : p.host.size();
    *launches = (int)n;
  }
  if (groups) *groups = b->ngroups;
  return kOk;
}

extern "C" int rgs_slab_batch_destroy(void* batch) {
  if (!batch) return kOk;
  SlabBatch* b = static_cast<SlabBatch*>(batch);
  hipError_t e = b->dev ? hipFree(b->dev) : hipSuccess;
  delete b;
  if (e != hipSuccess) return fail(kHip, std::string("slab_batch free: ") + hipGetErrorString(e));
  return kOk;
}
