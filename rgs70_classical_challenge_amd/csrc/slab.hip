// Slab pipeline on one 2^nbits complex64 state: fused init / Walsh /
// diagonal phase / real rotations / FP64 masses, a few HBM passes in total.
//
// Gates.  Every 2x2 unitary is factored on the host as u = D1 R D2 (real
// rotation R, diagonals D1 and D2 = diag(1, e^{ib})).  Because diagonals on
// different qubits commute with everything on other qubits, the product of all
// gates is (x)D1 . (x)R . (x)D2: the D2 layer is one pointwise phase
// e^{i sum_q b_q x_q} applied with the graph phase, the rotations cost four
// FMAs per amplitude and qubit, and the D1 layer, which leaves every
// probability unchanged, is applied when amplitudes are read back.
//
// Tiles.  A pass works on tiles of 2^TB amplitudes, one workgroup of
// 2^(TB-RB) threads per tile with 2^RB amplitudes per thread in registers.
// Tile-local bit L maps to a global bit: in the contiguous ("low") pass the
// map is the identity; in a strided pass local bits 0..c-1 are the contiguous
// lane bits 0..c-1 (rows of 2^c amplitudes) and local bits c..TB-1 are the
// target bits qstart..qstart+TB-c-1.  The workgroup index supplies the rest.
//
// Layouts.  In layout s the registers of a thread hold local bits s..s+RB-1
// and the thread index enumerates the other bits in increasing order, so lane
// bits are always the lowest thread bits and a wavefront reads whole rows in
// the load/store layout.  Butterflies act on register bits; a layout change
// goes through LDS (one padded float2 per 16).  Layouts visited: s_1 =
// max(TB-RB, c), s_{k+1} = max(s_k - RB, c), ...; set k owns the local bits
// [s_k, s_{k-1}) (s_0 = TB), so every bit is operated on exactly once.  When
// a tile does not fit in LDS, each layout change runs in 2^rb rounds, one per
// value of the top rb lane bits (thread bits in every layout).
#include "rgs_common.h"
#include "rgs_amd.h"

#include <algorithm>
#include <cstdio>
#include <vector>

namespace rgs {

enum : int { F_LOAD = 1, F_WHT = 2, F_PHASE = 4, F_GATES = 8, F_MASS = 16 };

constexpr size_t kLdsBudget = 160 * 1024;

struct PassArgs {
  float2* state;
  const uint64_t* init_lut;   // GRAPH: phase words [ic][256]; COV: map+phase [ic][256][2]
  const float2* init_w;       // COV: [ic][256]
  const uint64_t* phase_lut;  // [pc][256] sign words, or null
  const float2* dphase;       // [dc][256] complex diagonal phase, or null
  float2 rot[16];             // (cos, sin) of local target bit i (identity beyond gate_limit)
  double* mass;
  float init_scale;
  int nbits;
  int lane_bits;   // c
  int qstart;      // global bit of local bit c (0 for the low tile)
  int flags;
  int init_kind;
  int init_chunks;
  int phase_chunks;
  int dphase_chunks;
  int gate_limit;
  int mass_bits;
  int round_bits;  // rb
};

__device__ __forceinline__ uint32_t phase_parity(uint64_t x, const uint64_t* lut, int chunks) {
  uint64_t marked = x | (1ull << 63);
  uint64_t acc = 0;
  for (int c = 0; c < chunks; ++c) acc ^= lut[(c << 8) | ((x >> (8 * c)) & 255u)] & marked;
  return (uint32_t)__popcll(acc) & 1u;
}

__device__ __forceinline__ float2 dphase_of(uint64_t x, const float2* lut, int chunks) {
  float2 v = lut[x & 255u];
  for (int c = 1; c < chunks; ++c) v = cmul(v, lut[(c << 8) | ((x >> (8 * c)) & 255u)]);
  return v;
}

__device__ __forceinline__ float2 init_value(uint64_t x, const PassArgs& a, const uint64_t* lut,
                                             const float2* w) {
  if (a.init_kind == RGS_INIT_GRAPH) {
    uint32_t par = phase_parity(x, lut, a.init_chunks);
    return cneg_if(make_float2(a.init_scale, 0.f), par);
  }
  if (a.init_kind == RGS_INIT_COV) {
    uint64_t marked = x | (1ull << 63);
    uint64_t b = 0, acc = 0;
    for (int c = 0; c < a.init_chunks; ++c) {
      uint32_t e = (c << 8) | ((x >> (8 * c)) & 255u);
      b ^= lut[2 * e];
      acc ^= lut[2 * e + 1] & marked;
    }
    float2 val = w[b & 255u];
    for (int c = 1; c < a.init_chunks; ++c) val = cmul(val, w[(c << 8) | ((b >> (8 * c)) & 255u)]);
    return cneg_if(val, (uint32_t)__popcll(acc) & 1u);
  }
  // RGS_INIT_DELTA
  return x == 0 ? make_float2(a.init_scale, 0.f) : make_float2(0.f, 0.f);
}

template <int TB, int RB>
struct Shape {
  static constexpr int R = 1 << RB;
  static constexpr int NT = 1 << (TB - RB);
};

// The layout sequence s_0 = max(TB-RB, C), s_{k+1} = max(s_k - RB, C) as
// constant expressions; set k owns the local bits [at(k), hi(k)).
template <int TB, int RB, int C>
struct Seq {
  static constexpr int first() { return TB - RB > C ? TB - RB : C; }
  static constexpr int next(int s) { return s - RB > C ? s - RB : C; }
  static constexpr int at(int k) {
    int s = first();
    for (int i = 0; i < k; ++i) s = next(s);
    return s;
  }
  static constexpr int len() {
    int n = 1;
    for (int s = first(); s > C; s = next(s)) ++n;
    return n;
  }
  static constexpr int hi(int k) { return k == 0 ? TB : at(k - 1); }
};

template <int RB, int B>
__device__ __forceinline__ void wht_reg(float2 (&v)[1 << RB]) {
#pragma unroll
  for (int j = 0; j < (1 << RB); ++j) {
    if (!(j & (1 << B))) {
      float2 x = v[j], y = v[j | (1 << B)];
      v[j] = cadd(x, y);
      v[j | (1 << B)] = csub(x, y);
    }
  }
}

// Real rotation [[c, -s], [s, c]] on register bit B.
template <int RB, int B>
__device__ __forceinline__ void rot_reg(float2 (&v)[1 << RB], float2 cs) {
#pragma unroll
  for (int j = 0; j < (1 << RB); ++j) {
    if (!(j & (1 << B))) {
      float2 x = v[j], y = v[j | (1 << B)];
      v[j] = make_float2(cs.x * x.x - cs.y * y.x, cs.x * x.y - cs.y * y.y);
      v[j | (1 << B)] = make_float2(cs.y * x.x + cs.x * y.x, cs.y * x.y + cs.x * y.y);
    }
  }
}

template <int RB>
__device__ __forceinline__ void wht_set(float2 (&v)[1 << RB], int s, int lo, int hi) {
#define RGS_W(B) \
  if (B < RB && s + B >= lo && s + B < hi) wht_reg<RB, (B < RB ? B : 0)>(v);
  RGS_W(0) RGS_W(1) RGS_W(2) RGS_W(3) RGS_W(4) RGS_W(5)
#undef RGS_W
}

template <int RB>
__device__ __forceinline__ void rot_set(float2 (&v)[1 << RB], int s, int lo, int hi, int c, const PassArgs& a) {
#define RGS_R(B)                                                      \
  if (B < RB && s + B >= lo && s + B < hi) {                          \
    if (a.qstart + (s + B - c) < a.gate_limit)                        \
      rot_reg<RB, (B < RB ? B : 0)>(v, a.rot[s + B - c]);             \
  }
  RGS_R(0) RGS_R(1) RGS_R(2) RGS_R(3) RGS_R(4) RGS_R(5)
#undef RGS_R
}

// Thread coordinates.  With LDS rounds the round bits (the top rb lane bits,
// local bits [c-rb, c)) are the TOP thread-index bits `tr`, uniform across a
// wavefront, and `tl` enumerates the remaining non-register bits in order.
// "Compressed" local indices drop the round bits; LDS holds one round's
// compressed indices.
struct Thr {
  uint32_t tl;
  uint32_t tr;
};

// Compressed local index of register 0 in layout s (window at s - rb).
template <int RB>
__device__ __forceinline__ uint32_t comp_base(uint32_t tl, int s, int rb) {
  const int sc = s - rb;
  return (tl & ((1u << sc) - 1u)) | ((tl >> sc) << (sc + RB));
}

// Full local index of register 0 in layout s.
template <int RB>
__device__ __forceinline__ uint32_t thread_local_base(Thr th, int s, int c, int rb) {
  const uint32_t x = comp_base<RB>(th.tl, s, rb);
  const int lo = c - rb;
  return (x & ((1u << lo) - 1u)) | (th.tr << lo) | ((x >> lo) << c);
}

// Register-part offset of register j in layout s: J + J/16, J = j << (s - rb).
__device__ __forceinline__ uint32_t reg_slot(uint32_t j, int sh) {
  uint32_t J = j << sh;
  return J + (J >> 4);
}

template <int TB, int RB>
__device__ __forceinline__ void exchange(float2 (&v)[1 << RB], float2* lds, Thr th, int s_from, int s_to,
                                         int c, int rb) {
  constexpr int R = 1 << RB;
  const uint32_t wc = comp_base<RB>(th.tl, s_from, rb), rc = comp_base<RB>(th.tl, s_to, rb);
  const uint32_t wbase = wc + (wc >> 4);
  const uint32_t rbase = rc + (rc >> 4);
  const int wsh = s_from - rb, rsh = s_to - rb;
  const int rounds = 1 << rb;
  for (int r = 0; r < rounds; ++r) {
    if (th.tr == (uint32_t)r) {
      if (wsh >= 4) {
        const uint32_t K = (1u << wsh) + (1u << (wsh - 4));
#pragma unroll
        for (int j = 0; j < R; ++j) lds[wbase + j * K] = v[j];
      } else {
#pragma unroll
        for (int j = 0; j < R; ++j) lds[wbase + reg_slot(j, wsh)] = v[j];
      }
    }
    __syncthreads();
    if (th.tr == (uint32_t)r) {
      if (rsh >= 4) {
        const uint32_t K = (1u << rsh) + (1u << (rsh - 4));
#pragma unroll
        for (int j = 0; j < R; ++j) v[j] = lds[rbase + j * K];
      } else {
#pragma unroll
        for (int j = 0; j < R; ++j) v[j] = lds[rbase + reg_slot(j, rsh)];
      }
    }
    __syncthreads();
  }
}

template <int TB, int RB, int CFIX = -1, int RBFIX = -1>
__global__ void __launch_bounds__(1 << (TB - RB)) k_pass(PassArgs a) {
  constexpr int R = Shape<TB, RB>::R;
  constexpr int NT = Shape<TB, RB>::NT;
  extern __shared__ __align__(16) unsigned char smem[];
  const int c = CFIX >= 0 ? CFIX : a.lane_bits;
  const int rb = RBFIX >= 0 ? RBFIX : a.round_bits;
  const uint32_t tile_slots = (1u << (TB - rb)) + (1u << (TB - rb - 4));
  float2* lds = reinterpret_cast<float2*>(smem);
  uint64_t* lut_phase = reinterpret_cast<uint64_t*>(smem + (size_t)tile_slots * sizeof(float2));
  float2* lut_d = reinterpret_cast<float2*>(lut_phase + (size_t)a.phase_chunks * 256);
  uint64_t* lut_init = reinterpret_cast<uint64_t*>(lut_d + (size_t)a.dphase_chunks * 256);
  float2* lut_w =
      reinterpret_cast<float2*>(lut_init + (size_t)a.init_chunks * 256 * (a.init_kind == RGS_INIT_COV ? 2 : 1));
  const uint32_t tid = threadIdx.x;
  Thr th;
  {
    const int tlb = (TB - RB) - rb;  // >= 6 whenever rb > 0, so tr is wave-uniform
    th.tl = tid & ((1u << tlb) - 1u);
    th.tr = rb ? __builtin_amdgcn_readfirstlane(tid >> tlb) : 0u;
  }

  const bool gen = !(a.flags & F_LOAD);
  bool staged = false;
  if (a.phase_chunks) {
    for (int i = tid; i < a.phase_chunks * 256; i += NT) lut_phase[i] = a.phase_lut[i];
    staged = true;
  }
  if (a.dphase_chunks) {
    for (int i = tid; i < a.dphase_chunks * 256; i += NT) lut_d[i] = a.dphase[i];
    staged = true;
  }
  if (gen && a.init_chunks) {
    int words = a.init_chunks * 256 * (a.init_kind == RGS_INIT_COV ? 2 : 1);
    for (int i = tid; i < words; i += NT) lut_init[i] = a.init_lut[i];
    if (a.init_kind == RGS_INIT_COV)
      for (int i = tid; i < a.init_chunks * 256; i += NT) lut_w[i] = a.init_w[i];
    staged = true;
  }
  if (staged) __syncthreads();

  const int mid = a.qstart - c;
  const uint64_t blk = (uint64_t)blockIdx.x + ((uint64_t)blockIdx.y << 24);
  const uint64_t base = ((blk & ((1ull << mid) - 1ull)) << c) | ((blk >> mid) << (a.qstart + TB - c));
  const uint32_t cmask = (1u << c) - 1u;
  // global index of a local index with no register bits set
  auto gidx = [&](uint32_t l) -> uint64_t {
    return base | (uint64_t)(l & cmask) | ((uint64_t)(l >> c) << a.qstart);
  };

  float2 v[R];
  const int S1 = max(TB - RB, c);
  const uint64_t g0 = gidx(thread_local_base<RB>(th, S1, c, rb));
  const uint32_t gsh = a.qstart + S1 - c;   // global bit of register bit 0 in layout S1
  if (!gen) {
#pragma unroll
    for (int j = 0; j < R; ++j) v[j] = a.state[g0 + ((uint64_t)j << gsh)];
  } else {
#pragma unroll
    for (int j = 0; j < R; ++j) v[j] = init_value(g0 + ((uint64_t)j << gsh), a, lut_init, lut_w);
  }

  auto apply_phase = [&](int sl) {
    const uint64_t gb = gidx(thread_local_base<RB>(th, sl, c, rb));
    const uint32_t sh = a.qstart + sl - c;
#pragma unroll
    for (int j = 0; j < R; ++j) {
      const uint64_t x = gb + ((uint64_t)j << sh);
      float2 val = v[j];
      if (a.phase_chunks) val = cneg_if(val, phase_parity(x, lut_phase, a.phase_chunks));
      if (a.dphase_chunks) val = cmul(val, dphase_of(x, lut_d, a.dphase_chunks));
      v[j] = val;
    }
  };

  int s = S1;   // current layout
  int hi = TB;  // upper end of the bits owned by the current set
  if (a.flags & F_WHT) {
    while (true) {
      wht_set<RB>(v, s, s, hi);
      if (s <= c) break;
      int nx = max(s - RB, c);
      exchange<TB, RB>(v, lds, th, s, nx, c, rb);
      hi = s;
      s = nx;
    }
  }
  if (a.flags & F_PHASE) apply_phase(s);
  if (a.flags & F_GATES) {
    if (s != S1) {
      // Walk back up through the parents: the parent of the set at s starts at hi.
      while (true) {
        rot_set<RB>(v, s, s, hi, c, a);
        if (s == S1) break;
        int up = hi;
        int up_hi = min(hi + RB, TB);
        exchange<TB, RB>(v, lds, th, s, up, c, rb);
        s = up;
        hi = up_hi;
      }
    } else {
      while (true) {
        rot_set<RB>(v, s, s, hi, c, a);
        if (s <= c) break;
        int nx = max(s - RB, c);
        exchange<TB, RB>(v, lds, th, s, nx, c, rb);
        hi = s;
        s = nx;
      }
    }
  }
  if (RB == 4 && (a.flags & F_MASS)) {
    // low tile only (c = 0, 16 registers): registers must hold local bits 0..3
    if (s != 0) {
      exchange<TB, RB>(v, lds, th, s, 0, c, rb);
      s = 0;
    }
    double p[R];
#pragma unroll
    for (int j = 0; j < R; ++j) p[j] = prob64(v[j]);
#pragma unroll
    for (int w = 1; w < R; w <<= 1) {
#pragma unroll
      for (int j = 0; j < R; j += 2 * w) p[j] = __dadd_rn(p[j], p[j + w]);
    }
    double m = p[0];
    const int sb = a.mass_bits - RB;
    for (int b = 0; b < sb; ++b) m = __dadd_rn(m, __shfl_xor(m, 1 << b));
    if ((tid & ((1u << sb) - 1u)) == 0) a.mass[(base >> a.mass_bits) + (tid >> sb)] = m;
  }
  if (s != S1) exchange<TB, RB>(v, lds, th, s, S1, c, rb);
#pragma unroll
  for (int j = 0; j < R; ++j) a.state[g0 + ((uint64_t)j << gsh)] = v[j];
}

// ------------------------------------------------------------------ host side
struct PassPlan {
  int flags;
  int low;     // 1: contiguous tile
  int qstart;
  int count;   // target bits
  int tb;      // tile bits
  int rbits;   // log2 registers per thread
  int c;       // lane bits
  int rb;      // log2 LDS rounds
};

static int chunks_of(int bits) { return (bits + 7) / 8; }

static size_t lds_bytes(const PassPlan& p, int nbits, const rgs_slab_desc* d) {
  size_t slots = ((size_t)1 << (p.tb - p.rb)) + ((size_t)1 << (p.tb - p.rb - 4));
  size_t bytes = slots * 8;
  if (p.flags & F_PHASE) {
    if (d->phase_lut) bytes += (size_t)chunks_of(nbits) * 256 * 8;
    if (d->dphase_lut) bytes += (size_t)chunks_of(nbits) * 256 * 8;
  }
  if (!(p.flags & F_LOAD)) {
    if (d->init_kind == RGS_INIT_GRAPH) bytes += (size_t)chunks_of(nbits) * 256 * 8;
    if (d->init_kind == RGS_INIT_COV) bytes += (size_t)chunks_of(nbits) * 256 * 24;
  }
  return bytes;
}

static int plan_passes(const rgs_slab_desc* d, std::vector<PassPlan>& out, std::string& err) {
  const int n = d->nbits;
  if (n < 6 || n > 40) { err = "nbits must be in 6..40"; return kInvalid; }
  int T0 = std::min(d->low_bits, n);
  if (T0 < 6 || T0 > 14) { err = "low_bits must be in 6..14"; return kInvalid; }
  const int TS = d->tile_bits;
  const int L = d->lane_bits;
  std::vector<PassPlan> groups;
  if (n > T0) {
    if (TS < 6 || TS > 14) { err = "tile_bits must be in 6..14"; return kInvalid; }
    const int RBs = 4;
    if (L < 1 || TS - L < RBs) { err = "lane_bits must leave at least 4 target bits"; return kInvalid; }
    if (n - T0 < RBs) T0 = n - RBs;  // a strided group needs a full register window
    if (T0 < 6) { err = "state too small for a strided pass; raise low_bits"; return kInvalid; }
    const int F = TS - L;
    const int rem = n - T0;
    const int ng = (rem + F - 1) / F;
    int q = T0;
    for (int i = 0; i < ng; ++i) {
      int cnt = rem / ng + (i < rem % ng ? 1 : 0);
      PassPlan g{};
      g.low = 0;
      g.qstart = q;
      g.count = cnt;
      g.c = L;
      g.tb = L + cnt;
      g.rbits = 4;
      if (cnt < g.rbits) { err = "strided group narrower than a register window"; return kInvalid; }
      if (g.c > q) { err = "lane bits overlap the target bits; lower lane_bits"; return kInvalid; }
      g.rb = 0;
      if (g.rb > g.c) { err = "too few lane bits for the LDS rounds"; return kInvalid; }
      groups.push_back(g);
      q += cnt;
    }
  }
  PassPlan low{};
  low.low = 1;
  low.qstart = 0;
  low.count = T0;
  low.tb = T0;
  low.rbits = 4;
  low.c = 0;
  low.rb = 0;
  const bool walsh = d->do_walsh != 0;
  const bool phase = d->phase_lut != nullptr || d->dphase_lut != nullptr;
  const bool gates = d->rot != nullptr;
  const bool mass = d->mass != nullptr;
  const int src = d->init_kind == RGS_INIT_LOAD ? F_LOAD : 0;
  const int k = (int)groups.size();
  auto with = [](PassPlan p, int f) { p.flags = f; return p; };
  out.clear();
  if (k == 0) {
    out.push_back(with(low, src | (walsh ? F_WHT : 0) | (phase ? F_PHASE : 0) | (gates ? F_GATES : 0) | (mass ? F_MASS : 0)));
  } else if (walsh) {
    out.push_back(with(low, src | F_WHT));
    for (int i = 0; i + 1 < k; ++i) out.push_back(with(groups[i], F_LOAD | F_WHT));
    out.push_back(with(groups[k - 1], F_LOAD | F_WHT | (phase ? F_PHASE : 0) | (gates ? F_GATES : 0)));
    if (gates)
      for (int i = k - 2; i >= 0; --i) out.push_back(with(groups[i], F_LOAD | F_GATES));
    if (gates || mass) out.push_back(with(low, F_LOAD | (gates ? F_GATES : 0) | (mass ? F_MASS : 0)));
  } else {
    out.push_back(with(groups[0], src | (phase ? F_PHASE : 0) | (gates ? F_GATES : 0)));
    if (gates)
      for (int i = 1; i < k; ++i) out.push_back(with(groups[i], F_LOAD | F_GATES));
    if (gates || mass) out.push_back(with(low, F_LOAD | (gates ? F_GATES : 0) | (mass ? F_MASS : 0)));
  }
  std::vector<PassPlan> kept;
  for (auto& p : out)
    if ((p.flags & ~F_LOAD) || !(p.flags & F_LOAD)) kept.push_back(p);
  out.swap(kept);
  for (auto& p : out) {
    if (lds_bytes(p, n, d) > kLdsBudget) { err = "tile plus tables exceed 160 KiB of LDS; lower the tile bits"; return kInvalid; }
    if ((p.flags & F_MASS) && (d->mass_bits < p.rbits || d->mass_bits > std::min(p.tb, p.rbits + 6))) {
      err = "mass_bits must be in 4..min(low tile bits, 10)";
      return kInvalid;
    }
  }
  return kOk;
}

template <int TB, int RB, int CFIX = -1, int RBFIX = -1>
static hipError_t launch_tb(const PassArgs& a, size_t shmem, hipStream_t st) {
  uint64_t tiles = 1ull << (a.nbits - TB);
  dim3 grid((unsigned)std::min<uint64_t>(tiles, 1ull << 24), (unsigned)std::max<uint64_t>(1, tiles >> 24));
  static bool attr_set = false;
  if (!attr_set) {
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&k_pass<TB, RB, CFIX, RBFIX>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)kLdsBudget);
    attr_set = true;
  }
  hipLaunchKernelGGL((k_pass<TB, RB, CFIX, RBFIX>), grid, dim3(Shape<TB, RB>::NT), shmem, st, a);
  return hipGetLastError();
}

static hipError_t launch_pass(const PassArgs& a, int tb, int rbits, size_t shmem, hipStream_t st) {
  if (rbits != 4) return hipErrorInvalidValue;
  switch (tb) {
    case 6: return launch_tb<6, 4>(a, shmem, st);
    case 7: return launch_tb<7, 4>(a, shmem, st);
    case 8: return launch_tb<8, 4>(a, shmem, st);
    case 9: return launch_tb<9, 4>(a, shmem, st);
    case 10: return launch_tb<10, 4>(a, shmem, st);
    case 11: return launch_tb<11, 4>(a, shmem, st);
    case 12: return launch_tb<12, 4>(a, shmem, st);
    case 13: return launch_tb<13, 4>(a, shmem, st);
    case 14: return launch_tb<14, 4>(a, shmem, st);
    default: return hipErrorInvalidValue;
  }
}

static int run_passes(const rgs_slab_desc* d, int first, int last, void* stream) {
  if (!d || !d->state) return fail(kInvalid, "null descriptor or state");
  std::vector<PassPlan> plan;
  std::string err;
  int rc = plan_passes(d, plan, err);
  if (rc) return fail(rc, err);
  if (first >= (int)plan.size()) return fail(kInvalid, "pass index beyond the schedule");
  last = std::min(last, (int)plan.size() - 1);
  const int n = d->nbits;
  if (d->init_kind < 0 || d->init_kind > 3) return fail(kInvalid, "bad init_kind");
  if ((d->init_kind == RGS_INIT_GRAPH || d->init_kind == RGS_INIT_COV) && !d->init_lut)
    return fail(kInvalid, "init tables missing");
  if (d->init_kind == RGS_INIT_COV && !d->init_wlut) return fail(kInvalid, "covariant weights missing");
  if ((d->phase_lut || d->dphase_lut || d->init_kind == RGS_INIT_GRAPH || d->init_kind == RGS_INIT_COV) && n > 63)
    return fail(kInvalid, "lookup tables need nbits <= 63");
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  for (int pi = first; pi <= last; ++pi) {
    const PassPlan& p = plan[pi];
    PassArgs a{};
    a.state = reinterpret_cast<float2*>(d->state);
    a.init_lut = d->init_lut;
    a.init_w = reinterpret_cast<const float2*>(d->init_wlut);
    a.phase_lut = d->phase_lut;
    a.dphase = reinterpret_cast<const float2*>(d->dphase_lut);
    // rotations of this pass's target bits travel in the kernel arguments
    const int t0 = p.low ? 0 : p.qstart;
    for (int i = 0; i < 16; ++i) {
      const int q = t0 + i;
      a.rot[i] = (d->rot && i < p.count && q < n) ? make_float2(d->rot[2 * q], d->rot[2 * q + 1])
                                                  : make_float2(1.f, 0.f);
    }
    a.mass = d->mass;
    a.init_scale = d->init_scale;
    a.nbits = n;
    a.lane_bits = p.c;
    a.qstart = p.low ? 0 : p.qstart;
    a.flags = p.flags;
    a.init_kind = d->init_kind;
    const bool gen = !(p.flags & F_LOAD);
    a.init_chunks = (gen && (d->init_kind == RGS_INIT_GRAPH || d->init_kind == RGS_INIT_COV)) ? chunks_of(n) : 0;
    a.phase_chunks = ((p.flags & F_PHASE) && d->phase_lut) ? chunks_of(n) : 0;
    a.dphase_chunks = ((p.flags & F_PHASE) && d->dphase_lut) ? chunks_of(n) : 0;
    a.gate_limit = d->rot ? std::min(d->gate_limit, n) : 0;
    a.mass_bits = d->mass_bits;
    a.round_bits = p.rb;
    hipError_t e = launch_pass(a, p.tb, p.rbits, lds_bytes(p, n, d), st);
    if (e != hipSuccess) return fail(kHip, std::string("k_pass launch: ") + hipGetErrorString(e));
  }
  return kOk;
}

}  // namespace rgs

using namespace rgs;

extern "C" int rgs_slab_describe(const rgs_slab_desc* d, int* passes, char* text, int text_len) {
  if (!d) return fail(kInvalid, "null descriptor");
  std::vector<PassPlan> plan;
  std::string err;
  int rc = plan_passes(d, plan, err);
  if (rc) return fail(rc, err);
  if (passes) *passes = (int)plan.size();
  if (text && text_len > 0) {
    std::string s;
    for (auto& p : plan) {
      char buf[160];
      std::snprintf(buf, sizeof buf, "%s%s%s%s%s%s:%d+%d/tb%d/c%d/r%d ", p.low ? "low" : "strided",
                    (p.flags & F_LOAD) ? "" : "[init]", (p.flags & F_WHT) ? "[wht]" : "",
                    (p.flags & F_PHASE) ? "[phase]" : "", (p.flags & F_GATES) ? "[gates]" : "",
                    (p.flags & F_MASS) ? "[mass]" : "", p.qstart, p.count, p.tb, p.c, 1 << p.rbits);
      s += buf;
    }
    std::snprintf(text, text_len, "%s", s.c_str());
  }
  return kOk;
}

extern "C" int rgs_slab_pipeline(const rgs_slab_desc* d, void* stream) { return run_passes(d, 0, 1 << 30, stream); }

extern "C" int rgs_slab_pass(const rgs_slab_desc* d, int pass, void* stream) {
  if (pass < 0) return fail(kInvalid, "negative pass index");
  return run_passes(d, pass, pass, stream);
}
