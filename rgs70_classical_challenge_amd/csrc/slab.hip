// Slab pipeline on one 2^nbits complex64 state: fused init / Walsh / phase /
// 2x2 gates / FP64 masses, plus the mass pyramid and the device draw.
//
// Tiles.  A pass works on tiles of 2^TB amplitudes, one workgroup of
// NT = 2^(TB-4) threads per tile, 16 amplitudes per thread in registers.
// Tile-local bit L maps to a global bit: for the contiguous ("low") pass the
// map is the identity; for a strided pass the first `c` local bits are the
// contiguous lane bits 0..c-1 and local bits c..TB-1 are the target bits
// qstart..qstart+TB-c-1.  The workgroup index supplies the remaining bits.
//
// Layouts.  In layout s the 16 registers of a thread hold local bits s..s+3
// and the thread index enumerates the other bits in increasing order, so a
// wavefront always spans the lowest local bits outside the register window
// (coalesced global access in layout TB-4).  Butterflies act on register
// bits; a layout change is one padded LDS round trip.  Layouts visited:
// TB-4, TB-8, ..., down to the lowest target bit (clamped at 0); set k owns
// the bits [max(s_k, tb0), s_{k-1}) with s_0 = TB, so every bit is touched
// exactly once per operation.
#include "rgs_common.h"
#include "rgs_amd.h"

#include <algorithm>
#include <cstdio>
#include <vector>

namespace rgs {

enum : int { F_LOAD = 1, F_WHT = 2, F_PHASE = 4, F_GATES = 8, F_MASS = 16 };

struct PassArgs {
  float2* state;
  const uint64_t* init_lut;   // GRAPH: phase words [ic][256]; COV: map+phase [ic][256][2]
  const float2* init_w;       // COV: [ic][256]
  const uint64_t* phase_lut;  // [pc][256]
  const float2* gates;        // [nbits][4]
  double* mass;
  float init_scale;
  int nbits;
  int lane_bits;   // c
  int qstart;      // global bit of local bit c (0 for the low tile)
  int flags;
  int init_kind;
  int init_chunks;
  int phase_chunks;
  int gate_limit;
  int mass_bits;
};

// Padded LDS index: one float2 of padding per 16 keeps every layout change
// free of bank conflicts for 8-byte accesses.
__device__ __forceinline__ uint32_t lds_pad(uint32_t l) { return l + (l >> 4); }

// Local index of register j of thread tid in layout s (16 registers).
__device__ __forceinline__ uint32_t local_of(uint32_t tid, uint32_t j, int s) {
  uint32_t lowmask = (1u << s) - 1u;
  return (tid & lowmask) | (j << s) | ((tid >> s) << (s + 4));
}

__device__ __forceinline__ uint32_t phase_parity(uint64_t x, const uint64_t* lut, int chunks) {
  uint64_t marked = x | (1ull << 63);
  uint64_t acc = 0;
  for (int c = 0; c < chunks; ++c) acc ^= lut[(c << 8) | ((x >> (8 * c)) & 255u)] & marked;
  return (uint32_t)__popcll(acc) & 1u;
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

template <int TB>
struct Tile {
  static constexpr int NT = 1 << (TB - 4);
  static constexpr int SIZE = 1 << TB;
  static constexpr int PADDED = SIZE + (SIZE >> 4);
};

// Apply the Walsh butterfly on register bit r for all pairs.
template <int R>
__device__ __forceinline__ void wht_reg(float2 (&v)[16]) {
#pragma unroll
  for (int j = 0; j < 16; ++j) {
    if (!(j & (1 << R))) {
      float2 x = v[j], y = v[j | (1 << R)];
      v[j] = cadd(x, y);
      v[j | (1 << R)] = csub(x, y);
    }
  }
}

template <int R>
__device__ __forceinline__ void gate_reg(float2 (&v)[16], const float2* g) {
  const float2 u00 = g[0], u01 = g[1], u10 = g[2], u11 = g[3];
#pragma unroll
  for (int j = 0; j < 16; ++j) {
    if (!(j & (1 << R))) {
      float2 x = v[j], y = v[j | (1 << R)];
      v[j] = cmac2(u00, x, u01, y);
      v[j | (1 << R)] = cmac2(u10, x, u11, y);
    }
  }
}

// Ops on the register bits of layout s owning local bits [lo, hi).
__device__ __forceinline__ void wht_set(float2 (&v)[16], int s, int lo, int hi) {
  if (s + 0 >= lo && s + 0 < hi) wht_reg<0>(v);
  if (s + 1 >= lo && s + 1 < hi) wht_reg<1>(v);
  if (s + 2 >= lo && s + 2 < hi) wht_reg<2>(v);
  if (s + 3 >= lo && s + 3 < hi) wht_reg<3>(v);
}

__device__ __forceinline__ int qubit_of(int L, const PassArgs& a) {
  return L < a.lane_bits ? L : a.qstart + (L - a.lane_bits);
}

__device__ __forceinline__ void gate_set(float2 (&v)[16], int s, int lo, int hi, const PassArgs& a) {
#define RGS_GATE_BIT(R)                                              \
  if (s + R >= lo && s + R < hi) {                                   \
    int q = qubit_of(s + R, a);                                      \
    if (q < a.gate_limit) gate_reg<R>(v, a.gates + 4 * q);           \
  }
  RGS_GATE_BIT(0)
  RGS_GATE_BIT(1)
  RGS_GATE_BIT(2)
  RGS_GATE_BIT(3)
#undef RGS_GATE_BIT
}

template <int TB>
__device__ __forceinline__ void exchange(float2 (&v)[16], float2* lds, uint32_t tid, int s_from, int s_to) {
#pragma unroll
  for (int j = 0; j < 16; ++j) lds[lds_pad(local_of(tid, j, s_from))] = v[j];
  __syncthreads();
#pragma unroll
  for (int j = 0; j < 16; ++j) v[j] = lds[lds_pad(local_of(tid, j, s_to))];
  __syncthreads();
}

template <int TB>
__global__ void __launch_bounds__(Tile<TB>::NT) k_tile(PassArgs a) {
  constexpr int PADDED = Tile<TB>::PADDED;
  extern __shared__ __align__(16) unsigned char smem[];
  float2* lds = reinterpret_cast<float2*>(smem);
  uint64_t* lut_phase = reinterpret_cast<uint64_t*>(smem + (size_t)PADDED * sizeof(float2));
  uint64_t* lut_init = lut_phase + (size_t)a.phase_chunks * 256;
  float2* lut_w = reinterpret_cast<float2*>(lut_init + (size_t)a.init_chunks * 256 * (a.init_kind == RGS_INIT_COV ? 2 : 1));
  const uint32_t tid = threadIdx.x;

  // Stage the lookup tables in LDS.
  const bool gen = !(a.flags & F_LOAD);
  if (a.flags & F_PHASE)
    for (int i = tid; i < a.phase_chunks * 256; i += Tile<TB>::NT) lut_phase[i] = a.phase_lut[i];
  if (gen && (a.init_kind == RGS_INIT_GRAPH || a.init_kind == RGS_INIT_COV)) {
    int words = a.init_chunks * 256 * (a.init_kind == RGS_INIT_COV ? 2 : 1);
    for (int i = tid; i < words; i += Tile<TB>::NT) lut_init[i] = a.init_lut[i];
    if (a.init_kind == RGS_INIT_COV)
      for (int i = tid; i < a.init_chunks * 256; i += Tile<TB>::NT) lut_w[i] = a.init_w[i];
  }
  if ((a.flags & F_PHASE) || gen) __syncthreads();

  const int c = a.lane_bits;
  const int tb0 = c;  // lowest target bit (0 for the low tile)
  const int mid = a.qstart - c;
  const uint64_t blk = (uint64_t)blockIdx.x + ((uint64_t)blockIdx.y << 24);
  const uint64_t base = ((blk & ((1ull << mid) - 1ull)) << c) | ((blk >> mid) << (a.qstart + TB - c));
  const uint32_t cmask = (1u << c) - 1u;
  auto gidx = [&](uint32_t l) -> uint64_t {
    return base | (uint64_t)(l & cmask) | ((uint64_t)(l >> c) << a.qstart);
  };

  float2 v[16];
  const int S1 = TB - 4;
  if (!gen) {
#pragma unroll
    for (int j = 0; j < 16; ++j) v[j] = a.state[gidx(local_of(tid, j, S1))];
  } else {
#pragma unroll
    for (int j = 0; j < 16; ++j) v[j] = init_value(gidx(local_of(tid, j, S1)), a, lut_init, lut_w);
  }

  int s = S1;           // current layout
  int hi = TB;          // upper end of the current set's bit range
  // set sequence helpers: next(s) = max(s-4, 0), stop once s <= tb0
  if (a.flags & F_WHT) {
    while (true) {
      wht_set(v, s, std::max(s, tb0), hi);
      if (s <= tb0) break;
      int nx = std::max(s - 4, 0);
      exchange<TB>(v, lds, tid, s, nx);
      hi = s;
      s = nx;
    }
  }
  if (a.flags & F_PHASE) {
#pragma unroll
    for (int j = 0; j < 16; ++j) {
      uint32_t par = phase_parity(gidx(local_of(tid, j, s)), lut_phase, a.phase_chunks);
      v[j] = cneg_if(v[j], par);
    }
  }
  if (a.flags & F_GATES) {
    if (s != S1) {
      // Walk back up: the set owning [max(s,tb0), hi) first, then its parents.
      while (true) {
        gate_set(v, s, std::max(s, tb0), hi, a);
        if (s == S1) break;
        int up = hi;             // the parent set starts at `hi` ...
        int up_hi = std::min(hi + 4, TB);
        // ... and owns [hi, parent's upper end); parents are hi = s_{k-1}.
        exchange<TB>(v, lds, tid, s, up);
        s = up;
        hi = up_hi;
      }
    } else {
      while (true) {
        gate_set(v, s, std::max(s, tb0), hi, a);
        if (s <= tb0) break;
        int nx = std::max(s - 4, 0);
        exchange<TB>(v, lds, tid, s, nx);
        hi = s;
        s = nx;
      }
    }
  }
  if (a.flags & F_MASS) {
    if (s != 0) {
      exchange<TB>(v, lds, tid, s, 0);
      s = 0;
    }
    // registers hold local bits 0..3; lanes hold bits 4..9
    double p[16];
#pragma unroll
    for (int j = 0; j < 16; ++j) p[j] = prob64(v[j]);
#pragma unroll
    for (int w = 1; w < 16; w <<= 1) {
#pragma unroll
      for (int j = 0; j < 16; j += 2 * w) p[j] = __dadd_rn(p[j], p[j + w]);
    }
    double m = p[0];
    const int sb = a.mass_bits - 4;
    for (int b = 0; b < sb; ++b) {
      double o = __shfl_xor(m, 1 << b);
      m = __dadd_rn(m, o);
    }
    if ((tid & ((1u << sb) - 1u)) == 0) a.mass[(base >> a.mass_bits) + (tid >> sb)] = m;
  }
  if (s != S1) exchange<TB>(v, lds, tid, s, S1);
#pragma unroll
  for (int j = 0; j < 16; ++j) a.state[gidx(local_of(tid, j, S1))] = v[j];
}

// ------------------------------------------------------------------ pyramid
// One workgroup reduces 2^k consecutive entries of level L into levels L+1..L+k.
__global__ void __launch_bounds__(256) k_pyramid(double* pyr, uint64_t level_off, int level_bits, int k,
                                                 uint64_t out_off_first) {
  __shared__ double s[512];
  const uint32_t tid = threadIdx.x;
  const uint64_t group = blockIdx.x;
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

// ------------------------------------------------------------------ draw
// One wavefront per draw.  Lanes hold 4 consecutive amplitudes of the leaf
// block (mass_bits <= 8 -> 256 amplitudes).  Level l of the pyramid (0 = base
// with 2^B entries) starts at 2^(B+1) - 2^(B+1-l).
__device__ __forceinline__ uint64_t level_off(int B, int l) {
  return (1ull << (B + 1)) - (1ull << (B + 1 - l));
}

__global__ void __launch_bounds__(64) k_draw(const float2* state, int nbits, const double* pyr, int mass_bits,
                                             const double* uniforms, int count, int64_t* out_index,
                                             double* out_p) {
  const int d = blockIdx.x;
  if (d >= count) return;
  const int lane = threadIdx.x;
  const int B = nbits - mass_bits;
  double t = uniforms[d] * pyr[level_off(B, B)];
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
  // Leaf block: recompute the pairwise tree with the same association.
  const int bsize = 1 << mass_bits;
  const uint64_t bbase = idx << mass_bits;
  double p[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    int l = lane * 4 + j;
    p[j] = l < bsize ? prob64(state[bbase + l]) : 0.0;
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
    out_index[d] = (int64_t)(bbase + (uint64_t)(node * 4 + sub));
    out_p[d] = ps;
  }
}

__global__ void k_gather_prob(const float2* state, const int64_t* index, int count, double* out) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) out[i] = prob64(state[index[i]]);
}

// ------------------------------------------------------------------ host side
struct PassPlan {
  int flags;
  int low;        // 1: contiguous tile
  int qstart;
  int count;      // target bits
  int tb;         // tile bits
  int c;          // lane bits
};

static int plan_passes(const rgs_slab_desc* d, std::vector<PassPlan>& out, std::string& err) {
  const int n = d->nbits;
  if (n < 6 || n > 40) { err = "nbits must be in 6..40"; return kInvalid; }
  const int T0 = std::min(d->low_bits, n);
  if (T0 < 6 || T0 > 14) { err = "low_bits must be in 6..14"; return kInvalid; }
  const int TS = d->tile_bits;
  if (n > T0 && (TS < 6 || TS > 14)) { err = "tile_bits must be in 6..14"; return kInvalid; }
  if (n > T0 && (d->lane_bits < 1 || d->lane_bits >= TS)) { err = "lane_bits must be in 1..tile_bits-1"; return kInvalid; }
  // strided groups over [T0, n)
  std::vector<PassPlan> groups;
  int F = TS - d->lane_bits;
  for (int q = T0; q < n; q += F) {
    PassPlan g{};
    g.low = 0;
    g.qstart = q;
    g.count = std::min(F, n - q);
    g.tb = TS;
    g.c = TS - g.count;
    if (g.c > q) { err = "strided tile wider than the bits below it; lower tile_bits"; return kInvalid; }
    groups.push_back(g);
  }
  PassPlan low{};
  low.low = 1;
  low.qstart = 0;
  low.count = T0;
  low.tb = T0;
  low.c = 0;
  const bool walsh = d->do_walsh != 0;
  const bool phase = d->phase_lut != nullptr;
  const bool gates = d->gates != nullptr;
  const bool mass = d->mass != nullptr;
  const int src = d->init_kind == RGS_INIT_LOAD ? F_LOAD : 0;
  const int k = (int)groups.size();
  auto with = [](PassPlan p, int f) { p.flags = f; return p; };
  out.clear();
  if (k == 0) {
    out.push_back(with(low, src | (walsh ? F_WHT : 0) | (phase ? F_PHASE : 0) | (gates ? F_GATES : 0) | (mass ? F_MASS : 0)));
    return kOk;
  }
  if (walsh) {
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
  // drop passes that do nothing
  std::vector<PassPlan> kept;
  for (auto& p : out)
    if ((p.flags & ~F_LOAD) || !(p.flags & F_LOAD)) kept.push_back(p);
  out.swap(kept);
  return kOk;
}

static size_t lds_bytes(int tb, const PassArgs& a) {
  size_t tile = ((size_t)1 << tb) + (((size_t)1 << tb) >> 4);
  size_t bytes = tile * sizeof(float2);
  if (a.flags & F_PHASE) bytes += (size_t)a.phase_chunks * 256 * 8;
  if (!(a.flags & F_LOAD)) {
    if (a.init_kind == RGS_INIT_GRAPH) bytes += (size_t)a.init_chunks * 256 * 8;
    if (a.init_kind == RGS_INIT_COV) bytes += (size_t)a.init_chunks * 256 * 16 + (size_t)a.init_chunks * 256 * 8;
  }
  return bytes;
}

template <int TB>
static hipError_t launch_tb(const PassArgs& a, size_t shmem, hipStream_t st) {
  uint64_t tiles = 1ull << (a.nbits - TB);
  dim3 grid((unsigned)std::min<uint64_t>(tiles, 1ull << 24), (unsigned)std::max<uint64_t>(1, tiles >> 24));
  static bool attr_set = false;
  if (!attr_set) {
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&k_tile<TB>), hipFuncAttributeMaxDynamicSharedMemorySize,
                        160 * 1024);
    attr_set = true;
  }
  hipLaunchKernelGGL(k_tile<TB>, grid, dim3(Tile<TB>::NT), shmem, st, a);
  return hipGetLastError();
}

static hipError_t launch_pass(const PassArgs& a, int tb, size_t shmem, hipStream_t st) {
  switch (tb) {
    case 6: return launch_tb<6>(a, shmem, st);
    case 7: return launch_tb<7>(a, shmem, st);
    case 8: return launch_tb<8>(a, shmem, st);
    case 9: return launch_tb<9>(a, shmem, st);
    case 10: return launch_tb<10>(a, shmem, st);
    case 11: return launch_tb<11>(a, shmem, st);
    case 12: return launch_tb<12>(a, shmem, st);
    case 13: return launch_tb<13>(a, shmem, st);
    case 14: return launch_tb<14>(a, shmem, st);
    default: return hipErrorInvalidValue;
  }
}

static int chunks_of(int bits) { return (bits + 7) / 8; }

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
      char buf[128];
      std::snprintf(buf, sizeof buf, "%s%s%s%s%s%s:%d+%d/tb%d/c%d ", p.low ? "low" : "strided",
                    (p.flags & F_LOAD) ? "" : "[init]", (p.flags & F_WHT) ? "[wht]" : "",
                    (p.flags & F_PHASE) ? "[phase]" : "", (p.flags & F_GATES) ? "[gates]" : "",
                    (p.flags & F_MASS) ? "[mass]" : "", p.qstart, p.count, p.tb, p.c);
      s += buf;
    }
    std::snprintf(text, text_len, "%s", s.c_str());
  }
  return kOk;
}

extern "C" int rgs_slab_pipeline(const rgs_slab_desc* d, void* stream) {
  if (!d || !d->state) return fail(kInvalid, "null descriptor or state");
  std::vector<PassPlan> plan;
  std::string err;
  int rc = plan_passes(d, plan, err);
  if (rc) return fail(rc, err);
  const int n = d->nbits;
  if (d->init_kind < 0 || d->init_kind > 3) return fail(kInvalid, "bad init_kind");
  if ((d->init_kind == RGS_INIT_GRAPH || d->init_kind == RGS_INIT_COV) && !d->init_lut)
    return fail(kInvalid, "init tables missing");
  if (d->init_kind == RGS_INIT_COV && (!d->init_wlut || n > 63)) return fail(kInvalid, "covariant weights missing");
  if (d->phase_lut && n > 63) return fail(kInvalid, "phase tables need nbits <= 63");
  if (d->mass && (d->mass_bits < 4 || d->mass_bits > 10 || d->mass_bits > std::min(d->low_bits, n)))
    return fail(kInvalid, "mass_bits must be in 4..min(10, low_bits)");
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  for (auto& p : plan) {
    PassArgs a{};
    a.state = reinterpret_cast<float2*>(d->state);
    a.init_lut = d->init_lut;
    a.init_w = reinterpret_cast<const float2*>(d->init_wlut);
    a.phase_lut = d->phase_lut;
    a.gates = reinterpret_cast<const float2*>(d->gates);
    a.mass = d->mass;
    a.init_scale = d->init_scale;
    a.nbits = n;
    a.lane_bits = p.c;
    a.qstart = p.low ? 0 : p.qstart;
    a.flags = p.flags;
    a.init_kind = d->init_kind;
    a.init_chunks = chunks_of(n);
    a.phase_chunks = (p.flags & F_PHASE) ? chunks_of(n) : 0;
    if (p.flags & F_LOAD) a.init_chunks = 0;
    a.gate_limit = d->gates ? std::min(d->gate_limit, n) : 0;
    a.mass_bits = d->mass_bits;
    size_t shmem = lds_bytes(p.tb, a);
    if (shmem > 160 * 1024) return fail(kInvalid, "tile plus tables exceed 160 KiB of LDS; lower the tile bits");
    hipError_t e = launch_pass(a, p.tb, shmem, st);
    if (e != hipSuccess) return fail(kHip, std::string("k_tile launch: ") + hipGetErrorString(e));
  }
  return kOk;
}

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
                        int count, int64_t* out_index, double* out_p, void* stream) {
  if (!state || !pyr || !uniforms || !out_index || !out_p) return fail(kInvalid, "null pointer");
  if (mass_bits < 2 || mass_bits > 8 || nbits - mass_bits < 0 || nbits - mass_bits > 47)
    return fail(kInvalid, "draw needs 2 <= mass_bits <= 8");
  if (count <= 0) return kOk;
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  hipLaunchKernelGGL(k_draw, dim3(count), dim3(64), 0, st, reinterpret_cast<const float2*>(state), nbits, pyr,
                     mass_bits, uniforms, count, out_index, out_p);
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
