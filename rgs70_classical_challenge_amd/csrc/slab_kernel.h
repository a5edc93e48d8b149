// Tile kernel of the slab pipeline, specialised at compile time on the tile
// bits TB, the lane bits C and the LDS round bits RBW.
//
// Tiles.  A pass works on tiles of 2^TB amplitudes; a workgroup of 2^(TB-4)
// threads holds one tile, 16 amplitudes per thread in registers.  Tile-local
// bit L maps to a global bit: local bits 0..C-1 are the contiguous global bits
// 0..C-1 (rows of 2^C amplitudes; C = 0 for the contiguous "low" pass) and
// local bits C..TB-1 are the target bits qstart..qstart+TB-C-1.  The tile
// index supplies the remaining global bits.
//
// Layouts.  In layout s the registers hold local bits s..s+3 and the thread
// index enumerates the other local bits in increasing order, so the lane bits
// are the lowest thread bits and a wavefront touches whole rows when it loads
// or stores.  Butterflies act on register bits; a layout change goes through
// LDS (one padded float2 per 16).  The layouts visited are s_0 = max(TB-4, C),
// s_{k+1} = max(s_k - 4, C); layout k owns the local bits [s_k, s_{k-1})
// (s_{-1} = TB), so every bit is operated on exactly once.  With RBW > 0 each
// layout change runs in 2^RBW rounds, one per value of the top RBW lane bits
// (the top thread bits in every layout), so LDS holds 2^(TB-RBW) amplitudes.
//
// One workgroup per tile.  At 14 tile bits a workgroup fills a CU's LDS and
// its 4 waves per SIMD, so loads, butterflies and stores of a CU run one
// after another; at 12-13 tile bits two to four workgroups share a CU and
// overlap.  (A second register set to prefetch the next tile does not fit
// the 128-VGPR budget of 1024-thread workgroups.)
#pragma once
#include "rgs_common.h"
#include "rgs_amd.h"

#include <type_traits>

namespace rgs {

// F_INIT_UP: a contiguous pass that makes the state and only applies the
// Walsh butterflies builds its values in the lowest layout and walks up to
// the store layout (one LDS exchange fewer than walking down and back).
// F_NOSTORE: the contiguous mass pass writes the masses only (the state
// keeps the values it loaded; its tiles are recomputed on demand).
// F_LAYER: one walk down the layouts, each applying the Walsh butterflies,
// the diagonal phases (PassArgs::lph) and the rotations of the register bits
// it owns (a pass of a fuse_walsh schedule: per qubit Walsh -> phase ->
// rotation, and operations on different qubits commute).
enum : int {
  F_LOAD = 1, F_WHT = 2, F_PHASE = 4, F_GATES = 8, F_MASS = 16, F_INIT_UP = 32, F_NOSTORE = 64, F_LAYER = 128
};

constexpr size_t kLdsBudget = 160 * 1024;

// A quadratic sign form over GF(2) restricted to the local bits of a tile:
// up[L] holds bit L (the linear term of local bit L) and bits L' > L (the
// cross terms), in local coordinates; grow[L] is the form's global row of the
// global bit that local bit L maps to.
struct LocalForm {
  uint32_t up[16];
  uint64_t grow[16];
};

struct PassArgs {
  float2* state;
  const uint64_t* init_lut;   // GRAPH: phase words [ic][256]; COV: map+phase [ic][256][2]
  const float2* init_w;       // COV: [ic][256]
  const uint64_t* phase_lut;  // [pc][256] sign words, or null
  const float2* dphase;       // [dc][256] complex diagonal phase, or null
  double* mass;
  float tng[16];              // tangent t of target bit i's rotation M(t) (0 beyond gate_limit)
  // Separable sign evaluation of the init / phase_lut forms over the four
  // register bits (global bits p_0..p_3) of the layout where each is applied:
  // q[i] = symmetric form row of p_i, freg bit j = f(r_j) ^ f(0) for the
  // register offset r_j of register j.
  uint64_t iq[4];
  uint64_t pq[4];
  uint64_t icol[4];           // COV: map columns N e_{p_i} of the register bits
  // The graph-init and phase_lut sign forms over the tile's local bits
  // (split_form_bit).
  LocalForm iform;
  LocalForm pform;
  uint32_t ifreg;
  uint32_t pfreg;
  float init_scale;           // GRAPH / DELTA value, the rotations' scale included
  float wscale;               // COV: the rotations' scale, applied to the staged weights
  float lscale;               // LOAD: the rotations' scale, applied as the first pass loads (else 1)
  int nbits;
  int qstart;      // global bit of local bit C (0 for the low tile)
  int flags;
  int init_kind;
  int init_chunks;
  int phase_chunks;
  int dphase_chunks;
  int gate_limit;
  int mass_bits;
  int dphase_global;  // read dphase from global memory instead of staging it
  int wbits;          // COV weight chunk bits (8 or 9)
  int wchunks;        // COV weight chunks
  int wpop;           // COV popcount table: its row stride m + 1 (0: chunk products)
  int wentries;       // COV weight entries staged in LDS
  int wstride;        // COV popcount table as staged: row stride m - |o| + 1
  uint64_t womask;    // COV popcount table: the B outcome bits o
  uint64_t wnmask;    // COV popcount table: the other B bits (~o, m bits)
  uint64_t mflip;     // mass blocks: block b of the state is outcome block b ^ mflip
  // F_LAYER with F_PHASE: lph[K][m] = the diagonal phase of register offset m
  // over the bits that layout K owns (m a subset of them; lph[K][0] = 1)
  float2 lph[4][16];
};

__device__ __forceinline__ uint32_t phase_parity(uint64_t x, const uint64_t* lut, int chunks) {
  uint64_t acc = 0;
  for (int c = 0; c < chunks; ++c) acc ^= lut[(c << 8) | ((x >> (8 * c)) & 255u)];
  return (uint32_t)__popcll(acc & (x | (1ull << 63))) & 1u;
}

// A quadratic form over GF(2) satisfies, for x and y with no common set bits,
//   f(x ^ y) = f(x) ^ f(y) ^ f(0) ^ B(x, y),   B(x, y) = sum_{a in y} parity(x & row_a).
// Split a thread's index x0 = base ^ t, with base the tile's own bits (the
// same for every thread of the workgroup) and t the global image of the
// thread's local bits lb: f(base) is one table evaluation at a uniform
// address (fbase), the cross term is parity(lb & rb) with
// rb_L = parity(base & grow[L]) (uniform), and f(t) ^ f(0) =
// sum_{L in lb} parity(up[L] & lb) - the sign table is neither gathered per
// thread nor staged per tile.
template <int TB>
__device__ __forceinline__ uint32_t split_form_bit(uint32_t fbase, uint64_t base, uint32_t lb, const LocalForm& F) {
  uint32_t rb = 0;
#pragma unroll
  for (int L = 0; L < TB; ++L) rb |= (uint32_t)(__popcll(base & F.grow[L]) & 1) << L;
  uint32_t s = fbase ^ ((uint32_t)__popc(lb & rb) & 1u);
#pragma unroll
  for (int L = 0; L < TB; ++L) s ^= (lb >> L) & (uint32_t)__popc(lb & F.up[L]) & 1u;
  return s;
}

// Signs (-1)^{f(x0 ^ r_j)} of the 16 register values of a thread as a 16-bit
// mask, from the sign s0 = f(x0) of register 0: the register offsets add
// f(r_j) ^ f(0) (freg) and B(x0, r_j) = sum_i j_i parity(x0 & q[i]).
__device__ __forceinline__ uint32_t register_signs(uint32_t s0, uint64_t x0, const uint64_t* q, uint32_t freg) {
  uint32_t e = freg ^ (s0 ? 0xFFFFu : 0u);
  if (__popcll(x0 & q[0]) & 1) e ^= 0xAAAAu;
  if (__popcll(x0 & q[1]) & 1) e ^= 0xCCCCu;
  if (__popcll(x0 & q[2]) & 1) e ^= 0xF0F0u;
  if (__popcll(x0 & q[3]) & 1) e ^= 0xFF00u;
  return e;
}

// Address-space-qualified views of a table, so that a table read from LDS
// and the same table read from global memory compile to LDS and global
// instructions on their own code paths.
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef const f32x2 __attribute__((address_space(3)))* lds_f2;
typedef const f32x2 __attribute__((address_space(1)))* glb_f2;

__device__ __forceinline__ float2 tab_at(const float2* p, uint32_t i) { return p[i]; }
template <typename P>
__device__ __forceinline__ float2 tab_at(P p, uint32_t i) {
  const f32x2 t = p[i];
  return make_float2(t.x, t.y);
}

template <typename P>
__device__ __forceinline__ float2 dphase_of(uint64_t x, P lut, int chunks) {
  float2 v = tab_at(lut, (uint32_t)(x & 255u));
  for (int c = 1; c < chunks; ++c) v = cmul(v, tab_at(lut, (c << 8) | ((x >> (8 * c)) & 255u)));
  return v;
}

// Covariant init g(z) = (-1)^{Q(z)} prod_c w_c[(N z)_c] of a thread's 16
// registers z_j = x0 ^ r_j.  The map and the sign form are evaluated once per
// thread from the chunk tables; the registers follow from linearity,
// N z_j = N x0 ^ sum_i j_i col_i, and from the form rows (register_signs).  Only
// the weight product is looked up per amplitude (w staged in LDS).
__device__ __forceinline__ void cov_init(float2 (&v)[16], uint64_t x0, const PassArgs& a, const uint64_t* lut,
                                         const float2* w) {
  uint64_t b0 = 0, acc = 0;
  for (int c = 0; c < a.init_chunks; ++c) {
    const uint32_t e = (c << 8) | ((x0 >> (8 * c)) & 255u);
    b0 ^= lut[2 * e];
    acc ^= lut[2 * e + 1];
  }
  const uint32_t sg = register_signs((uint32_t)__popcll(acc & (x0 | (1ull << 63))) & 1u, x0, a.iq, a.ifreg);
  uint64_t bj[16];
  bj[0] = b0;
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 1 << i; j < (2 << i); ++j) bj[j] = bj[j - (1 << i)] ^ a.icol[i];
  const int wb = a.wbits;
  const uint32_t wm = (1u << wb) - 1u;
#pragma unroll
  for (int j = 0; j < 16; ++j) {
    const uint64_t b = bj[j];
    float2 val = w[b & wm];
    for (int c = 1; c < a.wchunks; ++c) val = cmul(val, w[(c << wb) | ((b >> (wb * c)) & wm)]);
    v[j] = cneg_if(val, (sg >> j) & 1u);
  }
}

// The same init with the chunk count fixed at compile time.  Register j's
// chunk-c index is the thread's index (N x0)_c XOR the uniform (M j)_c, M the
// register columns, so each lookup is one XOR of a per-thread byte offset
// (chunk base folded in: the index bits lie below it) with a scalar one,
// instead of a 64-bit shift and mask per lookup.  The products are taken in
// the same order as cov_init, so the results are the same bits.
template <int WCH>
__device__ __forceinline__ void cov_init_fixed(float2 (&v)[16], uint64_t x0, const PassArgs& a,
                                               const uint64_t* lut, const float2* w) {
  uint64_t b0 = 0, acc = 0;
  for (int c = 0; c < a.init_chunks; ++c) {
    const uint32_t e = (c << 8) | ((x0 >> (8 * c)) & 255u);
    b0 ^= lut[2 * e];
    acc ^= lut[2 * e + 1];
  }
  const uint32_t sg = register_signs((uint32_t)__popcll(acc & (x0 | (1ull << 63))) & 1u, x0, a.iq, a.ifreg);
  const int wb = a.wbits;
  const uint32_t wm = (1u << wb) - 1u;
  uint32_t off[WCH];   // byte offset of the thread's entry in chunk c
#pragma unroll
  for (int c = 0; c < WCH; ++c) off[c] = ((uint32_t)(b0 >> (wb * c)) & wm) << 3 | (uint32_t)c << (wb + 3);
  const unsigned char* wb8 = reinterpret_cast<const unsigned char*>(w);
#pragma unroll
  for (int j = 0; j < 16; ++j) {
    uint64_t dj = 0;   // uniform: M j
#pragma unroll
    for (int i = 0; i < 4; ++i)
      if (j & (1 << i)) dj ^= a.icol[i];
    float2 val = make_float2(0.f, 0.f);
#pragma unroll
    for (int c = 0; c < WCH; ++c) {
      const uint32_t d8 = ((uint32_t)(dj >> (wb * c)) & wm) << 3;
      const float2 e = *reinterpret_cast<const float2*>(wb8 + (off[c] ^ d8));
      val = c ? cmul(val, e) : e;
    }
    v[j] = cneg_if(val, (sg >> j) & 1u);
  }
}

// The covariant init with identical B unitaries: the weight of b depends on
// h = popc(b & o) <= |o| and l = popc(b & ~o) <= m - |o| only (o: the B
// outcome bits), so the table is staged in LDS as [h][l] - at most
// (m/2 + 1)^2 entries instead of (m + 1)^2 - and read once per amplitude;
// map, sign and register offsets as in cov_init_fixed.
__device__ __forceinline__ void cov_init_popc(float2 (&v)[16], uint64_t x0, const PassArgs& a,
                                              const uint64_t* lut, const float2* w) {
  uint64_t b0 = 0, acc = 0;
  for (int c = 0; c < a.init_chunks; ++c) {
    const uint32_t e = (c << 8) | ((x0 >> (8 * c)) & 255u);
    b0 ^= lut[2 * e];
    acc ^= lut[2 * e + 1];
  }
  const uint32_t sg = register_signs((uint32_t)__popcll(acc & (x0 | (1ull << 63))) & 1u, x0, a.iq, a.ifreg);
  const uint64_t o = a.womask, no = a.wnmask;
  const uint32_t st = (uint32_t)a.wstride;
#pragma unroll
  for (int j = 0; j < 16; ++j) {
    uint64_t dj = 0;   // uniform: M j
#pragma unroll
    for (int i = 0; i < 4; ++i)
      if (j & (1 << i)) dj ^= a.icol[i];
    const uint64_t b = b0 ^ dj;
    const uint32_t h = (uint32_t)__popcll(b & o), l = (uint32_t)__popcll(b & no);
    v[j] = cneg_if(w[h * st + l], (sg >> j) & 1u);
  }
}

// State traffic is streamed once per pass: non-temporal loads and stores keep
// it from displacing the lookup tables and the neighbouring tiles' rows in
// the caches (measured: 4-7% faster strided passes at n = 34).
__device__ __forceinline__ float2 load_stream(const float2* p) {
  const f32x2 t = __builtin_nontemporal_load(reinterpret_cast<const f32x2*>(p));
  return make_float2(t.x, t.y);
}
__device__ __forceinline__ void store_stream(float2* p, float2 v) {
  const f32x2 t = {v.x, v.y};
  __builtin_nontemporal_store(t, reinterpret_cast<f32x2*>(p));
}

// Wait for this wave's LDS traffic, then meet the other waves (global memory
// needs no ordering across the barrier: every workgroup owns its tile).
__device__ __forceinline__ void lds_barrier() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }

constexpr int cmax(int a, int b) { return a > b ? a : b; }
constexpr int cmin(int a, int b) { return a < b ? a : b; }

template <int TB, int C, int RBW>
struct Geo {
  static constexpr int R = 16;
  static constexpr int NT = 1 << (TB - 4);
  static constexpr int TLB = TB - 4 - RBW;  // thread bits below the round bits
  static constexpr int S1 = cmax(TB - 4, C);
  static constexpr int next(int s) { return cmax(s - 4, C); }
  static constexpr int count() {
    int k = 1;
    for (int s = S1; s > C; s = next(s)) ++k;
    return k;
  }
  static constexpr int at(int k) {
    int s = S1;
    for (int i = 0; i < k; ++i) s = next(s);
    return s;
  }
  static constexpr int hi(int k) { return k == 0 ? TB : at(k - 1); }
  static constexpr uint32_t slots = (1u << (TB - RBW)) + (1u << (TB - RBW - 4));
  static_assert(RBW == 0 || TLB >= 6, "round bits must be wave-uniform thread bits");
  static_assert(RBW <= C, "round bits are lane bits");
  static_assert(TB - C >= 4, "a strided tile needs four target bits");
};

// Compressed local index (round bits removed) of register 0 in layout S.
template <int S, int RBW>
__device__ __forceinline__ uint32_t comp_base(uint32_t tl) {
  constexpr int sc = S - RBW;
  return (tl & ((1u << sc) - 1u)) | ((tl >> sc) << (sc + 4));
}

// Full local index of register 0 in layout S.
template <int S, int C, int RBW>
__device__ __forceinline__ uint32_t local_base(uint32_t tl, uint32_t tr) {
  const uint32_t x = comp_base<S, RBW>(tl);
  constexpr int lo = C - RBW;
  return (x & ((1u << lo) - 1u)) | (tr << lo) | ((x >> lo) << C);
}

template <int S, int RBW>
__device__ __forceinline__ constexpr uint32_t slot_off(int j) {
  return ((uint32_t)j << (S - RBW)) + (((uint32_t)j << (S - RBW)) >> 4);
}

template <int TB, int C, int RBW, int SF, int ST>
__device__ __forceinline__ void exchange(float2 (&v)[16], float2* lds, uint32_t tl, uint32_t tr) {
  if constexpr (SF != ST) {
    const uint32_t wc = comp_base<SF, RBW>(tl), rc = comp_base<ST, RBW>(tl);
    const uint32_t wb = wc + (wc >> 4), rbs = rc + (rc >> 4);
#pragma unroll
    for (int r = 0; r < (1 << RBW); ++r) {
      if (tr == (uint32_t)r) {
#pragma unroll
        for (int j = 0; j < 16; ++j) lds[wb + slot_off<SF, RBW>(j)] = v[j];
      }
      lds_barrier();
      if (tr == (uint32_t)r) {
#pragma unroll
        for (int j = 0; j < 16; ++j) v[j] = lds[rbs + slot_off<ST, RBW>(j)];
      }
      lds_barrier();
    }
  }
}

template <int B>
__device__ __forceinline__ void wht_reg(float2 (&v)[16]) {
#pragma unroll
  for (int j = 0; j < 16; ++j)
    if (!(j & (1 << B))) {
      const float2 x = v[j], y = v[j | (1 << B)];
      v[j] = cadd(x, y);
      v[j | (1 << B)] = csub(x, y);
    }
}

// Rotation M(t) = [[1, -t], [t, 1]] on register bit B (|t| <= 1; the
// rotation's scale (1 + t^2)^(-1/2) is applied once for all qubits): two
// fused multiply-adds per amplitude.
template <int B>
__device__ __forceinline__ void rot_reg(float2 (&v)[16], float t) {
#pragma unroll
  for (int j = 0; j < 16; ++j)
    if (!(j & (1 << B))) {
      const float2 x = v[j], y = v[j | (1 << B)];
      // rounded as written (see cmul), whatever kernel inlines this
      v[j] = make_float2(__fmaf_rn(-t, y.x, x.x), __fmaf_rn(-t, y.y, x.y));
      v[j | (1 << B)] = make_float2(__fmaf_rn(t, x.x, y.x), __fmaf_rn(t, x.y, y.y));
    }
}

// Butterflies of layout K on the register bits it owns.
template <typename G, int K>
__device__ __forceinline__ void wht_layout(float2 (&v)[16]) {
  constexpr int s = G::at(K), h = G::hi(K);
  if constexpr (s + 0 < h) wht_reg<0>(v);
  if constexpr (s + 1 < h) wht_reg<1>(v);
  if constexpr (s + 2 < h) wht_reg<2>(v);
  if constexpr (s + 3 < h) wht_reg<3>(v);
}

// Every target bit of the layout is rotated: a qubit at or above gate_limit
// has t = 0, the identity (x - 0 y = x and y + 0 x = y exactly, up to the sign
// of a zero).  A runtime skip per bit would make the compiler copy all 32
// rotated registers back at the branch's merge.
template <typename G, int C, int K>
__device__ __forceinline__ void rot_layout(float2 (&v)[16], const PassArgs& a) {
  constexpr int s = G::at(K), h = G::hi(K);
  // target bit index of register bit B: s + B - C
  if constexpr (s + 0 < h) rot_reg<0>(v, a.tng[s + 0 - C]);
  if constexpr (s + 1 < h) rot_reg<1>(v, a.tng[s + 1 - C]);
  if constexpr (s + 2 < h) rot_reg<2>(v, a.tng[s + 2 - C]);
  if constexpr (s + 3 < h) rot_reg<3>(v, a.tng[s + 3 - C]);
}

// Walk down the layouts from K applying the Walsh butterflies (ends at layout count-1).
template <int TB, int C, int RBW, int K>
__device__ __forceinline__ void walk_wht(float2 (&v)[16], float2* lds, uint32_t tl, uint32_t tr) {
  using G = Geo<TB, C, RBW>;
  wht_layout<G, K>(v);
  if constexpr (K + 1 < G::count()) {
    exchange<TB, C, RBW, G::at(K), G::at(K + 1)>(v, lds, tl, tr);
    walk_wht<TB, C, RBW, K + 1>(v, lds, tl, tr);
  }
}

// Walk up from layout K applying the Walsh butterflies (ends at layout 0 = S1).
template <int TB, int C, int RBW, int K>
__device__ __forceinline__ void walk_wht_up(float2 (&v)[16], float2* lds, uint32_t tl, uint32_t tr) {
  using G = Geo<TB, C, RBW>;
  wht_layout<G, K>(v);
  if constexpr (K > 0) {
    exchange<TB, C, RBW, G::at(K), G::at(K - 1)>(v, lds, tl, tr);
    walk_wht_up<TB, C, RBW, K - 1>(v, lds, tl, tr);
  }
}

// Rotations walking down from layout K (ends at layout count-1).
template <int TB, int C, int RBW, int K>
__device__ __forceinline__ void walk_rot_down(float2 (&v)[16], float2* lds, uint32_t tl, uint32_t tr,
                                              const PassArgs& a) {
  using G = Geo<TB, C, RBW>;
  rot_layout<G, C, K>(v, a);
  if constexpr (K + 1 < G::count()) {
    exchange<TB, C, RBW, G::at(K), G::at(K + 1)>(v, lds, tl, tr);
    walk_rot_down<TB, C, RBW, K + 1>(v, lds, tl, tr, a);
  }
}

// Rotations walking up from layout K (ends at layout 0 = S1).
template <int TB, int C, int RBW, int K>
__device__ __forceinline__ void walk_rot_up(float2 (&v)[16], float2* lds, uint32_t tl, uint32_t tr,
                                            const PassArgs& a) {
  using G = Geo<TB, C, RBW>;
  rot_layout<G, C, K>(v, a);
  if constexpr (K > 0) {
    exchange<TB, C, RBW, G::at(K), G::at(K - 1)>(v, lds, tl, tr);
    walk_rot_up<TB, C, RBW, K - 1>(v, lds, tl, tr, a);
  }
}

// Diagonal phases of the register bits layout K owns (the low h - s bits of
// the register index): register j takes lph[K][j & own].
template <typename G, int K>
__device__ __forceinline__ void phase_layout(float2 (&v)[16], const PassArgs& a) {
  constexpr int s = G::at(K), h = G::hi(K);
  constexpr int own = (1 << cmin(4, h - s)) - 1;
#pragma unroll
  for (int j = 1; j < 16; ++j)
    if (j & own) v[j] = cmul(v[j], a.lph[K][j & own]);
}

// One walk down from layout K (ends at layout count-1): each layout applies
// the Walsh butterflies, the diagonal phases and the rotations of the bits it
// owns, in that order (F_LAYER).
template <int TB, int C, int RBW, int K>
__device__ __forceinline__ void walk_layer(float2 (&v)[16], float2* lds, uint32_t tl, uint32_t tr,
                                           const PassArgs& a, int flags) {
  using G = Geo<TB, C, RBW>;
  static_assert(G::count() <= 4, "lph holds four layouts");
  if (flags & F_WHT) wht_layout<G, K>(v);
  if (flags & F_PHASE) phase_layout<G, K>(v, a);
  if (flags & F_GATES) rot_layout<G, C, K>(v, a);
  if constexpr (K + 1 < G::count()) {
    exchange<TB, C, RBW, G::at(K), G::at(K + 1)>(v, lds, tl, tr);
    walk_layer<TB, C, RBW, K + 1>(v, lds, tl, tr, a, flags);
  }
}

// FP64 masses of the 2^mass_bits blocks of a contiguous tile at state index
// base, from registers in layout 0 (local bits 0..3): the pairwise tree over
// the thread's 16 values, then across the lanes.  A block's pairwise sum is
// the same in state and outcome order (a XOR relabelling keeps every pair),
// so only its slot moves.
__device__ __forceinline__ void write_masses(const float2 (&v)[16], const PassArgs& a, uint64_t base,
                                             uint32_t tid) {
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
  for (int b = 0; b < sb; ++b) m = __dadd_rn(m, __shfl_xor(m, 1 << b));
  if ((tid & ((1u << sb) - 1u)) == 0) a.mass[((base >> a.mass_bits) + (tid >> sb)) ^ a.mflip] = m;
}

// One tile of a pass: tile index blk of the state described by a.  With
// LEAF (contiguous passes only) the tile's values go to leaf[0 .. 2^TB)
// instead of back into the state.  FL != 0 fixes the pass's flags at compile
// time (they must equal a.flags): the branches of the other pass kinds drop
// out of the kernel.
template <int TB, int C, int RBW, bool LEAF = false, int FL = 0>
__device__ __forceinline__ void tile_run(const PassArgs& a, uint64_t blk, float2* leaf = nullptr) {
  static_assert(!LEAF || C == 0, "leaf tiles are contiguous");
  const int flags = FL ? FL : a.flags;
  using G = Geo<TB, C, RBW>;
  constexpr int S1 = G::S1;
  constexpr int SL = G::at(G::count() - 1);  // == C
  extern __shared__ __align__(16) unsigned char smem[];
  float2* lds = reinterpret_cast<float2*>(smem);
  // LDS after the exchange slots: the diagonal-phase table of the phase
  // pass and the covariant init's weights (read per amplitude).  Sign tables
  // are only read at the tile's own (uniform) index (split_form_bit), and the
  // covariant map once per thread, from global memory.
  float2* lut_d = reinterpret_cast<float2*>(smem + (size_t)G::slots * sizeof(float2));
  float2* lut_w = lut_d + (a.dphase_global ? 0 : (size_t)a.dphase_chunks * 256);
  const uint64_t* lut_init = a.init_lut;
  const uint32_t tid = threadIdx.x;
  const uint32_t tl = tid & ((1u << G::TLB) - 1u);
  const uint32_t tr = RBW ? __builtin_amdgcn_readfirstlane(tid >> G::TLB) : 0u;

  const bool cov = !(flags & F_LOAD) && a.init_kind == RGS_INIT_COV;
  // the diagonal-phase table is read through LDS or straight from global
  // memory (a per-plan choice: which is faster depends on the pass's other
  // work); each has its own code path, so LDS reads stay LDS instructions
  if ((a.dphase_chunks && !a.dphase_global) || cov) {
    if (!a.dphase_global)
      for (int i = tid; i < a.dphase_chunks * 256; i += G::NT) lut_d[i] = a.dphase[i];
    if (cov)
      for (int i = tid; i < a.wentries; i += G::NT) {
        // the rotations' scale enters every weight once: the popcount table's
        // entries, or the first chunk table of a product of chunk tables.
        // Popcount entry (h, l) is the table's [h + l][h].
        uint32_t src = (uint32_t)i;
        if (a.wpop) {
          const uint32_t h = (uint32_t)i / (uint32_t)a.wstride, l = (uint32_t)i - h * (uint32_t)a.wstride;
          src = (h + l) * (uint32_t)a.wpop + h;
        }
        const float2 w = a.init_w[src];
        const float sc = (a.wpop || i < (1 << a.wbits)) ? a.wscale : 1.f;
        lut_w[i] = make_float2(__fmul_rn(w.x, sc), __fmul_rn(w.y, sc));
      }
    __syncthreads();
  }

  const int mid = a.qstart - C;
  const uint32_t cmask = (1u << C) - 1u;
  const uint64_t base = ((blk & ((1ull << mid) - 1ull)) << C) | ((blk >> mid) << (a.qstart + TB - C));
  auto gidx = [&](uint32_t l) -> uint64_t {
    return base | (uint64_t)(l & cmask) | ((uint64_t)(l >> C) << a.qstart);
  };
  // the load/store layout S1: thread offset and the global stride of its registers
  const uint64_t g0 = gidx(local_base<S1, C, RBW>(tl, tr));
  const int gsh = a.qstart + S1 - C;

  // a state made in place starts in layout SL when it walks up, else in S1
  const bool up = C == 0 && (flags & F_INIT_UP);
  const uint64_t gi = up ? gidx(local_base<SL, C, RBW>(tl, tr)) : g0;
  const int gish = up ? a.qstart + SL - C : gsh;

  float2 v[16];
  if (flags & F_LOAD) {
#pragma unroll
    for (int j = 0; j < 16; ++j) v[j] = load_stream(a.state + g0 + ((uint64_t)j << gsh));
    if (a.lscale != 1.f) {
#pragma unroll
      for (int j = 0; j < 16; ++j) v[j] = make_float2(__fmul_rn(v[j].x, a.lscale), __fmul_rn(v[j].y, a.lscale));
    }
  } else if (a.init_kind == RGS_INIT_GRAPH) {
    const uint32_t li = up ? local_base<SL, C, RBW>(tl, tr) : local_base<S1, C, RBW>(tl, tr);
    const uint32_t fb = phase_parity(base, a.init_lut, a.init_chunks);
    const uint32_t e = register_signs(split_form_bit<TB>(fb, base, li, a.iform), gi, a.iq, a.ifreg);
#pragma unroll
    for (int j = 0; j < 16; ++j) v[j] = cneg_if(make_float2(a.init_scale, 0.f), (e >> j) & 1u);
  } else if (a.init_kind == RGS_INIT_COV) {
    if (a.wpop) {
      cov_init_popc(v, gi, a, lut_init, lut_w);
    } else if constexpr (C == 0) {
      // the contiguous pass that makes the state: chunk counts of m <= 45
      switch (a.wchunks) {
        case 1: cov_init_fixed<1>(v, gi, a, lut_init, lut_w); break;
        case 2: cov_init_fixed<2>(v, gi, a, lut_init, lut_w); break;
        case 3: cov_init_fixed<3>(v, gi, a, lut_init, lut_w); break;
        case 4: cov_init_fixed<4>(v, gi, a, lut_init, lut_w); break;
        case 5: cov_init_fixed<5>(v, gi, a, lut_init, lut_w); break;
        default: cov_init(v, gi, a, lut_init, lut_w); break;
      }
    } else {
      cov_init(v, gi, a, lut_init, lut_w);
    }
  } else {  // RGS_INIT_DELTA
#pragma unroll
    for (int j = 0; j < 16; ++j)
      v[j] = (gi | ((uint64_t)j << gish)) == 0 ? make_float2(a.init_scale, 0.f) : make_float2(0.f, 0.f);
  }

  // Diagonal phases on the registers of layout S.  Both are separable over
  // the register bits: the sign through register_signs, the complex phase as the
  // thread's base phase times the single-bit phases of the register bits.
  auto apply_phase = [&](auto s_tag) {
    constexpr int S = decltype(s_tag)::value;
    const uint64_t gb = gidx(local_base<S, C, RBW>(tl, tr));
    const uint32_t sh = a.qstart + S - C;
    auto apply_d = [&](auto tab) {
      float2 ph[16];
      ph[0] = dphase_of(gb, tab, a.dphase_chunks);
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const uint32_t p = sh + i;
        const float2 f = tab_at(tab, ((p >> 3) << 8) | (1u << (p & 7)));
#pragma unroll
        for (int j = 1 << i; j < (2 << i); ++j) ph[j] = cmul(ph[j - (1 << i)], f);
      }
#pragma unroll
      for (int j = 0; j < 16; ++j) v[j] = cmul(v[j], ph[j]);
    };
    if (a.dphase_chunks) {
      if (a.dphase_global)
        apply_d((glb_f2)a.dphase);
      else
        apply_d((lds_f2)lut_d);
    }
    if (a.phase_chunks) {
      const uint32_t fb = phase_parity(base, a.phase_lut, a.phase_chunks);
      const uint32_t e = register_signs(split_form_bit<TB>(fb, base, local_base<S, C, RBW>(tl, tr), a.pform), gb,
                                        a.pq, a.pfreg);
#pragma unroll
      for (int j = 0; j < 16; ++j) v[j] = cneg_if(v[j], (e >> j) & 1u);
    }
  };

  // After the walks the registers are in layout S1 (top) or SL (bottom).
  bool at_top;
  if (up) {
    if constexpr (C == 0) walk_wht_up<TB, C, RBW, G::count() - 1>(v, lds, tl, tr);
    at_top = true;
  } else if (flags & F_LAYER) {
    walk_layer<TB, C, RBW, 0>(v, lds, tl, tr, a, flags);
    at_top = (SL == S1);
  } else if (flags & F_WHT) {
    walk_wht<TB, C, RBW, 0>(v, lds, tl, tr);
    if (flags & F_PHASE) apply_phase(std::integral_constant<int, SL>());
    if (flags & F_GATES) {
      walk_rot_up<TB, C, RBW, G::count() - 1>(v, lds, tl, tr, a);
      at_top = true;
    } else {
      at_top = (SL == S1);
    }
  } else {
    if (flags & F_PHASE) apply_phase(std::integral_constant<int, S1>());
    if (flags & F_GATES) {
      walk_rot_down<TB, C, RBW, 0>(v, lds, tl, tr, a);
      at_top = (SL == S1);
    } else {
      at_top = true;
    }
  }
  if constexpr (C == 0) {
    if (flags & F_MASS) {
      // registers must hold local bits 0..3 (layout SL = 0)
      if (at_top) exchange<TB, C, RBW, S1, SL>(v, lds, tl, tr);
      at_top = (SL == S1);
      write_masses(v, a, base, tid);
      if (flags & F_NOSTORE) return;
    }
  }
  if (!at_top) {
    if constexpr (C > 0) {
      // every strided layout keeps the row bits on the lowest lanes, so the
      // tile is stored from layout SL as coalesced as from S1
      const uint64_t gl = gidx(local_base<SL, C, RBW>(tl, tr));
      const int lsh = a.qstart + SL - C;
#pragma unroll
      for (int j = 0; j < 16; ++j) store_stream(a.state + gl + ((uint64_t)j << lsh), v[j]);
      return;
    } else {
      exchange<TB, C, RBW, SL, S1>(v, lds, tl, tr);
    }
  }
  if constexpr (LEAF) {
    // g0 - base: the thread's offset inside the tile
#pragma unroll
    for (int j = 0; j < 16; ++j) leaf[(g0 - base) + ((uint64_t)j << gsh)] = v[j];
  } else {
#pragma unroll
    for (int j = 0; j < 16; ++j) store_stream(a.state + g0 + ((uint64_t)j << gsh), v[j]);
  }
}

// The contiguous pass's values of listed tiles: workgroup b recomputes tile
// tiles[b] into out[b * 2^TB ..] (an index past the state's tiles is skipped).
template <int TB, int RBW>
__global__ void __launch_bounds__(1 << (TB - 4)) k_tile_leaf(PassArgs a, const int64_t* __restrict__ tiles,
                                                            float2* __restrict__ out) {
  const uint64_t t = (uint64_t)tiles[blockIdx.x];
  if (t >> (a.nbits - TB)) return;
  tile_run<TB, 0, RBW, true>(a, t, out + ((uint64_t)blockIdx.x << TB));
}

template <int TB, int RBW>
hipError_t launch_tile_leaf(const PassArgs& a, const int64_t* tiles, int count, float2* out, size_t shmem,
                            hipStream_t st) {
  static bool attr_set = false;
  if (!attr_set) {
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_tile_leaf<TB, RBW>),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)kLdsBudget);
    if (e != hipSuccess) return e;
    attr_set = true;
  }
  hipLaunchKernelGGL((k_tile_leaf<TB, RBW>), dim3((unsigned)count), dim3(1 << (TB - 4)), shmem, st, a, tiles, out);
  return hipGetLastError();
}

template <int TB, int C, int RBW>
__global__ void __launch_bounds__(1 << (TB - 4)) k_tile(PassArgs a) {
  tile_run<TB, C, RBW>(a, (uint64_t)blockIdx.x + ((uint64_t)blockIdx.y << 24));
}

// The contiguous mass pass with its flags fixed (the gates and masses of the
// last pass, storing the state or not): 49.0 -> 46.3 ms for the storing pass
// at n = 34.  (A masses-only variant that loads each next tile while it
// rotates the current one measured the same, 29.6-29.8 ms at n = 34: the
// pass is bound by instruction issue, not by load latency.)
constexpr int kMassPass = F_LOAD | F_GATES | F_MASS;
constexpr int kMassOnlyPass = F_LOAD | F_GATES | F_MASS | F_NOSTORE;
// The passes of a fuse_walsh schedule that load the state: strided
// [wht][phase][gates] and the contiguous [wht][phase][gates][mass] (storing
// the state or not).
constexpr int kFusedPass = F_LOAD | F_WHT | F_PHASE | F_GATES | F_LAYER;
constexpr int kFusedMassPass = kFusedPass | F_MASS;
constexpr int kFusedMassOnlyPass = kFusedPass | F_MASS | F_NOSTORE;
template <int TB, int RBW, int FL>
__global__ void __launch_bounds__(1 << (TB - 4)) k_tile_mass(PassArgs a) {
  tile_run<TB, 0, RBW, false, FL>(a, (uint64_t)blockIdx.x + ((uint64_t)blockIdx.y << 24));
}
template <int TB, int C, int RBW, int FL>
__global__ void __launch_bounds__(1 << (TB - 4)) k_tile_fixed(PassArgs a) {
  tile_run<TB, C, RBW, false, FL>(a, (uint64_t)blockIdx.x + ((uint64_t)blockIdx.y << 24));
}

// The same pass over many states of one plan in one launch.  The states of
// a plan group have the same size, so each owns 2^shift consecutive tiles:
// workgroup b runs tile b & (2^shift - 1) of state b >> shift, with that
// state's arguments read from global memory.
template <int TB, int C, int RBW>
__global__ void __launch_bounds__(1 << (TB - 4)) k_tile_multi(const PassArgs* __restrict__ args, int shift,
                                                             int count) {
  const uint64_t b = (uint64_t)blockIdx.x + ((uint64_t)blockIdx.y << 24);
  const uint64_t s = b >> shift;
  if (s >= (uint64_t)count) return;   // the grid's last row may run past the tiles
  tile_run<TB, C, RBW>(args[s], b & ((1ull << shift) - 1ull));
}

template <int TB, int C, int RBW>
hipError_t launch_tile_multi(const PassArgs* args, int shift, int count, size_t shmem, hipStream_t st) {
  static bool attr_set = false;
  if (!attr_set) {
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_tile_multi<TB, C, RBW>),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)kLdsBudget);
    if (e != hipSuccess) return e;
    attr_set = true;
  }
  const uint64_t tiles = (uint64_t)count << shift;
  dim3 grid((unsigned)(tiles < (1ull << 24) ? tiles : (1ull << 24)), (unsigned)((tiles + (1ull << 24) - 1) >> 24));
  hipLaunchKernelGGL((k_tile_multi<TB, C, RBW>), grid, dim3(1 << (TB - 4)), shmem, st, args, shift, count);
  return hipGetLastError();
}

// One workgroup per tile: kernel K over the state's 2^(nbits - TB) tiles.
template <int TB>
hipError_t launch_tiles(const void* kernel, const PassArgs& a, size_t shmem, hipStream_t st, bool& attr_set) {
  if (!attr_set) {
    hipError_t e = hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)kLdsBudget);
    if (e != hipSuccess) return e;
    attr_set = true;
  }
  const uint64_t tiles = 1ull << (a.nbits - TB);
  dim3 grid((unsigned)(tiles < (1ull << 24) ? tiles : (1ull << 24)), (unsigned)(tiles >> 24 ? tiles >> 24 : 1));
  void* args[] = {const_cast<PassArgs*>(&a)};
  return hipLaunchKernel(kernel, grid, dim3(1 << (TB - 4)), args, shmem, st);
}

template <int TB, int C, int RBW>
hipError_t launch_tile(const PassArgs& a, size_t shmem, hipStream_t st) {
  if constexpr (C == 0) {
    static bool mass_attr = false, mass_only_attr = false, fmass_attr = false, fmass_only_attr = false;
    if (a.flags == kMassPass)
      return launch_tiles<TB>(reinterpret_cast<const void*>(&k_tile_mass<TB, RBW, kMassPass>), a, shmem, st,
                              mass_attr);
    if (a.flags == kMassOnlyPass)
      return launch_tiles<TB>(reinterpret_cast<const void*>(&k_tile_mass<TB, RBW, kMassOnlyPass>), a, shmem, st,
                              mass_only_attr);
    if (a.flags == kFusedMassPass)
      return launch_tiles<TB>(reinterpret_cast<const void*>(&k_tile_mass<TB, RBW, kFusedMassPass>), a, shmem, st,
                              fmass_attr);
    if (a.flags == kFusedMassOnlyPass)
      return launch_tiles<TB>(reinterpret_cast<const void*>(&k_tile_mass<TB, RBW, kFusedMassOnlyPass>), a, shmem,
                              st, fmass_only_attr);
  } else {
    static bool fused_attr = false;
    if (a.flags == kFusedPass)
      return launch_tiles<TB>(reinterpret_cast<const void*>(&k_tile_fixed<TB, C, RBW, kFusedPass>), a, shmem, st,
                              fused_attr);
  }
  static bool attr_set = false;
  return launch_tiles<TB>(reinterpret_cast<const void*>(&k_tile<TB, C, RBW>), a, shmem, st, attr_set);
}

// Dispatch over the instantiated specialisations of one lane-bit count
// (defined in slab_k<C>.hip).
hipError_t launch_c0(int tb, int rbw, const PassArgs& a, size_t shmem, hipStream_t st);
hipError_t launch_c2(int tb, int rbw, const PassArgs& a, size_t shmem, hipStream_t st);
hipError_t launch_c3(int tb, int rbw, const PassArgs& a, size_t shmem, hipStream_t st);
hipError_t launch_c4(int tb, int rbw, const PassArgs& a, size_t shmem, hipStream_t st);
hipError_t launch_c5(int tb, int rbw, const PassArgs& a, size_t shmem, hipStream_t st);
hipError_t launch_c6(int tb, int rbw, const PassArgs& a, size_t shmem, hipStream_t st);
hipError_t launch_c7(int tb, int rbw, const PassArgs& a, size_t shmem, hipStream_t st);
// The leaf-tile launches of the contiguous specialisations (slab_k0.hip).
hipError_t launch_c0_leaf(int tb, int rbw, const PassArgs& a, const int64_t* tiles, int count, float2* out,
                          size_t shmem, hipStream_t st);
// The multi-state launches of the same specialisations (count states of
// 2^shift tiles each).
#define RGS_MULTI_DECL(NAME) \
  hipError_t NAME(int tb, int rbw, const PassArgs* args, int shift, int count, size_t shmem, hipStream_t st);
RGS_MULTI_DECL(launch_c0_multi)
RGS_MULTI_DECL(launch_c2_multi)
RGS_MULTI_DECL(launch_c3_multi)
RGS_MULTI_DECL(launch_c4_multi)
RGS_MULTI_DECL(launch_c5_multi)
RGS_MULTI_DECL(launch_c6_multi)
RGS_MULTI_DECL(launch_c7_multi)
#undef RGS_MULTI_DECL

}  // namespace rgs
