// Batched linear rank-width scan: p(x) of many proposals in one persistent
// launch.  One workgroup of W wavefronts owns P proposals at a time (one or
// two) and sweeps all n steps with their environment tables resident in LDS
// (2^rmax entries each); the table addresses and signs of a step depend on
// the plan only and are shared by the P proposals.  Outputs of a step are
// gathered into registers (K per lane), the waves combine their maximum
// modulus through LDS, and the outputs are written back in place, their scale
// carried as an exact power of two, so there is no host round trip between
// steps.  Sums and weights are FP64; the table is stored as complex64
// (S = float2) or complex128 (S = double2, the audit precision).
#include "rgs_common.h"
#include "rgs_amd.h"

#include <algorithm>
#include <mutex>
#include <string>

namespace rgs {

template <typename S>
__device__ __forceinline__ double2 load_c(const S& s) {
  return make_double2((double)s.x, (double)s.y);
}

// An entry widened to FP64 with its sign flipped when sgn = 0x80000000.
__device__ __forceinline__ double2 load_signed(const float2& s, uint32_t sgn) {
  return make_double2((double)__uint_as_float(__float_as_uint(s.x) ^ sgn),
                      (double)__uint_as_float(__float_as_uint(s.y) ^ sgn));
}
__device__ __forceinline__ double2 load_signed(const double2& s, uint32_t sgn) {
  const uint64_t m = (uint64_t)sgn << 32;
  return make_double2(__longlong_as_double(__double_as_longlong(s.x) ^ m),
                      __longlong_as_double(__double_as_longlong(s.y) ^ m));
}

// x with its sign flipped when neg is 1 (an integer XOR on the sign bit: exact,
// and scalar when x and neg are uniform).
__device__ __forceinline__ double flip_sign(double x, uint32_t neg) {
  return __longlong_as_double(__double_as_longlong(x) ^ ((long long)neg << 63));
}

template <typename S>
__device__ __forceinline__ S make_s(double re, double im);
template <>
__device__ __forceinline__ float2 make_s<float2>(double re, double im) {
  return make_float2((float)re, (float)im);
}
template <>
__device__ __forceinline__ double2 make_s<double2>(double re, double im) {
  return make_double2(re, im);
}

// Maximum of a non-negative 32-bit value over the wavefront: DPP within each
// row of 16 lanes (quad swaps, then rotations by 4 and 8), then the four row
// results read into scalars.  All 64 lanes must be active.
__device__ __forceinline__ uint32_t wave_max_u32(uint32_t v) {
  v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0xB1, 0xF, 0xF, false));    // quad_perm 1,0,3,2
  v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x4E, 0xF, 0xF, false));    // quad_perm 2,3,0,1
  v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x124, 0xF, 0xF, false));   // row_ror 4
  v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x128, 0xF, 0xF, false));   // row_ror 8
  const uint32_t r0 = __builtin_amdgcn_readlane(v, 0), r1 = __builtin_amdgcn_readlane(v, 16);
  const uint32_t r2 = __builtin_amdgcn_readlane(v, 32), r3 = __builtin_amdgcn_readlane(v, 48);
  return max(max(r0, r1), max(r2, r3));
}

// The high word of a non-negative double orders like the double itself, so
// the maximum of |out|^2 is tracked as the maximum high word; only its
// exponent E is used.  The renormalisation exponent is e = floor(E / 2) + 1
// (the frexp exponent of sqrt(max |out|^2)), so max |out| 2^-e lies in
// [1/2, 1); a zero or subnormal maximum leaves the table unscaled.
__device__ __forceinline__ uint32_t sq_hi(double re, double im) {
  return (uint32_t)(__double_as_longlong(fma(re, re, im * im)) >> 32);
}
__device__ __forceinline__ int scale_exponent(uint32_t mxh) {
  const int biased = (int)(mxh >> 20);
  return biased ? ((biased - 1023) >> 1) + 1 : 0;
}

// One step's uniform constants.  The step's 2^nker kernel combinations (y = 1
// in the upper half when split) are passed beside them, from the plan in
// global memory or from an LDS copy.
struct ScanStepArgs {
  uint32_t beta, mask;
  int r_in, nker;
  double2 f0, f1;
};

// Outputs of one step for this lane.  The wave's blocks are visited in Gray
// order g(kk) = kk ^ (kk >> 1), block k = wave + W g(kk), so that each block's
// p0 = lo ^ hi[k] differs from the previous one by a single uniform word
// hstep[ctz(kk)] = hi[W 2^ctz(kk)]; lo_w = lo ^ hi[wave] is per lane.  Blocks
// kk < nkw are live (nkw uniform).  Term counts 2^NKER (NKER <= 2 in every
// packed plan) are compile-time constants, so the table reads of all outputs
// are straight-line loads.
// Split steps: every p0 has y = 0; the first half of the combinations keeps
// y = 0 (weight f0, no sign), the second has y = 1 (weight f1, sign
// (-1)^(c . beta)).  Plain steps: every term of an output has y = y(p0),
// weight f_y and sign (-1)^(y (c . beta)); the lane's part of y is folded
// into the per-lane weights fa / fb and sign masks ba / bb once per step.
// Returns the lane's maximum high word of |out|^2.
template <typename S, int K, int NKER, bool SPLIT, typename CP>
__device__ __forceinline__ uint32_t scan_outputs(const S* E, const ScanStepArgs& a, CP combos, uint32_t lo_w,
                                                 const uint32_t* hstep, int nkw, S (&outv)[K]) {
#pragma clang fp contract(off)
  constexpr int NC = 1 << NKER, HALF = SPLIT ? NC / 2 : NC;
  uint32_t cmb[NC];
#pragma unroll
  for (int t = 0; t < NC; ++t) cmb[t] = (uint32_t)combos[t];
  const uint32_t ylo = SPLIT ? 0u : (lo_w >> a.r_in) & 1u;
  const double2 fa = ylo ? a.f1 : a.f0, fb = ylo ? a.f0 : a.f1;   // weights for y(hi part) = 0 / 1
  const uint32_t ba = ylo ? a.beta : 0u, bb = a.beta ^ ba;
  uint32_t mxh = 0, hs = 0;
#pragma unroll
  for (int kk = 0; kk < K; ++kk) {
    if (kk) hs ^= hstep[__builtin_ctz(kk)];
    if (kk < nkw) {
      const uint32_t p0 = lo_w ^ hs;
      const bool yh = !SPLIT && ((hs >> a.r_in) & 1u);
      const uint32_t bm = yh ? bb : ba;
      double s0r = 0.0, s0i = 0.0, s1r = 0.0, s1i = 0.0;
#pragma unroll
      for (int t = 0; t < NC; ++t) {
        const uint32_t c = (p0 ^ cmb[t]) & a.mask;
        const uint32_t sg = t < HALF ? (SPLIT ? 0u : (uint32_t)__popc(c & bm) << 31)
                                     : (uint32_t)__popc(c & a.beta) << 31;
        const double2 e = load_signed(E[c], sg);
        if (t == 0) {
          s0r = e.x;
          s0i = e.y;
        } else if (t < HALF) {
          s0r += e.x;
          s0i += e.y;
        } else if (t == HALF) {
          s1r = e.x;
          s1i = e.y;
        } else {
          s1r += e.x;
          s1i += e.y;
        }
      }
      // explicit fused multiply-adds (contraction is off in these helpers):
      // every kernel that sums through them rounds the same way
      double orr, oi;
      if (SPLIT) {
        orr = fma(s1r, a.f1.x, fma(-s1i, a.f1.y, fma(s0r, a.f0.x, -s0i * a.f0.y)));
        oi = fma(s1r, a.f1.y, fma(s1i, a.f1.x, fma(s0r, a.f0.y, s0i * a.f0.x)));
      } else {
        const double2 f = yh ? fb : fa;
        orr = fma(s0r, f.x, -s0i * f.y);
        oi = fma(s0r, f.y, s0i * f.x);
      }
      mxh = max(mxh, sq_hi(orr, oi));
      outv[kk] = make_s<S>(orr, oi);
    }
  }
  return mxh;
}

// Outputs of one step for this lane on a step whose table coordinates make
// y(p0) and the sign form vanish on the lane bits (the packer's uniform
// flag): within a 64-output block every output has the same weight and every
// term t the same sign, decided from the block word b = hi[wave] ^ hs alone.
// A term's table index is the lane's masked word lm ^ ((b ^ combo_t) & mask);
// its sign is uniform, so the signed sums become fused multiply-adds by a
// scalar +-1 and the first term's sign moves into the weight.  The indices
// and signs depend on the plan only, so P proposals (tables TBYTES apart,
// weights f0[q], f1[q]) share them.  With TRACK, sets each proposal's maximum
// high word of |out|^2 over the lane's outputs.
#define LDS_AS __attribute__((address_space(3)))
// lm8: the lane's masked word as the LDS byte address of its entry in the
// first table (the tables start at the LDS base, which is 0: the kernel has
// no static LDS), so a term's address is one XOR with a uniform offset.
template <typename S, int K, int NKER, bool SPLIT, int P, int TBYTES, bool TRACK, typename CP>
__device__ __forceinline__ void scan_outputs_u(const ScanStepArgs& a, const double2 (&f0)[P], const double2 (&f1)[P],
                                               CP combos, uint32_t lm8, uint32_t hw, const uint32_t* hstep, int nkw,
                                               S (&outv)[P][K], uint32_t (&mxh)[P]) {
#pragma clang fp contract(off)
  constexpr int NC = 1 << NKER, HALF = SPLIT ? NC / 2 : NC;
  constexpr int SH = sizeof(S) == 8 ? 3 : 4;   // log2 of an entry's bytes
  uint32_t cm8[NC], par[NC];
#pragma unroll
  for (int t = 0; t < NC; ++t) {
    const uint32_t cm = (uint32_t)combos[t] & a.mask;
    cm8[t] = cm << SH;
    par[t] = (uint32_t)__popc(cm & a.beta) & 1u;
  }
#pragma unroll
  for (int q = 0; q < P; ++q) mxh[q] = 0;
  uint32_t hs = 0;
#pragma unroll
  for (int kk = 0; kk < K; ++kk) {
    if (kk) hs ^= hstep[__builtin_ctz(kk)];
    if (kk < nkw) {
      const uint32_t b = hw ^ hs;
      const uint32_t bm = b & a.mask;
      const uint32_t yb = SPLIT ? 0u : (b >> a.r_in) & 1u;
      const uint32_t sb = (uint32_t)__popc(bm & a.beta) & 1u;
      // sign of term t: split steps sign only the second half (y = 1);
      // plain steps sign every term when y = 1
      uint32_t neg[NC];
#pragma unroll
      for (int t = 0; t < NC; ++t) neg[t] = SPLIT ? (t < HALF ? 0u : sb ^ par[t]) : (yb & (sb ^ par[t]));
      const uint32_t row8 = lm8 ^ (bm << SH);
      double s0r[P], s0i[P], s1r[P], s1i[P];
#pragma unroll
      for (int t = 0; t < NC; ++t) {
        // one address for every proposal: proposal q's table is TBYTES
        // further on (a constant offset of the LDS read)
        const uint32_t at = row8 ^ cm8[t];
        const int first = t < HALF ? 0 : HALF;
        const double rho = (neg[t] ^ neg[first]) ? -1.0 : 1.0;
#pragma unroll
        for (int q = 0; q < P; ++q) {
          const double2 e = load_c(*reinterpret_cast<const S LDS_AS*>(at + q * TBYTES));
          if (t == 0) {
            s0r[q] = e.x;
            s0i[q] = e.y;
          } else if (t < HALF) {
            s0r[q] = fma(e.x, rho, s0r[q]);
            s0i[q] = fma(e.y, rho, s0i[q]);
          } else if (t == HALF) {
            s1r[q] = e.x;
            s1i[q] = e.y;
          } else {
            s1r[q] = fma(e.x, rho, s1r[q]);
            s1i[q] = fma(e.y, rho, s1i[q]);
          }
        }
      }
#pragma unroll
      for (int q = 0; q < P; ++q) {
        double orr, oi;
        if (SPLIT) {
          const double g1x = flip_sign(f1[q].x, neg[HALF]), g1y = flip_sign(f1[q].y, neg[HALF]);
          orr = fma(s1r[q], g1x, fma(-s1i[q], g1y, fma(s0r[q], f0[q].x, -s0i[q] * f0[q].y)));
          oi = fma(s1r[q], g1y, fma(s1i[q], g1x, fma(s0r[q], f0[q].y, s0i[q] * f0[q].x)));
        } else {
          const double fx = flip_sign(yb ? f1[q].x : f0[q].x, neg[0]);
          const double fy = flip_sign(yb ? f1[q].y : f0[q].y, neg[0]);
          orr = fma(s0r[q], fx, -s0i[q] * fy);
          oi = fma(s0r[q], fy, s0i[q] * fx);
        }
        if (TRACK) mxh[q] = max(mxh[q], sq_hi(orr, oi));
        outv[q][kk] = make_s<S>(orr, oi);
      }
    }
  }
}

// Wave w of the workgroup computes the output blocks k = w + W g of a step
// (at most K of them per lane).  Output c' sums E[c] over the kernel
// combinations t of p = p0(c') ^ combos[t] = (c, y), each term weighted by
// f_y and negated when y = 1 and c . beta is odd; the packed plan orders the
// combinations so that each output takes its weights once (split and plain
// steps above).  Indices fit 32 bits (rank <= 13).  Every read of the step's
// input table ends at the barrier that also publishes the waves' maxima, so
// the outputs overwrite the table in place.  W waves per proposal keep the
// LDS per proposal while giving each CU W times the wavefronts to hide LDS
// and FP64 latency, and need W times fewer output registers per lane.
// Outputs of a launch whose plan exceeds the kernel's rank: NaN values and
// the smallest exponent, written by the first thread of each workgroup over
// a grid stride (the host reads NaN as "plan rank above the rmax passed").
__device__ inline void reject_outputs(int count, double2* out_val, int32_t* out_exp) {
  if (threadIdx.x != 0) return;
  for (int i = blockIdx.x; i < count; i += gridDim.x) {
    out_val[i] = make_double2(__builtin_nan(""), __builtin_nan(""));
    out_exp[i] = INT32_MIN;
  }
}

constexpr int kStepHdrWords = 8;     // per step: header (2), beta, offset, 4 combinations

// The tables are renormalised by a power of two on every kRenormSteps-th
// step of rank 8 or more (and on every smaller step and the last); between
// those the outputs keep their scale.  A power of two changes no mantissa
// (sums are FP64, the table complex64 or complex128, and a step changes
// magnitudes by a few bits at most, far from the formats' ranges), and the
// last step's renormalisation fixes the returned (value, exponent) pair, so
// the results are the same bits as renormalising every step.
constexpr int kRenormSteps = 8;
constexpr int kMaxPlanSteps = 128;   // packed plans have at most 128 steps

template <typename S, int K, int W, int P, int R>
__global__ void __launch_bounds__(64 * W) k_scan_batch(const uint64_t* __restrict__ plan,
                                                      const double2* __restrict__ u,
                                                      const uint64_t* __restrict__ bits, int count,
                                                      double2* __restrict__ out_val, int32_t* __restrict__ out_exp) {
  static_assert((K & (K - 1)) == 0 && (W & (W - 1)) == 0, "K and W are powers of two");
  static_assert(P * W <= 16, "the waves' maxima hold 16 words");
  constexpr int KB = K > 1 ? __builtin_ctz(K) : 0;   // Gray steps within the wave's blocks
  constexpr int SH = sizeof(S) == 8 ? 3 : 4;   // log2 of an entry's bytes
  constexpr int TN = 1 << R;                   // entries per table
  constexpr int TBYTES = TN * (int)sizeof(S);
  // P tables from LDS byte 0 (so an entry's byte offset is its index
  // shifted, and proposal q's entry is TBYTES further on), the waves' maxima
  // after them
  extern __shared__ __align__(16) unsigned char smem[];
  S* E = reinterpret_cast<S*>(smem);
  uint32_t* wmx = reinterpret_cast<uint32_t*>(smem + (size_t)P * TBYTES);
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);   // uniform: plan reads stay scalar
  const int n = (int)plan[0];
  // the tables' LDS byte address; terms XOR their offsets into it, which
  // equals adding them because each table is aligned to its own size
  const uint32_t lds_base = (uint32_t)(uintptr_t)(LDS_AS unsigned char*)smem;
  if ((int)plan[1] > R || (lds_base & ((2u * TBYTES) - 1u))) {
    // the plan's tables would not fit (or the table is not size-aligned):
    // mark every output, touch no table
    reject_outputs(count, out_val, out_exp);
    return;
  }
  // Each step's header words, staged in LDS once per workgroup (after the
  // tables and the maxima), so a step starts from an LDS read instead of a
  // chain of dependent global reads (offset, then header, then combinations).
  uint32_t* shdr = wmx + 16;
  for (int i = threadIdx.x; i < n; i += 64 * W) {
    const uint64_t off = plan[2 + i];
    const uint64_t* st = plan + off;
    const uint64_t hdr = st[0];
    const int r_out = (int)((hdr >> 16) & 255u), nker = (int)((hdr >> 24) & 255u);
    const uint64_t* cb = st + 66 + (1ull << (r_out > 6 ? r_out - 6 : 0));
    uint32_t* h = shdr + kStepHdrWords * i;
    h[0] = (uint32_t)hdr;
    h[1] = (uint32_t)(hdr >> 32);
    h[2] = (uint32_t)st[1];
    h[3] = (uint32_t)off;
    for (int t = 0; t < 4; ++t) h[4 + t] = t < (1 << nker) ? (uint32_t)cb[t] : 0u;
  }
  // P proposals at a time; past the end of the batch a workgroup repeats
  // its first proposal and discards the copy
  for (int pb = blockIdx.x * P; pb < count; pb += gridDim.x * P) {
    uint64_t x0[P], x1[P];
#pragma unroll
    for (int q = 0; q < P; ++q) {
      const int pq = pb + q < count ? pb + q : pb;
      x0[q] = bits[2 * pq];
      x1[q] = bits[2 * pq + 1];
    }
    if (threadIdx.x < P) E[threadIdx.x * TN] = make_s<S>(1.0, 0.0);
    __syncthreads();
    int esum[P], e_prev[P];
#pragma unroll
    for (int q = 0; q < P; ++q) esum[q] = e_prev[q] = 0;
    // The tables hold each step's outputs unscaled; the power of two that
    // brings the step's largest |output| into [1/2, 1) is applied to the
    // next step's weights instead (exact in FP64, so every sum rounds as if
    // the table had been scaled).
    for (int j = 0; j < n; ++j) {
      const uint32_t* h = shdr + kStepHdrWords * j;
      const uint32_t hlo = __builtin_amdgcn_readfirstlane(h[0]), hhi = __builtin_amdgcn_readfirstlane(h[1]);
      const uint64_t hdr = (uint64_t)hlo | ((uint64_t)hhi << 32);
      const uint64_t* st = plan + __builtin_amdgcn_readfirstlane(h[3]);
      uint32_t combos[4];
#pragma unroll
      for (int t = 0; t < 4; ++t) combos[t] = __builtin_amdgcn_readfirstlane(h[4 + t]);
      const int v = (int)(hdr & 255u), r_out = (int)((hdr >> 16) & 255u);
      const bool split = (hdr >> 32) & 1u;
      ScanStepArgs a;
      a.r_in = (int)((hdr >> 8) & 255u);
      a.nker = (int)((hdr >> 24) & 255u);
      a.beta = __builtin_amdgcn_readfirstlane(h[2]);
      a.mask = (1u << a.r_in) - 1u;
      double2 f0[P], f1[P];
#pragma unroll
      for (int q = 0; q < P; ++q) {
        const int xv = (int)((v < 64 ? (x0[q] >> v) : (x1[q] >> (v - 64))) & 1u);
        f0[q] = u[(v * 2 + xv) * 2 + 0];
        f1[q] = u[(v * 2 + xv) * 2 + 1];
        f0[q].x = ldexp(f0[q].x, -e_prev[q]);
        f0[q].y = ldexp(f0[q].y, -e_prev[q]);
        f1[q].x = ldexp(f1[q].x, -e_prev[q]);
        f1[q].y = ldexp(f1[q].y, -e_prev[q]);
      }
      const int nk = r_out > 6 ? 1 << (r_out - 6) : 1;
      const int nkw = nk >= W ? nk / W : (wave < nk ? 1 : 0);
      const bool lane_ok = r_out >= 6 || lane < (1 << r_out);
      // renormalise on every kRenormSteps-th step, on small steps and on the
      // last (see kRenormSteps)
      const bool renorm = r_out < 8 || j % kRenormSteps == kRenormSteps - 1 || j == n - 1;
      // The end of a step: the waves' maxima, the barrier that ends every
      // read of the tables, then the outputs written back in place.  Each
      // kind of step below calls it with its own outputs, so the compiler
      // does not merge the output registers of the different kinds (a
      // register copy per output at every step).  Lanes past 2^r_out (steps
      // of rank below six) computed in-range values that are neither counted
      // nor stored.
      auto finish = [&](auto& outv, auto& mxh, auto track) {
        constexpr bool T = decltype(track)::value;
        if (T) {
#pragma unroll
          for (int q = 0; q < P; ++q) {
            mxh[q] = wave_max_u32(lane_ok ? mxh[q] : 0u);
            if (W > 1 && lane == 0) wmx[q * W + wave] = mxh[q];
          }
        }
        __syncthreads();  // every read of the tables done; the waves' maxima visible
#pragma unroll
        for (int q = 0; q < P; ++q) {
          int e = 0;
          if (T) {
            if (W > 1) {
#pragma unroll
              for (int w2 = 0; w2 < W; ++w2) mxh[q] = max(mxh[q], wmx[q * W + w2]);
            }
            e = scale_exponent(mxh[q]);
          }
          if (lane_ok) {
#pragma unroll
            for (int kk = 0; kk < K; ++kk) {
              if (kk < nkw) E[q * TN + lane + 64 * (wave + W * (kk ^ (kk >> 1)))] = outv[q][kk];
            }
          }
          esum[q] += e;
          e_prev[q] = e;
        }
        __syncthreads();
      };
      // the packed plan has nker <= 2, and split steps nker >= 1
      if ((hdr >> 33) & 1u) {
        uint32_t hstep[KB > 0 ? KB : 1];
#pragma unroll
        for (int b = 0; b < KB; ++b) hstep[b] = (W << b) < nk ? (uint32_t)st[66 + (W << b)] : 0u;
        uint32_t lm8 = (((uint32_t)st[2 + lane] & a.mask) << SH) + lds_base;
        asm volatile("" : "+v"(lm8));   // keep the shifted word: terms then need one XOR each
        const uint32_t hw = wave < nk ? (uint32_t)st[66 + wave] : 0u;
        auto run = [&](auto nker_c, auto split_c, auto track) {
          constexpr int NK = decltype(nker_c)::value;
          constexpr bool SP = decltype(split_c)::value;
          constexpr bool T = decltype(track)::value;
          S outv[P][K];
          uint32_t mxh[P];
          scan_outputs_u<S, K, NK, SP, P, TBYTES, T>(a, f0, f1, combos, lm8, hw, hstep, nkw, outv, mxh);
          finish(outv, mxh, track);
        };
        using I0 = std::integral_constant<int, 0>;
        using I1 = std::integral_constant<int, 1>;
        using I2 = std::integral_constant<int, 2>;
        auto by_kind = [&](auto track) {
          if (split) {
            if (a.nker == 1)
              run(I1(), std::true_type(), track);
            else
              run(I2(), std::true_type(), track);
          } else {
            if (a.nker == 0)
              run(I0(), std::false_type(), track);
            else if (a.nker == 1)
              run(I1(), std::false_type(), track);
            else
              run(I2(), std::false_type(), track);
          }
        };
        if (renorm)
          by_kind(std::true_type());
        else
          by_kind(std::false_type());
      } else {
        // steps without the uniform flag have r_out < 8 (always renormalised):
        // at most two blocks, so at most one per wave (nkw <= 1); each
        // proposal in turn
        const uint32_t lo_w = (uint32_t)st[2 + lane] ^ (wave < nk ? (uint32_t)st[66 + wave] : 0u);
        const uint32_t hstep1[1] = {0u};
        S outv[P][K];
        uint32_t mxh[P];
#pragma unroll
        for (int q = 0; q < P; ++q) {
          ScanStepArgs aq = a;
          aq.f0 = f0[q];
          aq.f1 = f1[q];
          S (&o1)[1] = *reinterpret_cast<S(*)[1]>(&outv[q][0]);
          const S* Eq = E + q * TN;
          if (split) {
            mxh[q] = a.nker == 1 ? scan_outputs<S, 1, 1, true>(Eq, aq, combos, lo_w, hstep1, nkw, o1)
                                 : scan_outputs<S, 1, 2, true>(Eq, aq, combos, lo_w, hstep1, nkw, o1);
          } else {
            mxh[q] = a.nker == 0   ? scan_outputs<S, 1, 0, false>(Eq, aq, combos, lo_w, hstep1, nkw, o1)
                     : a.nker == 1 ? scan_outputs<S, 1, 1, false>(Eq, aq, combos, lo_w, hstep1, nkw, o1)
                                   : scan_outputs<S, 1, 2, false>(Eq, aq, combos, lo_w, hstep1, nkw, o1);
          }
        }
        finish(outv, mxh, std::true_type());
      }
    }
    if (threadIdx.x < P && pb + (int)threadIdx.x < count) {
      const int q = threadIdx.x;
      int ep = e_prev[0], es = esum[0];
#pragma unroll
      for (int q2 = 1; q2 < P; ++q2)
        if (q == q2) {
          ep = e_prev[q2];
          es = esum[q2];
        }
      double2 v = load_c(E[q * TN]);
      out_val[pb + q] = make_double2(ldexp(v.x, -ep), ldexp(v.y, -ep));
      out_exp[pb + q] = es;
    }
    __syncthreads();
  }
}

template <typename S, int K, int W, int P, int R>
static hipError_t launch_scan(const uint64_t* plan, const double2* u, const uint64_t* bits, int count,
                                double2* val, int32_t* ex, hipStream_t st) {
  // the tables, the waves' maxima (64 bytes), then the step headers
  size_t shmem = (size_t)P * ((size_t)1 << R) * sizeof(S) + 64 + kStepHdrWords * 4 * kMaxPlanSteps;
  // at least two rounds of the workgroups that fit a CU at once (160 KiB of
  // LDS, 32 waves), and at least 16 per CU (even shares of large batches)
  const int per_cu = std::max(1, std::min((int)((160 * 1024) / shmem), 32 / W));
  const int groups = (count + P - 1) / P;
  const int grid = std::min(groups, 256 * std::max(16, 2 * per_cu));
  static bool attr = false;
  if (!attr) {
    // dynamic LDS up to the CU's 160 KiB less the kernel's own static array
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_scan_batch<S, K, W, P, R>),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024 - 256);
    if (e != hipSuccess) return e;
    attr = true;
  }
  hipLaunchKernelGGL((k_scan_batch<S, K, W, P, R>), dim3(grid), dim3(64 * W), shmem, st, plan, u, bits, count, val,
                     ex);
  return hipGetLastError();
}

// Wavefronts per proposal for each table precision (the fastest measured on
// 65,536 proposals of a 70-vertex plan: four for complex64 tables, four for
// complex128 tables up to rank 11 and eight at rank 12, whose 64 KiB table
// leaves room for two workgroups per CU); ranks with fewer 64-entry output
// blocks than that use one wave per block.
constexpr int kWavesC64 = 4, kWavesC128 = 4, kWavesC128R12 = 8;

// Proposals per workgroup: the plan-only index and sign work of an output is
// shared by this many proposals, their tables side by side in LDS.  Two for
// complex64 tables up to rank 11 (rank 11: 5.15 -> 4.63 ms per 65,536
// proposals); one at rank 12 and for complex128 tables, where the second
// table and its registers cost more occupancy than the sharing saves.
constexpr int kPropsC64 = 2;

// Tables of rank up to R: 2^(R-6) output blocks of 64 entries, shared by
// min(W, blocks) waves with K blocks each.
template <typename S, int R, int W, int P = 1>
static hipError_t launch_rank(const uint64_t* plan, const double2* u, const uint64_t* bits, int count, double2* val,
                              int32_t* ex, hipStream_t st) {
  constexpr int NB = R > 6 ? 1 << (R - 6) : 1;
  constexpr int WE = W < NB ? W : NB;
  return launch_scan<S, NB / WE, WE, P, R>(plan, u, bits, count, val, ex, st);
}

// ---------------------------------------------------------------- rank <= 6
// Plans of rank at most six have one output block of at most 64 entries per
// step.  A step of the general kernel spends one full 64-lane pass per
// proposal, though a step has only 2^r_out outputs (26 on average for a
// rank-6 plan of 34 vertices), and at this size the pass is bound by
// instruction issue (FP64 sums on every lane).  Here a wavefront owns P
// proposals at once and gives each a slot of G = max(2^r_out, 64 / P) lanes,
// so a step of 2^r_out <= 64 / P outputs serves all P proposals in one pass
// (64 / G proposals per pass, P G / 64 passes).  Each proposal's exponent is
// the maximum over its own slot, and every output is summed by scan_outputs
// with K = 1, so values and exponents are those of the general kernel bit for
// bit.  Each workgroup first copies every step's constants into LDS (header,
// split flag, beta, the kernel combinations, the 64 lane words with the
// step's block word folded in, and both weight pairs of the step's vertex);
// after that one barrier its PW wavefronts run independently.
constexpr int kSmallMaxSteps = 128;   // packed plans have at most 128 steps
struct SmallStep {
  uint32_t hdr, split, beta, pad;
  uint32_t combos[4];
  uint32_t lo[64];
};
constexpr int kSmallStepWords = (int)(sizeof(SmallStep) / 4);
static_assert(kSmallStepWords == 72, "SmallStep layout");
constexpr size_t kSmallPlanBytes = (size_t)kSmallMaxSteps * (sizeof(SmallStep) + 4 * sizeof(double2));
constexpr int kSmallP = 4;   // proposals per wavefront

// LDS of one wavefront: P tables of 64 entries, the P proposals' bit words,
// and their exponent sums.
template <typename S, int P>
constexpr size_t small_wave_bytes() {
  return (size_t)P * 64 * sizeof(S) + (size_t)P * 16 + 16;
}

// Orders this wavefront's LDS accesses across its lanes (a lane reads what
// another lane of the same wavefront wrote): no workgroup barrier is needed
// because a wavefront's LDS operations complete in program order.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Maximum over aligned segments of 2^lg lanes, lg in 4..6 (uniform): DPP
// within each row of 16 lanes, then lane swaps across rows.  All 64 lanes
// must be active.
__device__ __forceinline__ uint32_t segment_max_u32(uint32_t v, int lg) {
  v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0xB1, 0xF, 0xF, false));    // quad_perm 1,0,3,2
  v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x4E, 0xF, 0xF, false));    // quad_perm 2,3,0,1
  v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x124, 0xF, 0xF, false));   // row_ror 4
  v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x128, 0xF, 0xF, false));   // row_ror 8
  if (lg >= 5) v = max(v, (uint32_t)__shfl_xor((int)v, 16));
  if (lg >= 6) v = max(v, (uint32_t)__shfl_xor((int)v, 32));
  return v;
}

template <typename S, int PW, int P>
__global__ void __launch_bounds__(64 * PW) k_scan_small(const uint64_t* __restrict__ plan,
                                                       const double2* __restrict__ u,
                                                       const uint64_t* __restrict__ bits, int count,
                                                       double2* __restrict__ out_val, int32_t* __restrict__ out_exp) {
  static_assert(P == 1 || P == 2 || P == 4, "P divides the wavefront into slots of at least 16 lanes");
  constexpr int LOG2P = P == 4 ? 2 : P == 2 ? 1 : 0;
  extern __shared__ __align__(16) unsigned char smem[];
  SmallStep* ps = reinterpret_cast<SmallStep*>(smem);
  uint32_t* psw = reinterpret_cast<uint32_t*>(smem);
  double2* pf = reinterpret_cast<double2*>(smem + (size_t)kSmallMaxSteps * sizeof(SmallStep));
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  unsigned char* wb = smem + kSmallPlanBytes + (size_t)wave * small_wave_bytes<S, P>();
  S* E = reinterpret_cast<S*>(wb);                                            // [P][64]
  uint64_t* xw = reinterpret_cast<uint64_t*>(wb + (size_t)P * 64 * sizeof(S));  // [P][2]
  int* es = reinterpret_cast<int*>(wb + (size_t)P * 64 * sizeof(S) + (size_t)P * 16);   // [P]
  const int n = (int)plan[0];
  if ((int)plan[1] > 6) {   // this kernel's tables hold rank <= 6
    reject_outputs(count, out_val, out_exp);
    return;
  }
  // Step words: the plan layout at r_out <= 6 is header, beta, 64 lane words,
  // one block word, then 2^nker combinations.
  for (int i = threadIdx.x; i < n * kSmallStepWords; i += 64 * PW) {
    const int j = i / kSmallStepWords, k = i - j * kSmallStepWords;
    const uint64_t* st = plan + plan[2 + j];
    const uint64_t hdr = st[0];
    uint32_t w = 0;
    if (k == 0) w = (uint32_t)hdr;
    else if (k == 1) w = (uint32_t)(hdr >> 32) & 1u;
    else if (k == 2) w = (uint32_t)st[1];
    else if (k >= 4 && k < 8) w = (k - 4) < (1 << ((hdr >> 24) & 255u)) ? (uint32_t)st[67 + (k - 4)] : 0u;
    else if (k >= 8) w = (uint32_t)st[2 + (k - 8)] ^ (uint32_t)st[66];
    psw[i] = w;
  }
  // weights u[v][xv][0..1] of the step's vertex, both values of the bit
  for (int i = threadIdx.x; i < n * 4; i += 64 * PW) {
    const int v = (int)(plan[plan[2 + (i >> 2)]] & 255u);
    pf[i] = u[v * 4 + (i & 3)];
  }
  __syncthreads();
  for (int g = blockIdx.x * PW + wave; g * P < count; g += gridDim.x * PW) {
    const int pbase = g * P;
    if (lane < P) {
      const bool in = pbase + lane < count;   // slots past the batch run on zero bits, unstored
      xw[2 * lane] = in ? bits[2 * (size_t)(pbase + lane)] : 0ull;
      xw[2 * lane + 1] = in ? bits[2 * (size_t)(pbase + lane) + 1] : 0ull;
      es[lane] = 0;
      E[64 * lane] = make_s<S>(1.0, 0.0);
    }
    wave_lds_sync();
    for (int j = 0; j < n; ++j) {
      const SmallStep& sp = ps[j];
      const uint32_t hdr = __builtin_amdgcn_readfirstlane(sp.hdr);
      const bool split = __builtin_amdgcn_readfirstlane(sp.split) != 0u;
      const int v = (int)(hdr & 255u), r_out = (int)((hdr >> 16) & 255u);
      ScanStepArgs a;
      a.r_in = (int)((hdr >> 8) & 255u);
      a.nker = (int)((hdr >> 24) & 255u);
      a.beta = __builtin_amdgcn_readfirstlane(sp.beta);
      a.mask = (1u << a.r_in) - 1u;
      const int lg = r_out > 6 - LOG2P ? r_out : 6 - LOG2P;   // log2 of the slot width G
      const int passes = P >> (6 - lg);
      const int out = lane & ((1 << lg) - 1);
      const bool lane_ok = out < (1 << r_out);
      const uint32_t lo_w = sp.lo[out];
      for (int k = 0; k < passes; ++k) {
        const int p = (k << (6 - lg)) + (lane >> lg);
        const uint64_t x0 = xw[2 * p], x1 = xw[2 * p + 1];
        const int xv = (int)((v < 64 ? (x0 >> v) : (x1 >> (v - 64))) & 1u);
        a.f0 = pf[4 * j + 2 * xv];
        a.f1 = pf[4 * j + 2 * xv + 1];
        S* Ep = E + 64 * p;
        const uint32_t hstep[1] = {0u};
        S outv[1];
        uint32_t mxh;
        if (split) {
          mxh = a.nker == 1 ? scan_outputs<S, 1, 1, true>(Ep, a, sp.combos, lo_w, hstep, 1, outv)
                            : scan_outputs<S, 1, 2, true>(Ep, a, sp.combos, lo_w, hstep, 1, outv);
        } else {
          mxh = a.nker == 0   ? scan_outputs<S, 1, 0, false>(Ep, a, sp.combos, lo_w, hstep, 1, outv)
                : a.nker == 1 ? scan_outputs<S, 1, 1, false>(Ep, a, sp.combos, lo_w, hstep, 1, outv)
                              : scan_outputs<S, 1, 2, false>(Ep, a, sp.combos, lo_w, hstep, 1, outv);
        }
        mxh = segment_max_u32(lane_ok ? mxh : 0u, lg);
        const int e = scale_exponent(mxh);
        wave_lds_sync();   // every lane's reads of the pass's inputs precede the writes
        if (lane_ok) {
          S val = outv[0];
          val.x = ldexp(val.x, -e);
          val.y = ldexp(val.y, -e);
          Ep[out] = val;
        }
        if (out == 0) es[p] += e;
        wave_lds_sync();
      }
    }
    if (lane < P && pbase + lane < count) {
      out_val[pbase + lane] = load_c(E[64 * lane]);
      out_exp[pbase + lane] = es[lane];
    }
    wave_lds_sync();
  }
}

// Wavefronts per workgroup: the workgroup's plan copy is fixed (128 steps),
// so wider workgroups spread it over more proposals; two workgroups fit a CU
// at each precision (complex64 at 16 waves, complex128 at 8).
template <typename S>
static hipError_t launch_small(const uint64_t* plan, const double2* u, const uint64_t* bits, int count, double2* val,
                               int32_t* ex, hipStream_t st) {
  constexpr int PW = sizeof(S) == 8 ? 16 : 8, P = kSmallP;
  const size_t shmem = kSmallPlanBytes + (size_t)PW * small_wave_bytes<S, P>();
  static bool attr = false;
  if (!attr) {
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_scan_small<S, PW, P>),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)shmem);
    if (e != hipSuccess) return e;
    attr = true;
  }
  // one round of the workgroups that fit a CU at once (LDS bound), each
  // copying the plan once for its share of the proposals
  const int per_cu = std::max(1, (int)((160 * 1024) / shmem));
  const int grid = std::min((count + PW * P - 1) / (PW * P), 256 * per_cu);
  hipLaunchKernelGGL((k_scan_small<S, PW, P>), dim3(grid), dim3(64 * PW), shmem, st, plan, u, bits, count, val, ex);
  return hipGetLastError();
}

// ---------------------------------------------------------------- large rank
// One proposal at a time with the environment tables in HBM (rank up to 32):
// one launch per step over the 2^r_out outputs.  The power-of-two scale of
// step j (from the maximum |out|^2 of step j, kept as the bit pattern of a
// non-negative double so that an integer atomicMax orders it) is applied when
// step j+1 reads its input, so no step waits for the host.

__device__ __forceinline__ double inscale_of(const unsigned long long* mx) {
  const double m = __longlong_as_double((long long)*mx);
  int e = 0;
  if (m > 0.0) frexp(sqrt(m), &e);
  return ldexp(1.0, -e);
}

template <typename S>
__global__ void k_scan_init(S* e0, unsigned long long* mx, int n) {
  const int t = threadIdx.x;
  if (t == 0) e0[0] = make_s<S>(1.0, 0.0);
  for (int j = t; j < n; j += blockDim.x) mx[j] = 0ull;
}

// One step over the 2^r_out outputs.  Output c' sums E[c] over the kernel
// combinations of p = p0(c') ^ combos[t] = (c, y), weighted by f_y and
// negated when y = 1 and c . beta is odd.  When d != 0 (the twin of c' is
// c' ^ d, whose terms are the same entries with y flipped, see twin_of), a
// thread takes an output and its twin: it reads each entry once and weights
// it into both, halving the step's reads.
template <typename S>
__global__ void __launch_bounds__(256) k_scan_step(const uint64_t* __restrict__ st, const S* __restrict__ ein,
                                                   S* __restrict__ eout, double2 f0, double2 f1,
                                                   const unsigned long long* __restrict__ mx_prev,
                                                   unsigned long long* __restrict__ mx_out, uint32_t d) {
  __shared__ uint64_t tab[4 * 256];
  __shared__ double red[4];
  const uint64_t hdr = st[0];
  const int r_in = (int)((hdr >> 8) & 255u), r_out = (int)((hdr >> 16) & 255u);
  const int nker = (int)((hdr >> 24) & 255u), nbytes = (int)((hdr >> 32) & 255u);
  const uint64_t beta = st[1];
  for (int i = threadIdx.x; i < nbytes * 256; i += blockDim.x) tab[i] = st[2 + i];
  __syncthreads();
  const uint64_t* combos = st + 2 + 256ull * nbytes;
  const int ncomb = 1 << nker;
  const double insc = mx_prev ? inscale_of(mx_prev) : 1.0;
  const uint64_t mask = (1ull << r_in) - 1ull;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  double lmax = 0.0;
  if (d == 0) {
    const uint64_t total = 1ull << r_out;
    for (uint64_t cp = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; cp < total; cp += stride) {
      uint64_t p0 = 0;
      for (int b = 0; b < nbytes; ++b) p0 ^= tab[(b << 8) | ((cp >> (8 * b)) & 255u)];
      double ar = 0.0, ai = 0.0;
      for (int t = 0; t < ncomb; ++t) {
        const uint64_t p = p0 ^ combos[t];
        const uint64_t c = p & mask;
        const uint32_t y = (uint32_t)(p >> r_in) & 1u;
        const uint32_t neg = y & ((uint32_t)__popcll(c & beta) & 1u);
        const double2 w = y ? f1 : f0;
        const double2 e = load_c(ein[c]);
        const double pr = e.x * w.x - e.y * w.y;
        const double pi = e.x * w.y + e.y * w.x;
        if (neg) {
          ar -= pr;
          ai -= pi;
        } else {
          ar += pr;
          ai += pi;
        }
      }
      ar *= insc;
      ai *= insc;
      lmax = fmax(lmax, ar * ar + ai * ai);
      eout[cp] = make_s<S>(ar, ai);
    }
  } else {
    // outputs c' with bit hb of d clear (hb its highest set bit) and their twins
    const int hb = 31 - __clz(d);
    const uint64_t low = (1ull << hb) - 1ull, half = 1ull << (r_out - 1);
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < half; i += stride) {
      const uint64_t cp = ((i >> hb) << (hb + 1)) | (i & low);
      uint64_t p0 = 0;
      for (int b = 0; b < nbytes; ++b) p0 ^= tab[(b << 8) | ((cp >> (8 * b)) & 255u)];
      double ar = 0.0, ai = 0.0, br = 0.0, bi = 0.0;
      for (int t = 0; t < ncomb; ++t) {
        const uint64_t p = p0 ^ combos[t];
        const uint64_t c = p & mask;
        const uint32_t y = (uint32_t)(p >> r_in) & 1u;
        const uint32_t par = (uint32_t)__popcll(c & beta) & 1u;
        const double2 e = load_c(ein[c]);
        const double2 wa = y ? f1 : f0, wb = y ? f0 : f1;   // the twin sees the entry with y flipped
        const double pra = e.x * wa.x - e.y * wa.y, pia = e.x * wa.y + e.y * wa.x;
        const double prb = e.x * wb.x - e.y * wb.y, pib = e.x * wb.y + e.y * wb.x;
        if (y & par) {
          ar -= pra;
          ai -= pia;
        } else {
          ar += pra;
          ai += pia;
        }
        if ((y ^ 1u) & par) {
          br -= prb;
          bi -= pib;
        } else {
          br += prb;
          bi += pib;
        }
      }
      ar *= insc;
      ai *= insc;
      br *= insc;
      bi *= insc;
      lmax = fmax(lmax, fmax(ar * ar + ai * ai, br * br + bi * bi));
      eout[cp] = make_s<S>(ar, ai);
      eout[cp ^ d] = make_s<S>(br, bi);
    }
  }
  for (int off = 32; off > 0; off >>= 1) lmax = fmax(lmax, __shfl_xor(lmax, off));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = lmax;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
    if (m > 0.0) atomicMax(mx_out, (unsigned long long)__double_as_longlong(m));
  }
}

// value = E_last[0] * 2^-e_last, exponent = sum_j e_j.
template <typename S>
__global__ void k_scan_finish(const S* elast, const unsigned long long* mx, int n, double2* out_val,
                              int32_t* out_exp) {
  if (threadIdx.x != 0) return;
  int esum = 0;
  for (int j = 0; j < n; ++j) {
    const double m = __longlong_as_double((long long)mx[j]);
    int e = 0;
    if (m > 0.0) frexp(sqrt(m), &e);
    esum += e;
  }
  const double sc = inscale_of(mx + (n - 1));
  const double2 v = load_c(elast[0]);
  out_val[0] = make_double2(v.x * sc, v.y * sc);
  out_exp[0] = esum;
}

// The twin offset of a step: p0(c' ^ d) = p0(c') ^ e_y ^ k with k in the
// kernel, so the terms of output c' ^ d are those of c' with y flipped.  The
// lcols and the kernel basis together form a basis of p-space (r_in + 1
// vectors), so e_y = L d ^ K z has exactly one solution; d is its L part
// (0 when e_y lies in the kernel, and then outputs have no twin).
static uint32_t twin_of(const rgs_scan_step& st) {
  const int m = st.r_out + st.nker;
  if (st.r_out < 1 || m > 64) return 0;
  uint64_t vec[64], comb[64];
  int piv[64], nb = 0;
  for (int i = 0; i < m; ++i) {
    uint64_t v = i < st.r_out ? st.lcols[i] : st.kern[i - st.r_out];
    uint64_t c = 1ull << i;
    for (int k = 0; k < nb; ++k)
      if ((v >> piv[k]) & 1ull) {
        v ^= vec[k];
        c ^= comb[k];
      }
    if (v) {
      vec[nb] = v;
      comb[nb] = c;
      piv[nb] = 63 - __builtin_clzll(v);
      ++nb;
    }
  }
  uint64_t t = 1ull << st.r_in, tc = 0;
  for (int k = 0; k < nb; ++k)
    if ((t >> piv[k]) & 1ull) {
      t ^= vec[k];
      tc ^= comb[k];
    }
  if (t) return 0;   // not in the span: leave the step to the one-output path
  return (uint32_t)(tc & ((st.r_out >= 32 ? 0ull : (1ull << st.r_out)) - 1ull));
}

template <typename S>
static int scan_large(const rgs_scan_step* steps, int n, const uint64_t* plan_host_offsets, const uint64_t* plan,
                      const double2* u, const uint64_t* bits, void* buf0, void* buf1,
                      unsigned long long* mx, double2* out_val, int32_t* out_exp, hipStream_t st) {
  S* b[2] = {reinterpret_cast<S*>(buf0), reinterpret_cast<S*>(buf1)};
  hipLaunchKernelGGL((k_scan_init<S>), dim3(1), dim3(128), 0, st, b[0], mx, n);
  RGS_LAUNCH_CHECK("k_scan_init");
  for (int j = 0; j < n; ++j) {
    const rgs_scan_step& s = steps[j];
    const int v = s.v;
    const int xv = (int)((v < 64 ? (bits[0] >> v) : (bits[1] >> (v - 64))) & 1u);
    const double2 f0 = u[(v * 2 + xv) * 2 + 0], f1 = u[(v * 2 + xv) * 2 + 1];
    const uint32_t d = twin_of(s);
    const uint64_t outs = d ? 1ull << (s.r_out - 1) : 1ull << s.r_out;   // threads: output pairs or outputs
    const uint64_t blocks = std::min<uint64_t>((outs + 255) / 256, 256ull * 64);
    hipLaunchKernelGGL((k_scan_step<S>), dim3((unsigned)blocks), dim3(256), 0, st, plan + plan_host_offsets[j],
                       b[j & 1], b[(j + 1) & 1], f0, f1, j ? mx + (j - 1) : nullptr, mx + j, d);
    RGS_LAUNCH_CHECK("k_scan_step");
  }
  hipLaunchKernelGGL((k_scan_finish<S>), dim3(1), dim3(64), 0, st, b[n & 1], mx, n, out_val, out_exp);
  RGS_LAUNCH_CHECK("k_scan_finish");
  return kOk;
}

// ------------------------------------------------------------ decisions
// Capped-rejection decisions from the batch scan's outputs, one proposal per
// lane: p = (re^2 + im^2) 2^(2 e - n), acc = min(p * scale, 1) (uniform
// proposals, scale = 2^n / cap) or min(p / (scale * exp(log q)), 1) (scale =
// cap), keep = u < acc.  Every product, sum and quotient is rounded on its
// own (no fused multiply-add), as the same expressions over device tensors
// evaluate them; the power of two is exact.  Each wavefront adds its kept
// count to *accepted with one atomic.
__global__ void __launch_bounds__(256) k_capped_decide(const double2* __restrict__ val,
                                                       const int32_t* __restrict__ ex, const double* __restrict__ u,
                                                       const double* __restrict__ logq, int count, int n,
                                                       double scale, uint8_t* __restrict__ keep,
                                                       unsigned long long* __restrict__ accepted) {
  // hipcc contracts a * b + c into one fused multiply-add by default (also
  // through __dmul_rn / __dadd_rn once inlined); the decision must round the
  // squares and their sum separately
#pragma clang fp contract(off)
  const int i = blockIdx.x * 256 + threadIdx.x;
  bool k = false;
  if (i < count) {
    const double2 v = val[i];
    const double s = v.x * v.x + v.y * v.y;
    const double p = ldexp(s, 2 * ex[i] - n);
    const double acc = logq ? p / (scale * exp(logq[i])) : p * scale;
    k = u[i] < fmin(acc, 1.0);
    keep[i] = k ? 1 : 0;
  }
  const unsigned long long b = __ballot(k);
  if ((threadIdx.x & 63) == 0 && b) atomicAdd(accepted, (unsigned long long)__popcll(b));
}

}  // namespace rgs

using namespace rgs;

extern "C" int rgs_scan_batch(const uint64_t* plan_dev, int rmax, const double* u_dev, const uint64_t* bits_dev,
                              int count, int precision, double* out_val, int32_t* out_exp, void* stream) {
  if (!plan_dev || !u_dev || !bits_dev || !out_val || !out_exp) return fail(kInvalid, "scan_batch: null pointer");
  if (count <= 0) return kOk;
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  const double2* u = reinterpret_cast<const double2*>(u_dev);
  double2* val = reinterpret_cast<double2*>(out_val);
  hipError_t e;
  if (precision == 64) {
    if (rmax <= 6) e = launch_small<float2>(plan_dev, u, bits_dev, count, val, out_exp, st);
    else if (rmax <= 8) e = launch_rank<float2, 8, kWavesC64, kPropsC64>(plan_dev, u, bits_dev, count, val, out_exp, st);
    else if (rmax <= 10) e = launch_rank<float2, 10, kWavesC64, kPropsC64>(plan_dev, u, bits_dev, count, val, out_exp, st);
    else if (rmax <= 11) e = launch_rank<float2, 11, kWavesC64, kPropsC64>(plan_dev, u, bits_dev, count, val, out_exp, st);
    else if (rmax <= 12) e = launch_rank<float2, 12, kWavesC64>(plan_dev, u, bits_dev, count, val, out_exp, st);
    else if (rmax <= 13) e = launch_rank<float2, 13, kWavesC64>(plan_dev, u, bits_dev, count, val, out_exp, st);
    else return fail(kInvalid, "scan_batch: complex64 batch limited to rmax <= 13");
  } else if (precision == 128) {
    if (rmax <= 6) e = launch_small<double2>(plan_dev, u, bits_dev, count, val, out_exp, st);
    else if (rmax <= 8) e = launch_rank<double2, 8, kWavesC128>(plan_dev, u, bits_dev, count, val, out_exp, st);
    else if (rmax <= 10) e = launch_rank<double2, 10, kWavesC128>(plan_dev, u, bits_dev, count, val, out_exp, st);
    else if (rmax <= 11) e = launch_rank<double2, 11, kWavesC128>(plan_dev, u, bits_dev, count, val, out_exp, st);
    else if (rmax <= 12) e = launch_rank<double2, 12, kWavesC128R12>(plan_dev, u, bits_dev, count, val, out_exp, st);
    else return fail(kInvalid, "scan_batch: complex128 batch limited to rmax <= 12");
  } else {
    return fail(kInvalid, "precision must be 64 (complex64 tables) or 128 (complex128)");
  }
  if (e != hipSuccess) return fail(kHip, std::string(rmax <= 6 ? "k_scan_small: " : "k_scan_batch: ") + hipGetErrorString(e));
  return kOk;
}

extern "C" int rgs_capped_decide(const double* val, const int32_t* ex, const double* u, const double* logq,
                                 int count, int n, double scale, uint8_t* keep, unsigned long long* accepted,
                                 void* stream) {
  if (!val || !ex || !u || !keep || !accepted || count < 0 || n < 1 || n > 128 || !(scale > 0.0))
    return fail(kInvalid, "capped_decide: bad arguments");
  if (count == 0) return kOk;
  hipLaunchKernelGGL(k_capped_decide, dim3((unsigned)((count + 255) / 256)), dim3(256), 0,
                     reinterpret_cast<hipStream_t>(stream), reinterpret_cast<const double2*>(val), ex, u, logq,
                     count, n, scale, keep, accepted);
  RGS_LAUNCH_CHECK("k_capped_decide");
  return kOk;
}

extern "C" int rgs_scan_large(const rgs_scan_step* steps, int n, const uint64_t* plan_dev,
                              const uint64_t* plan_offsets, const double* u_host, const uint64_t* bits_host,
                              int precision, void* buf0, void* buf1, uint64_t buf_entries, uint64_t* mx_dev,
                              double* out_val, int32_t* out_exp, void* stream) {
  if (!steps || !plan_dev || !plan_offsets || !u_host || !bits_host || !buf0 || !buf1 || !mx_dev || !out_val ||
      !out_exp || n < 1 || n > 128)
    return fail(kInvalid, "scan_large: bad arguments");
  int rmax = 0;
  for (int j = 0; j < n; ++j) {
    if (steps[j].r_out > 32 || steps[j].r_in > 32) return fail(kInvalid, "scan_large: rank above 32");
    rmax = std::max(rmax, std::max(steps[j].r_in, steps[j].r_out));
  }
  if (buf_entries < (1ull << rmax)) return fail(kInvalid, "scan_large: buffers smaller than 2^rmax entries");
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  const double2* u = reinterpret_cast<const double2*>(u_host);
  unsigned long long* mx = reinterpret_cast<unsigned long long*>(mx_dev);
  double2* val = reinterpret_cast<double2*>(out_val);
  if (precision == 64)
    return scan_large<float2>(steps, n, plan_offsets, plan_dev, u, bits_host, buf0, buf1, mx, val, out_exp, st);
  if (precision == 128)
    return scan_large<double2>(steps, n, plan_offsets, plan_dev, u, bits_host, buf0, buf1, mx, val, out_exp, st);
  return fail(kInvalid, "precision must be 64 (complex64 tables) or 128 (complex128)");
}
// ------------------------------------------------- allocation-free variant
// rgs_scan_large_static runs the same kernels with its two tables and the
// step maxima in a fixed device array of this library (one copy per device,
// held until the process ends), for callers that cannot allocate.  Layout:
// the n <= 128 step maxima, then two tables of kStaticTableBytes each.
namespace rgs {

constexpr size_t kStaticMaxBytes = 128 * sizeof(unsigned long long);
constexpr size_t kStaticTableBytes = (size_t)1 << 24;   // 2^20 complex128 or 2^21 complex64 entries
constexpr size_t kStaticBytes = 4096 + 2 * kStaticTableBytes;
static_assert(kStaticMaxBytes <= 4096, "maxima fit the first 4 KiB");

static __device__ __align__(256) unsigned char scratch[kStaticBytes];

constexpr int kMaxDevices = 64;
static std::mutex scratch_lock;
static unsigned char* scratch_ptr[kMaxDevices];   // device address of `scratch` on each device

// The device address of `scratch` on the current device, looked up once per
// device (the array's host-side name is not a device address).
static int scratch_of_current(unsigned char** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return fail(kHip, std::string("scan_large_static: ") + hipGetErrorString(e));
  if (dev < 0 || dev >= kMaxDevices) return fail(kInvalid, "scan_large_static: device index above 63");
  std::lock_guard<std::mutex> g(scratch_lock);
  if (!scratch_ptr[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) return fail(kHip, std::string("scan_large_static: hipGetSymbolAddress: ") +
                                                     hipGetErrorString(e));
    scratch_ptr[dev] = static_cast<unsigned char*>(p);
  }
  *out = scratch_ptr[dev];
  return kOk;
}

}  // namespace rgs

extern "C" int rgs_scan_large_static(const rgs_scan_step* steps, int n, const uint64_t* plan_dev,
                                     const uint64_t* plan_offsets, const double* u_host, const uint64_t* bits_host,
                                     int precision, double* out_val, int32_t* out_exp, void* stream) {
  if (!steps || !plan_dev || !plan_offsets || !u_host || !bits_host || !out_val || !out_exp || n < 1 || n > 128)
    return fail(kInvalid, "scan_large_static: bad arguments");
  if (precision != 64 && precision != 128)
    return fail(kInvalid, "precision must be 64 (complex64 tables) or 128 (complex128)");
  int rmax = 0;
  for (int j = 0; j < n; ++j) rmax = std::max(rmax, std::max(steps[j].r_in, steps[j].r_out));
  const int limit = precision == 64 ? 21 : 20;
  if (rmax > limit)
    return fail(kInvalid, "scan_large_static: rank " + std::to_string(rmax) + " above the fixed tables' limit " +
                              std::to_string(limit) + " (use rgs_scan_large with caller buffers)");
  unsigned char* base = nullptr;
  int rc = scratch_of_current(&base);
  if (rc) return rc;
  unsigned long long* mx = reinterpret_cast<unsigned long long*>(base);
  void* buf0 = base + 4096;
  void* buf1 = base + 4096 + kStaticTableBytes;
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  const double2* u = reinterpret_cast<const double2*>(u_host);
  double2* val = reinterpret_cast<double2*>(out_val);
  if (precision == 64)
    return scan_large<float2>(steps, n, plan_offsets, plan_dev, u, bits_host, buf0, buf1, mx, val, out_exp, st);
  return scan_large<double2>(steps, n, plan_offsets, plan_dev, u, bits_host, buf0, buf1, mx, val, out_exp, st);
}
