// Kernels of the learned proposal: a masked autoregressive MLP (MADE) with
// BF16 weights and activations, FP32 accumulation on the MFMA units, and the
// elementwise pieces of training (loss, bias gradients, masked Adam) and of
// autoregressive sampling.
//
// GEMM: C[M x N] = op(A)[M x K] * op(B)[K x N], op = identity or transpose of
// a row-major BF16 matrix.  A 256-thread workgroup computes a 64 x 64 tile
// with v_mfma_f32_16x16x16_bf16 (each wavefront a 32 x 32 quarter, 2 x 2
// MFMA blocks), staging 64 x 32 slices of A and B through LDS.  Fragment
// layout of the 16x16x16 instruction: lane l supplies A[l % 16][4 (l / 16) + i]
// and B[4 (l / 16) + i][l % 16], i = 0..3, and receives
// D[4 (l / 16) + i][l % 16].  Epilogues: bias, ReLU, a ReLU-derivative mask
// from a saved activation, FP32 and/or BF16 stores, accumulate-into-C.
#include "rgs_common.h"
#include "rgs_amd.h"

#include <algorithm>

namespace rgs {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float bf16_to_f32(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }
// Round to nearest even (inputs here are finite).
__device__ __forceinline__ uint16_t f32_to_bf16(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

struct GemmArgs {
  const uint16_t* A;
  const uint16_t* B;
  float* C;
  uint16_t* Cb;
  const float* bias;
  const uint16_t* dmask;  // multiply by (dmask > 0) (ReLU derivative), or NULL
  int M, N, K, lda, ldb, ldc, ldcb, ldm;
  int ta, tb, relu, accumulate;
  float alpha;
};

constexpr int TM = 64, TN = 64, TK = 32;

// One thread's share of an operand tile: rows [r0, r0 + 64) x columns
// [k0, k0 + TK), zero outside rows x K, 8 elements per thread (TK = 32).  The
// operand is row-major with k contiguous (X[r ld + k], kcontig) or r
// contiguous (X[k ld + r]).  With vec (ld a multiple of 8, X 16-byte aligned,
// and the contiguous extent a multiple of 8) the 8 elements are one 16-byte
// load; otherwise one element per load.  load_tile fills registers (so the
// next tile's loads are in flight while the current one is multiplied),
// store_tile writes them into S[r][k] (k contiguous).
static_assert(64 * TK == 256 * 8, "8 elements per thread");
struct TileRegs {
  uint4 v;
};

__device__ __forceinline__ void load_tile(TileRegs& t, const uint16_t* X, int ld, bool kcontig, int rows, int K,
                                          int r0, int k0, bool vec) {
  const int tid = threadIdx.x;
  t.v = make_uint4(0u, 0u, 0u, 0u);
  if (vec && kcontig) {
    const int rr = tid / (TK / 8), kc = (tid % (TK / 8)) * 8;
    const int gr = r0 + rr, gk = k0 + kc;
    if (gr < rows && gk < K) t.v = *reinterpret_cast<const uint4*>(X + (size_t)gr * ld + gk);
  } else if (vec) {
    const int kk = tid >> 3, rc = (tid & 7) * 8;
    const int gk = k0 + kk, gr = r0 + rc;
    if (gk < K && gr < rows) t.v = *reinterpret_cast<const uint4*>(X + (size_t)gk * ld + gr);
  } else {
    uint32_t w[4] = {0u, 0u, 0u, 0u};
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const int e = tid + 256 * i;
      const int rr = e / TK, kk = e % TK;
      const int gr = r0 + rr, gk = k0 + kk;
      uint32_t v = 0;
      if (gr < rows && gk < K) v = kcontig ? X[(size_t)gr * ld + gk] : X[(size_t)gk * ld + gr];
      w[i >> 1] |= v << (16 * (i & 1));
    }
    t.v = make_uint4(w[0], w[1], w[2], w[3]);
  }
}

__device__ __forceinline__ void store_tile(uint16_t (*S)[TK + 4], const TileRegs& t, bool kcontig, bool vec) {
  const int tid = threadIdx.x;
  const uint32_t w[4] = {t.v.x, t.v.y, t.v.z, t.v.w};
  if (vec && kcontig) {
    const int rr = tid / (TK / 8), kc = (tid % (TK / 8)) * 8;
    // a row of S is 72 bytes: 8-byte aligned, so two 8-byte stores
    uint2* d = reinterpret_cast<uint2*>(&S[rr][kc]);
    d[0] = make_uint2(w[0], w[1]);
    d[1] = make_uint2(w[2], w[3]);
  } else if (vec) {
    const int kk = tid >> 3, rc = (tid & 7) * 8;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      S[rc + 2 * i][kk] = (uint16_t)(w[i] & 0xffffu);
      S[rc + 2 * i + 1][kk] = (uint16_t)(w[i] >> 16);
    }
  } else {
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const int e = tid + 256 * i;
      S[e / TK][e % TK] = (uint16_t)((w[i >> 1] >> (16 * (i & 1))) & 0xffffu);
    }
  }
}

__device__ __forceinline__ bool vec_ok(const uint16_t* X, int ld, int extent) {
  return (ld % 8) == 0 && (extent % 8) == 0 && ((uintptr_t)X % 16) == 0;
}

__global__ void __launch_bounds__(256) k_gemm(GemmArgs g) {
  // As[m][k], Bs[n][k]: k contiguous so a lane's four k values are adjacent.
  __shared__ uint16_t As[TM][TK + 4];
  __shared__ uint16_t Bs[TN][TK + 4];
  const int tid = threadIdx.x;
  const int wave = tid >> 6, lane = tid & 63;
  const int m0 = blockIdx.y * TM, n0 = blockIdx.x * TN;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  f32x4 acc[2][2];
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j) acc[i][j] = (f32x4){0.f, 0.f, 0.f, 0.f};
  // A is M x K (k contiguous unless ta), B is K x N given as N x K (k
  // contiguous) when tb
  const bool akc = !g.ta, bkc = g.tb != 0;
  const bool avec = vec_ok(g.A, g.lda, akc ? g.K : g.M), bvec = vec_ok(g.B, g.ldb, bkc ? g.K : g.N);

  TileRegs ra, rb;
  load_tile(ra, g.A, g.lda, akc, g.M, g.K, m0, 0, avec);
  load_tile(rb, g.B, g.ldb, bkc, g.N, g.K, n0, 0, bvec);
  for (int k0 = 0; k0 < g.K; k0 += TK) {
    store_tile(As, ra, akc, avec);
    store_tile(Bs, rb, bkc, bvec);
    __syncthreads();
    if (k0 + TK < g.K) {   // the next tile's loads overlap this tile's products
      load_tile(ra, g.A, g.lda, akc, g.M, g.K, m0, k0 + TK, avec);
      load_tile(rb, g.B, g.ldb, bkc, g.N, g.K, n0, k0 + TK, bvec);
    }
#pragma unroll
    for (int ks = 0; ks < TK; ks += 16) {
      const int kq = ks + 4 * (lane >> 4);
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        bf16x4 a;
        const int r = wm + 16 * i + (lane & 15);
        a[0] = (short)As[r][kq + 0];
        a[1] = (short)As[r][kq + 1];
        a[2] = (short)As[r][kq + 2];
        a[3] = (short)As[r][kq + 3];
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          bf16x4 b;
          const int cidx = wn + 16 * j + (lane & 15);
          b[0] = (short)Bs[cidx][kq + 0];
          b[1] = (short)Bs[cidx][kq + 1];
          b[2] = (short)Bs[cidx][kq + 2];
          b[3] = (short)Bs[cidx][kq + 3];
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[i][j], 0, 0, 0);
        }
      }
    }
    __syncthreads();
  }
  // epilogue
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const int row = m0 + wm + 16 * i + 4 * (lane >> 4) + t;
        const int col = n0 + wn + 16 * j + (lane & 15);
        if (row < g.M && col < g.N) {
          float v = g.alpha * acc[i][j][t];
          if (g.bias) v += g.bias[col];
          if (g.relu) v = fmaxf(v, 0.f);
          if (g.dmask && !(bf16_to_f32(g.dmask[(size_t)row * g.ldm + col]) > 0.f)) v = 0.f;
          if (g.C) {
            float* cp = g.C + (size_t)row * g.ldc + col;
            *cp = g.accumulate ? *cp + v : v;
          }
          if (g.Cb) g.Cb[(size_t)row * g.ldcb + col] = f32_to_bf16(v);
        }
      }
}

// Loss and its gradient for sigmoid outputs: per row, NLL = sum_o softplus(l)
// - x l (binary cross-entropy in logits), dlogit = (sigmoid(l) - x) * scale.
// One wavefront per row: lane o takes outputs o, o + 64, ...; the row's NLL
// is the wavefront's butterfly sum (a fixed order).
__global__ void __launch_bounds__(256) k_bce(const float* logits, const uint16_t* x, int B, int n, float scale,
                                             float* nll, float* dlog, uint16_t* dlog_b) {
  const int r = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (r >= B) return;   // whole wavefronts: no barrier follows
  float acc = 0.f;
  for (int o = lane; o < n; o += 64) {
    float l = logits[(size_t)r * n + o];
    float xv = bf16_to_f32(x[(size_t)r * n + o]);
    float sp = l > 0.f ? l + log1pf(expf(-l)) : log1pf(expf(l));
    acc += sp - xv * l;
    float s = 1.f / (1.f + expf(-l));
    float d = (s - xv) * scale;
    if (dlog) dlog[(size_t)r * n + o] = d;
    if (dlog_b) dlog_b[(size_t)r * n + o] = f32_to_bf16(d);
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
  if (nll && lane == 0) nll[r] = acc;
}

// out[c] (+)= sum_r in[r][c]  (bias gradients).  A workgroup takes 64
// columns; its 16 row lanes sum rows r = lane, lane + 16, ... and the partial
// sums are added in row-lane order, so the result does not depend on timing.
constexpr int kColRowLanes = 16;

template <typename T>
__device__ __forceinline__ float col_value(const T* p);
template <>
__device__ __forceinline__ float col_value<float>(const float* p) { return *p; }
template <>
__device__ __forceinline__ float col_value<uint16_t>(const uint16_t* p) { return bf16_to_f32(*p); }

template <typename T>
__global__ void __launch_bounds__(64 * kColRowLanes) k_colsum(const T* in, int rows, int cols, int ld, float* out,
                                                              int accumulate) {
  __shared__ float part[kColRowLanes][64];
  const int cl = threadIdx.x & 63, rl = threadIdx.x >> 6;
  const int c = blockIdx.x * 64 + cl;
  float s = 0.f;
  if (c < cols)
    for (int r = rl; r < rows; r += kColRowLanes) s += col_value<T>(in + (size_t)r * ld + c);
  part[rl][cl] = s;
  __syncthreads();
  if (rl == 0 && c < cols) {
    float t = part[0][cl];
    for (int k = 1; k < kColRowLanes; ++k) t += part[k][cl];
    out[c] = accumulate ? out[c] + t : t;
  }
}

__global__ void k_step_counter(int* t) { t[0] += 1; }

// Masked Adam on a flat parameter vector; writes the BF16 working copy.
__global__ void k_adam(float* w, const float* grad, const float* mask, float* m, float* v, uint16_t* wb,
                       long long count, const int* t, float lr, float b1, float b2, float eps) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const float step = (float)t[0];
  float g = grad[i] * mask[i];
  float mi = b1 * m[i] + (1.f - b1) * g;
  float vi = b2 * v[i] + (1.f - b2) * g * g;
  m[i] = mi;
  v[i] = vi;
  float mh = mi / (1.f - powf(b1, step));
  float vh = vi / (1.f - powf(b2, step));
  float wi = (w[i] - lr * mh / (sqrtf(vh) + eps)) * mask[i];
  w[i] = wi;
  wb[i] = f32_to_bf16(wi);
}

// One draw from a float logit l, in FP64: bit = [u < sigmoid(l)] and *term =
// log sigmoid(+-l) of the chosen bit.  With e = exp(-|l|), sigmoid(l) is
// 1 / (1 + e) for l >= 0 and e / (1 + e) otherwise, and log sigmoid(s) =
// min(s, 0) - log1p(e) for s = +-l, so one exponential serves both.  Every
// sampler (per step and fused) draws through this function, so they agree
// bit for bit.
__device__ __forceinline__ int made_bit(float lf, double u, double* term) {
  const double l = (double)lf;
  const double e = exp(-fabs(l));
  const double d = 1.0 + e;
  const double p = l >= 0.0 ? 1.0 / d : e / d;
  const int one = u < p;
  if (term) *term = fmin(one ? l : -l, 0.0) - log1p(e);
  return one;
}

// Autoregressive sampling step: x[r][o] = (u[r][o] < sigmoid(l)) for the
// column o just computed.
__global__ void k_sample_col(const float* logits, const double* u, int B, int n, int o, uint16_t* x) {
  int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= B) return;
  const int one = made_bit(logits[(size_t)r * n + o], u[(size_t)r * n + o], nullptr);
  x[(size_t)r * n + o] = one ? (uint16_t)0x3f80 : (uint16_t)0;  // bf16 1.0 / 0.0
}

// One autoregressive step: x[b][o] = (u[b][o] < sigmoid(l[b])), and log q
// accumulates log sigmoid(+-l[b]) of the chosen bit (in float64, from the
// same float logit that decided the bit).
__global__ void k_sample_step(const float* logit, int ldl, const double* u, int B, int n, int o, uint16_t* x,
                              double* logq, int first) {
  int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= B) return;
  double ls;
  const int one = made_bit(logit[(size_t)r * ldl], u[(size_t)r * n + o], &ls);
  x[(size_t)r * n + o] = one ? (uint16_t)0x3f80 : (uint16_t)0;
  logq[r] = first ? ls : logq[r] + ls;
}

// Four consecutive bf16 of an 8-byte aligned LDS or global address.
__device__ __forceinline__ bf16x4 load4(const uint16_t* p) { return *reinterpret_cast<const bf16x4*>(p); }

// Sampler for the degree-blocked layout: every degree d owns g16 whole
// aligned 16-unit blocks [(d-1) g16, d g16) of each hidden layer, either
// H = 16 g16 (n - 1) or H = 16 k with k < n - 1 (one block for each of the
// degrees 1 .. k, none above; steps past k add no units).  One
// wavefront draws 16 proposals with no workgroup barrier.  Products are taken
// transposed, D[unit][row] = sum_k W[unit][k] v[row][k], so a lane (row
// r = lane % 16, quarter q = lane / 16) receives units 4q .. 4q+3 of its row:
// exactly the B-operand fragment of v_mfma_f32_16x16x16_bf16 for that unit
// block.  First-layer blocks stay in the wave's LDS slice in fragment order
// (one 8-byte entry per lane and block); a second-layer block is used at once
// to add its contribution to the logit blocks L[ob] (outputs 16 ob .. 16 ob
// + 15) and is not kept.  At step o the degree-o blocks are formed (first
// layer over inputs < o, second layer over the blocks of degree <= o), then
// logit o = L[o / 16] + b3[o] decides bit o as k_sample_step does.  Every
// sum runs over ascending 16-wide K blocks into one accumulator with the
// epilogue of k_gemm, as the launch-per-step GEMM sequence does (weights
// outside a unit's mask are exact zeros there); log q accumulates in step
// order.  Weights are read as 8-byte fragments straight from L2.
__device__ __forceinline__ f32x4 mfma16(bf16x4 a, bf16x4 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// Logit o sits in component o % 4 of L[o / 16] in the lanes of quarter
// (o % 16) / 4; those lanes draw bit o as k_sample_step does and the bit and
// its log-probability term are broadcast to the row's four lanes.
// The four logit blocks are separate variables so that no dynamic index
// turns them into a scratch array.
struct LogitBlocks {
  f32x4 b0, b1, b2, b3;
};

__device__ __forceinline__ void made_logits(LogitBlocks& L, int nob, const bf16x4 (&w3)[4], bf16x4 h) {
  L.b0 = mfma16(w3[0], h, L.b0);
  if (nob > 1) L.b1 = mfma16(w3[1], h, L.b1);
  if (nob > 2) L.b2 = mfma16(w3[2], h, L.b2);
  if (nob > 3) L.b3 = mfma16(w3[3], h, L.b3);
}

__device__ __forceinline__ void made_draw(const LogitBlocks& L, const float* __restrict__ b3, int o, int q, int r,
                                          bool live, double uo, uint64_t& xb, double& lq) {
  const int ob = o >> 4, qq = (o & 15) >> 2, t = o & 3;
  const f32x4 lb = ob == 0 ? L.b0 : ob == 1 ? L.b1 : ob == 2 ? L.b2 : L.b3;
  const float acc = t == 0 ? lb[0] : t == 1 ? lb[1] : t == 2 ? lb[2] : lb[3];
  int one = 0;
  double term = 0.0;
  if (q == qq && live) one = made_bit(1.0f * acc + b3[o], uo, &term);
  one = __shfl(one, qq * 16 + r);
  term = __shfl(term, qq * 16 + r);
  lq += term;
  xb |= (uint64_t)(one & 1) << o;
}

__device__ __forceinline__ bf16x4 made_xfrag(uint64_t xb, int kb, int q) {
  bf16x4 b;
#pragma unroll
  for (int i = 0; i < 4; ++i) b[i] = (short)(((xb >> (kb * 16 + 4 * q + i)) & 1u) ? 0x3f80 : 0);
  return b;
}

__device__ __forceinline__ bf16x4 made_relu_bf16(f32x4 acc, const float* __restrict__ bias) {
  bf16x4 h;
#pragma unroll
  for (int t = 0; t < 4; ++t) h[t] = (short)f32_to_bf16(fmaxf(1.0f * acc[t] + bias[t], 0.f));
  return h;
}

constexpr int kW2Ahead = 8;   // second-layer fragments of the next step loaded a step ahead

// Weight fragments as the MFMA A operand reads them: lane (r, q) holds
// W[row block * 16 + r][K block * 16 + 4q + i], i = 0..3.  Raw row-major
// weights are read in place; the packed copy (rgs_made_sample_pack) stores
// each 16 x 16 tile as 64 consecutive 8-byte fragments in lane order, so a
// fragment load is one contiguous 512-byte read whatever H is (row strides of
// a power-of-two bytes otherwise put a load's 16 rows in the same cache sets).
// First-layer fragments are zero at inputs k >= the block's degree d (the
// unit sees inputs < d), as the step sequence's K = d product pads them.
struct MadeWeights {
  const uint16_t* W1;   // [H][n]
  const uint16_t* W2;   // [H][H]
  const uint16_t* W3;   // [n][H]
  const bf16x4* P;      // packed copy, or null
  int n, H, nb, nkb1;   // nb = H / 16 row blocks, nkb1 = ceil(n / 16) input blocks

  // kP: the packed copy is known to be present (no raw-layout code emitted)
  template <bool kP = false>
  __device__ __forceinline__ bf16x4 w1(int blk, int kb, int d, int r, int q, int lane) const {
    if (kP || P) return P[((size_t)blk * nkb1 + kb) * 64 + lane];
    const uint16_t* w1r = W1 + (size_t)(blk * 16 + r) * n;
    const int k = kb * 16 + 4 * q;
    const bf16x4 a = {(short)(k < d ? w1r[k] : 0), (short)(k + 1 < d ? w1r[k + 1] : 0),
                      (short)(k + 2 < d ? w1r[k + 2] : 0), (short)(k + 3 < d ? w1r[k + 3] : 0)};
    return a;
  }
  template <bool kP = false>
  __device__ __forceinline__ bf16x4 w2(int blk, int kb, int r, int q, int lane) const {
    if (kP || P) return P[((size_t)nb * nkb1 + (size_t)blk * nb + kb) * 64 + lane];
    return load4(W2 + (size_t)(blk * 16 + r) * H + kb * 16 + 4 * q);
  }
  template <bool kP = false>
  __device__ __forceinline__ bf16x4 w3(int ob, int blk, int r, int q, int lane) const {
    if (kP || P) return P[((size_t)nb * nkb1 + (size_t)nb * nb + (size_t)ob * nb + blk) * 64 + lane];
    const bf16x4 zero = {0, 0, 0, 0};
    return 16 * ob + r < n ? load4(W3 + (size_t)(16 * ob + r) * H + blk * 16 + 4 * q) : zero;
  }
};

// Weight fragments of the block that serves step st (one block per degree):
// first layer over inputs < st, the first AHEAD second-layer K blocks, and
// the logit rows.
template <bool kP, int AHEAD>
__device__ __forceinline__ void made_prefetch(const MadeWeights& wt, int nob, int st, int r, int q, int lane,
                                              bf16x4 (&w1n)[4], bf16x4 (&w2n)[AHEAD], bf16x4 (&w3n)[4]) {
  const bf16x4 zero = {0, 0, 0, 0};
  const int blk = st - 1;
#pragma unroll
  for (int kb = 0; kb < 4; ++kb) w1n[kb] = kb < wt.nkb1 ? wt.w1<kP>(blk, kb, st, r, q, lane) : zero;
#pragma unroll
  for (int kb = 0; kb < AHEAD; ++kb) w2n[kb] = kb < st ? wt.w2<kP>(blk, kb, r, q, lane) : zero;
#pragma unroll
  for (int ob = 0; ob < 4; ++ob) w3n[ob] = ob < nob ? wt.w3<kP>(ob, blk, r, q, lane) : zero;
}

// The n draws of one 16-row group when every degree owns one block (block
// s-1 serves step s).  A step's weight fragments (first layer, the first
// AHEAD second-layer K blocks, the logit rows) are loaded during the previous
// step, before its draw, so their L2 latency overlaps the draw's FP64 work;
// later second-layer K blocks are read AHEAD at a time.  The AHEAD grouping
// changes only when loads are issued: the MFMAs run in ascending K order
// into one accumulator whatever it is.
template <bool kP, int AHEAD>
__device__ __forceinline__ void made_rows_one(const MadeWeights& wt, const float* __restrict__ b1,
                                              const float* __restrict__ b2, const float* __restrict__ b3,
                                              const double* __restrict__ u, int row, bool live, int r, int q,
                                              int lane, int nob, int ndeg, bf16x4* h1, uint64_t& xb, double& lq) {
  const int n = wt.n;
  const bf16x4 zero = {0, 0, 0, 0};
  LogitBlocks L = {{0.f, 0.f, 0.f, 0.f}, {0.f, 0.f, 0.f, 0.f}, {0.f, 0.f, 0.f, 0.f}, {0.f, 0.f, 0.f, 0.f}};
  double un = live ? u[(size_t)row * n] : 0.0;   // uniform of the next draw, loaded a step ahead
  bf16x4 w1n[4], w2n[AHEAD], w3n[4];
  if (ndeg >= 1) made_prefetch<kP, AHEAD>(wt, nob, 1, r, q, lane, w1n, w2n, w3n);
  for (int o = 0; o < n; ++o) {
    const double uo = un;
    if (o + 1 < n && live) un = u[(size_t)row * n + o + 1];
    if (o > 0 && o <= ndeg) {
      const int blk = o - 1;
      f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int kb = 0; kb < 4; ++kb)
        if (kb * 16 < o) acc = mfma16(w1n[kb], made_xfrag(xb, kb, q), acc);
      h1[blk * 64 + lane] = made_relu_bf16(acc, b1 + blk * 16 + 4 * q);
      acc = (f32x4){0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int kb = 0; kb < AHEAD; ++kb)
        if (kb < o) acc = mfma16(w2n[kb], h1[kb * 64 + lane], acc);
      for (int kb = AHEAD; kb < o; kb += AHEAD) {
        bf16x4 wa[AHEAD];
#pragma unroll
        for (int i = 0; i < AHEAD; ++i) wa[i] = kb + i < o ? wt.w2<kP>(blk, kb + i, r, q, lane) : zero;
#pragma unroll
        for (int i = 0; i < AHEAD; ++i)
          if (kb + i < o) acc = mfma16(wa[i], h1[(kb + i) * 64 + lane], acc);
      }
      made_logits(L, nob, w3n, made_relu_bf16(acc, b2 + blk * 16 + 4 * q));
    }
    if (o + 1 <= ndeg) made_prefetch<kP, AHEAD>(wt, nob, o + 1, r, q, lane, w1n, w2n, w3n);
    made_draw(L, b3, o, q, r, live, uo, xb, lq);
  }
}

// A row's results: packed words (the scan's two-word layout), log q, and
// optionally the bf16 0/1 bits.
__device__ __forceinline__ void made_store(int n, int row, int q, uint64_t xb, double lq,
                                           uint64_t* __restrict__ words, uint16_t* __restrict__ xout,
                                           double* __restrict__ logq) {
  if (q == 0) {
    if (words) {
      words[2 * (size_t)row] = xb;
      words[2 * (size_t)row + 1] = 0;
    }
    if (logq) logq[row] = lq;
  }
  if (xout)
    for (int k = q; k < n; k += 4) xout[(size_t)row * n + k] = ((xb >> k) & 1u) ? (uint16_t)0x3f80 : (uint16_t)0;
}

__global__ void __launch_bounds__(64) k_made_sample16(MadeWeights wt, const float* __restrict__ b1,
                                                      const float* __restrict__ b2, const float* __restrict__ b3,
                                                      const double* __restrict__ u, int B,
                                                      uint64_t* __restrict__ words, uint16_t* __restrict__ xout,
                                                      double* __restrict__ logq) {
  const int n = wt.n, H = wt.H;
  extern __shared__ __align__(16) unsigned char smem[];
  bf16x4* h1 = reinterpret_cast<bf16x4*>(smem);   // [H / 16 blocks][64 lanes]
  const int lane = threadIdx.x;
  const int r = lane & 15, q = lane >> 4;
  // g16 blocks per degree for degrees 1 .. ndeg (ndeg = n - 1, or fewer when
  // the width has fewer blocks than degrees: one block each, none above)
  const int nb = H >> 4, g16 = nb >= n - 1 ? nb / (n - 1) : 1, ndeg = nb / g16, nob = (n + 15) >> 4;
  const bf16x4 zero = {0, 0, 0, 0};
  for (int base = blockIdx.x * 16; base < B; base += gridDim.x * 16) {
    const int row = base + r;
    const bool live = row < B;
    uint64_t xb = 0;   // bits drawn so far for this lane's row
    double lq = 0.0;
    if (g16 == 1) {
      made_rows_one<false, kW2Ahead>(wt, b1, b2, b3, u, row, live, r, q, lane, nob, ndeg, h1, xb, lq);
    } else {
      LogitBlocks L = {{0.f, 0.f, 0.f, 0.f}, {0.f, 0.f, 0.f, 0.f}, {0.f, 0.f, 0.f, 0.f}, {0.f, 0.f, 0.f, 0.f}};
      double un = live ? u[(size_t)row * n] : 0.0;   // uniform of the next draw, loaded a step ahead
      for (int o = 0; o < n; ++o) {
        const double uo = un;
        if (o + 1 < n && live) un = u[(size_t)row * n + o + 1];
        if (o > 0) {
          const int g0 = (o - 1) * g16;
          for (int j = 0; j < g16; ++j) {   // first layer: inputs 0 .. o-1
            const int blk = g0 + j;
            f32x4 acc = {0.f, 0.f, 0.f, 0.f};
            for (int kb = 0; kb * 16 < o; ++kb) acc = mfma16(wt.w1(blk, kb, o, r, q, lane), made_xfrag(xb, kb, q), acc);
            h1[blk * 64 + lane] = made_relu_bf16(acc, b1 + blk * 16 + 4 * q);
          }
          const int nkb = g0 + g16;   // first-layer blocks of degree <= o
          for (int j = 0; j < g16; ++j) {   // second layer, then its logit contributions
            const int blk = g0 + j;
            f32x4 acc = {0.f, 0.f, 0.f, 0.f};
            int kb = 0;
            for (; kb + 8 <= nkb; kb += 8) {   // eight fragment loads in flight, then the MFMAs in K order
              bf16x4 wa[8], hb[8];
#pragma unroll
              for (int i = 0; i < 8; ++i) {
                wa[i] = wt.w2(blk, kb + i, r, q, lane);
                hb[i] = h1[(kb + i) * 64 + lane];
              }
#pragma unroll
              for (int i = 0; i < 8; ++i) acc = mfma16(wa[i], hb[i], acc);
            }
            for (; kb < nkb; ++kb) acc = mfma16(wt.w2(blk, kb, r, q, lane), h1[kb * 64 + lane], acc);
            bf16x4 w3[4];
#pragma unroll
            for (int ob = 0; ob < 4; ++ob) w3[ob] = ob < nob ? wt.w3(ob, blk, r, q, lane) : zero;
            made_logits(L, nob, w3, made_relu_bf16(acc, b2 + blk * 16 + 4 * q));
          }
        }
        made_draw(L, b3, o, q, r, live, uo, xb, lq);
      }
    }
    if (live) made_store(n, row, q, xb, lq, words, xout, logq);
  }
}

// The packed-weight, one-block-per-degree case (H = 16 k, k < n - 1, as the
// 34 -> 512 -> 512 -> 34 proposal): the same draws as k_made_sample16, with
// the raw-layout paths compiled out and the register budget held to WAVES
// wavefronts per SIMD.  The draw chain of one 16-row group is latency-bound
// (n dependent steps of L2 reads, MFMA chains and FP64 draws), so more
// resident groups per SIMD hide more of it.
template <int AHEAD, int WAVES>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(WAVES, WAVES)))
k_made_sample_one(MadeWeights wt, const float* __restrict__ b1, const float* __restrict__ b2,
                  const float* __restrict__ b3, const double* __restrict__ u, int B, uint64_t* __restrict__ words,
                  uint16_t* __restrict__ xout, double* __restrict__ logq) {
  extern __shared__ __align__(16) unsigned char smem[];
  bf16x4* h1 = reinterpret_cast<bf16x4*>(smem);   // [H / 16 blocks][64 lanes]
  const int lane = threadIdx.x;
  const int r = lane & 15, q = lane >> 4;
  const int n = wt.n, ndeg = wt.nb, nob = (n + 15) >> 4;
  for (int base = blockIdx.x * 16; base < B; base += gridDim.x * 16) {
    const int row = base + r;
    const bool live = row < B;
    uint64_t xb = 0;
    double lq = 0.0;
    made_rows_one<true, AHEAD>(wt, b1, b2, b3, u, row, live, r, q, lane, nob, ndeg, h1, xb, lq);
    if (live) made_store(n, row, q, xb, lq, words, xout, logq);
  }
}

__global__ void k_f32_to_bf16(const float* in, uint16_t* out, long long count) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) out[i] = f32_to_bf16(in[i]);
}

// Fragment-major copy of the sampler's weights (see MadeWeights): first
// layer [nb][nkb1] tiles (zero at inputs >= the block's degree), second layer
// [nb][nb] tiles, logit rows [nob][nb] tiles (zero rows past n); each tile is
// 64 lanes x 4 bf16.
__global__ void k_made_pack(const uint16_t* __restrict__ W1, const uint16_t* __restrict__ W2,
                            const uint16_t* __restrict__ W3, int n, int H, int nb, int nkb1, int g16,
                            bf16x4* __restrict__ out, long long fragments) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= fragments * 64) return;
  const long long f = t >> 6;
  const int lane = (int)(t & 63), r = lane & 15, q = lane >> 4;
  bf16x4 v = {0, 0, 0, 0};
  if (f < (long long)nb * nkb1) {
    const int blk = (int)(f / nkb1), kb = (int)(f % nkb1);
    const int d = blk / g16 + 1;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int k = kb * 16 + 4 * q + i;
      v[i] = (short)(k < d && k < n ? W1[(size_t)(blk * 16 + r) * n + k] : 0);
    }
  } else if (f < (long long)nb * nkb1 + (long long)nb * nb) {
    const long long g = f - (long long)nb * nkb1;
    const int blk = (int)(g / nb), kb = (int)(g % nb);
    v = load4(W2 + (size_t)(blk * 16 + r) * H + kb * 16 + 4 * q);
  } else {
    const long long h = f - (long long)nb * nkb1 - (long long)nb * nb;
    const int ob = (int)(h / nb), blk = (int)(h % nb);
    if (16 * ob + r < n) v = load4(W3 + (size_t)(16 * ob + r) * H + blk * 16 + 4 * q);
  }
  out[t] = v;
}

static bool made_blocked(int n, int H) {
  return n >= 2 && n <= 64 && H >= 16 && H % 16 == 0 && H <= 2048 && ((H / 16) % (n - 1) == 0 || H / 16 < n - 1);
}

static long long made_pack_fragments(int n, int H) {
  const long long nb = H / 16, nkb1 = (n + 15) / 16, nob = (n + 15) / 16;
  return nb * nkb1 + nb * nb + nob * nb;
}

static int launch_made_sample16(const MadeWeights& wt, const float* b1, const float* b2, const float* b3,
                                const double* u, int B, uint64_t* words, uint16_t* xout, double* logq,
                                hipStream_t stream) {
  typedef void (*Kernel)(MadeWeights, const float*, const float*, const float*, const double*, int, uint64_t*,
                         uint16_t*, double*);
  constexpr int kOneAhead = 4, kOneWaves = 3;   // 165 VGPRs: three wavefronts per SIMD, no spills
  const bool one = wt.P && wt.nb <= wt.n - 1;   // packed copy, one block per degree
  const Kernel kern = one ? &k_made_sample_one<kOneAhead, kOneWaves> : &k_made_sample16;
  const size_t lds = (size_t)wt.nb * 64 * 8;
  const int groups = (B + 15) / 16;
  // one round of the workgroups that fit a CU at once (registers and LDS),
  // so every wave takes an even share of the 16-row groups
  static size_t cached_lds[2] = {0, 0};
  static int cached_per_cu[2] = {0, 0};
  if (lds != cached_lds[one]) {
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(kern), 64, lds) !=
            hipSuccess ||
        per_cu < 1)
      per_cu = std::max(1, (int)((160 * 1024) / lds));
    cached_lds[one] = lds;
    cached_per_cu[one] = per_cu;
  }
  const int grid = std::min(groups, 256 * cached_per_cu[one]);
  hipLaunchKernelGGL(kern, dim3(grid), dim3(64), lds, stream, wt, b1, b2, b3, u, B, words, xout, logq);
  RGS_LAUNCH_CHECK(one ? "k_made_sample_one" : "k_made_sample16");
  return kOk;
}

}  // namespace rgs

using namespace rgs;

extern "C" int rgs_gemm_bf16(int ta, int tb, int M, int N, int K, const uint16_t* A, int lda, const uint16_t* B,
                             int ldb, float alpha, float* C, int ldc, uint16_t* Cb, int ldcb, const float* bias,
                             int relu, const uint16_t* dmask, int ldm, int accumulate, void* stream) {
  // K = 0 is allowed: the product is empty and the epilogue sees zeros (bias only)
  if (!A || !B || (!C && !Cb) || M <= 0 || N <= 0 || K < 0) return fail(kInvalid, "gemm: bad arguments");
  GemmArgs g{A, B, C, Cb, bias, dmask, M, N, K, lda, ldb, ldc, ldcb, ldm, ta, tb, relu, accumulate, alpha};
  dim3 grid((N + TN - 1) / TN, (M + TM - 1) / TM);
  hipLaunchKernelGGL(k_gemm, grid, dim3(256), 0, reinterpret_cast<hipStream_t>(stream), g);
  RGS_LAUNCH_CHECK("k_gemm");
  return kOk;
}

extern "C" int rgs_made_bce(const float* logits, const uint16_t* x, int B, int n, float scale, float* nll,
                            float* dlogits, uint16_t* dlogits_bf16, void* stream) {
  if (!logits || !x || B <= 0 || n <= 0) return fail(kInvalid, "bce: bad arguments");
  hipLaunchKernelGGL(k_bce, dim3((B + 3) / 4), dim3(256), 0, reinterpret_cast<hipStream_t>(stream), logits, x,
                     B, n, scale, nll, dlogits, dlogits_bf16);
  RGS_LAUNCH_CHECK("k_bce");
  return kOk;
}

extern "C" int rgs_colsum(const float* in, int rows, int cols, int ld, float* out, int accumulate, void* stream) {
  if (!in || !out || rows <= 0 || cols <= 0) return fail(kInvalid, "colsum: bad arguments");
  hipLaunchKernelGGL(k_colsum<float>, dim3((cols + 63) / 64), dim3(64 * kColRowLanes), 0,
                     reinterpret_cast<hipStream_t>(stream), in, rows, cols, ld, out, accumulate);
  RGS_LAUNCH_CHECK("k_colsum");
  return kOk;
}

extern "C" int rgs_colsum_bf16(const uint16_t* in, int rows, int cols, int ld, float* out, void* stream) {
  if (!in || !out || rows <= 0 || cols <= 0) return fail(kInvalid, "colsum: bad arguments");
  hipLaunchKernelGGL(k_colsum<uint16_t>, dim3((cols + 63) / 64), dim3(64 * kColRowLanes), 0,
                     reinterpret_cast<hipStream_t>(stream), in, rows, cols, ld, out, 0);
  RGS_LAUNCH_CHECK("k_colsum_bf16");
  return kOk;
}

extern "C" int rgs_adam(float* w, const float* grad, const float* mask, float* m, float* v, uint16_t* wb,
                        long long count, int* t_dev, float lr, float b1, float b2, float eps, void* stream) {
  if (!w || !grad || !mask || !m || !v || !wb || !t_dev || count <= 0) return fail(kInvalid, "adam: bad arguments");
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  hipLaunchKernelGGL(k_step_counter, dim3(1), dim3(1), 0, st, t_dev);
  RGS_LAUNCH_CHECK("k_step_counter");
  hipLaunchKernelGGL(k_adam, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, st, w, grad, mask, m, v, wb, count,
                     t_dev, lr, b1, b2, eps);
  RGS_LAUNCH_CHECK("k_adam");
  return kOk;
}

extern "C" int rgs_sample_col(const float* logits, const double* u, int B, int n, int o, uint16_t* x, void* stream) {
  if (!logits || !u || !x || o < 0 || o >= n) return fail(kInvalid, "sample_col: bad arguments");
  hipLaunchKernelGGL(k_sample_col, dim3((B + 255) / 256), dim3(256), 0, reinterpret_cast<hipStream_t>(stream),
                     logits, u, B, n, o, x);
  RGS_LAUNCH_CHECK("k_sample_col");
  return kOk;
}

extern "C" int rgs_f32_to_bf16(const float* in, uint16_t* out, long long count, void* stream) {
  if (!in || !out || count <= 0) return fail(kInvalid, "f32_to_bf16: bad arguments");
  hipLaunchKernelGGL(k_f32_to_bf16, dim3((unsigned)((count + 255) / 256)), dim3(256), 0,
                     reinterpret_cast<hipStream_t>(stream), in, out, count);
  RGS_LAUNCH_CHECK("k_f32_to_bf16");
  return kOk;
}

extern "C" int rgs_made_sample_step(const float* logit, int ldl, const double* u, int B, int n, int o, uint16_t* x,
                                    double* logq, int first, void* stream) {
  if (!logit || !u || !x || !logq || B <= 0 || o < 0 || o >= n || ldl <= 0)
    return fail(kInvalid, "made_sample_step: bad arguments");
  hipLaunchKernelGGL(k_sample_step, dim3((B + 255) / 256), dim3(256), 0, reinterpret_cast<hipStream_t>(stream),
                     logit, ldl, u, B, n, o, x, logq, first);
  RGS_LAUNCH_CHECK("k_sample_step");
  return kOk;
}

extern "C" int rgs_made_sample(const uint16_t* W1, const float* b1, const uint16_t* W2, const float* b2,
                               const uint16_t* W3, const float* b3, const int* kpref, const double* u, int B, int n,
                               int H, uint64_t* words, uint16_t* xout, double* logq, void* stream) {
  if (!W1 || !b1 || !W2 || !b2 || !W3 || !b3 || !kpref || !u || B <= 0 || n < 1 || n > 64 || H < 16 ||
      H % 16 != 0 || (!words && !xout))
    return fail(kInvalid, "made_sample: bad arguments (n <= 64, H a multiple of 16)");
  if (!made_blocked(n, H))
    return fail(kInvalid,
                "made_sample: the fused sampler needs every degree to own whole aligned 16-unit blocks "
                "(H = 16 m (n - 1) or H = 16 k with k < n - 1, H <= 2048); use the step sequence for other widths");
  const MadeWeights wt{W1, W2, W3, nullptr, n, H, H / 16, (n + 15) / 16};
  return launch_made_sample16(wt, b1, b2, b3, u, B, words, xout, logq, reinterpret_cast<hipStream_t>(stream));
}

extern "C" int rgs_made_sample_pack(const uint16_t* W1, const uint16_t* W2, const uint16_t* W3, int n, int H,
                                    void* packed, uint64_t* bytes, void* stream) {
  if (!bytes) return fail(kInvalid, "made_sample_pack: null size");
  if (!made_blocked(n, H)) return fail(kInvalid, "made_sample_pack: width without aligned degree blocks");
  const long long fragments = made_pack_fragments(n, H);
  *bytes = (uint64_t)fragments * 64 * 8;
  if (!packed) return kOk;
  if (!W1 || !W2 || !W3) return fail(kInvalid, "made_sample_pack: null weights");
  const int nb = H / 16, g16 = nb >= n - 1 ? nb / (n - 1) : 1;
  const long long threads = fragments * 64;
  hipLaunchKernelGGL(k_made_pack, dim3((unsigned)((threads + 255) / 256)), dim3(256), 0,
                     reinterpret_cast<hipStream_t>(stream), W1, W2, W3, n, H, nb, (n + 15) / 16, g16,
                     reinterpret_cast<bf16x4*>(packed), fragments);
  RGS_LAUNCH_CHECK("k_made_pack");
  return kOk;
}

extern "C" int rgs_made_sample_packed(const void* packed, const float* b1, const float* b2, const float* b3,
                                      const double* u, int B, int n, int H, uint64_t* words, uint16_t* xout,
                                      double* logq, void* stream) {
  if (!packed || !b1 || !b2 || !b3 || !u || B <= 0 || (!words && !xout))
    return fail(kInvalid, "made_sample_packed: bad arguments");
  if (!made_blocked(n, H)) return fail(kInvalid, "made_sample_packed: width without aligned degree blocks");
  const MadeWeights wt{nullptr, nullptr, nullptr, reinterpret_cast<const bf16x4*>(packed), n, H, H / 16,
                       (n + 15) / 16};
  return launch_made_sample16(wt, b1, b2, b3, u, B, words, xout, logq, reinterpret_cast<hipStream_t>(stream));
}
