// Batched linear rank-width scan: p(x) of many proposals in one persistent
// launch.  One workgroup of W wavefronts owns one proposal at a time and
// sweeps all n steps with the environment table resident in LDS (2^rmax
// entries); outputs of a step are gathered into registers (K per lane), the
// waves combine their maximum modulus through LDS, and the outputs are written
// back in place scaled by an exact power of two, so there is no host round
// trip between steps.  Sums and weights are FP64; the table is stored as
// complex64 (S = float2) or complex128 (S = double2, the audit precision).
#include "rgs_common.h"
#include "rgs_amd.h"

#include <algorithm>

namespace rgs {

template <typename S>
__device__ __forceinline__ double2 load_c(const S& s) {
  return make_double2((double)s.x, (double)s.y);
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

// Wave w of the workgroup computes the output blocks k = w, w + W, ... of a
// step (at most K of them per lane).  Every read of the step's input table
// ends at the barrier that also publishes the waves' maxima, so the outputs
// overwrite the table in place.  W waves per proposal keep the LDS per
// proposal while giving each CU W times the wavefronts to hide LDS and FP64
// latency, and need W times fewer output registers per lane.
template <typename S, int K, int W>
__global__ void __launch_bounds__(64 * W) k_scan_batch(const uint64_t* __restrict__ plan,
                                                      const double2* __restrict__ u,
                                                      const uint64_t* __restrict__ bits, int count,
                                                      double2* __restrict__ out_val, int32_t* __restrict__ out_exp) {
  extern __shared__ __align__(16) unsigned char smem[];
  S* E = reinterpret_cast<S*>(smem);
  __shared__ double wmx[W];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int n = (int)plan[0];
  for (int pidx = blockIdx.x; pidx < count; pidx += gridDim.x) {
    const uint64_t x0 = bits[2 * pidx], x1 = bits[2 * pidx + 1];
    if (threadIdx.x == 0) E[0] = make_s<S>(1.0, 0.0);
    __syncthreads();
    int esum = 0;
    for (int j = 0; j < n; ++j) {
      const uint64_t* st = plan + plan[2 + j];
      const uint64_t hdr = st[0];
      const int v = (int)(hdr & 255u), r_in = (int)((hdr >> 8) & 255u), r_out = (int)((hdr >> 16) & 255u);
      const int nker = (int)((hdr >> 24) & 255u);
      const uint64_t beta = st[1];
      const int xv = (int)((v < 64 ? (x0 >> v) : (x1 >> (v - 64))) & 1u);
      const double2 f0 = u[(v * 2 + xv) * 2 + 0], f1 = u[(v * 2 + xv) * 2 + 1];
      const uint64_t lo = st[2 + lane];
      const uint64_t* hi = st + 66;
      const uint64_t* combos = st + 66 + (1ull << (r_out > 6 ? r_out - 6 : 0));
      const int ncomb = 1 << nker;
      const uint64_t mask = (1ull << r_in) - 1ull;
      const int nk = r_out > 6 ? 1 << (r_out - 6) : 1;
      const bool lane_ok = r_out >= 6 || lane < (1 << r_out);
      S outv[K];
      double mx = 0.0;
#pragma unroll
      for (int kk = 0; kk < K; ++kk) {
        const int k = wave + W * kk;
        if (k < nk && lane_ok) {
          const uint64_t p0 = lo ^ hi[k];
          double ar = 0.0, ai = 0.0;
          for (int t = 0; t < ncomb; ++t) {
            const uint64_t p = p0 ^ combos[t];
            const uint64_t c = p & mask;
            const uint32_t y = (uint32_t)(p >> r_in) & 1u;
            const uint32_t neg = y & ((uint32_t)__popcll(c & beta) & 1u);
            const double2 w = y ? f1 : f0;
            const double2 e = load_c(E[c]);
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
          mx = fmax(mx, ar * ar + ai * ai);
          outv[kk] = make_s<S>(ar, ai);
        }
      }
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) mx = fmax(mx, __shfl_xor(mx, off));
      if (lane == 0) wmx[wave] = mx;
      __syncthreads();  // every read of E done; the waves' maxima visible
#pragma unroll
      for (int w2 = 0; w2 < W; ++w2) mx = fmax(mx, wmx[w2]);
      int e = 0;
      if (mx > 0.0) frexp(sqrt(mx), &e);
#pragma unroll
      for (int kk = 0; kk < K; ++kk) {
        const int k = wave + W * kk;
        if (k < nk && lane_ok) {
          S val = outv[kk];
          val.x = ldexp(val.x, -e);
          val.y = ldexp(val.y, -e);
          E[lane + 64 * k] = val;
        }
      }
      esum += e;
      __syncthreads();
    }
    if (threadIdx.x == 0) {
      out_val[pidx] = load_c(E[0]);
      out_exp[pidx] = esum;
    }
    __syncthreads();
  }
}

template <typename S, int K, int W>
static hipError_t launch_scan(const uint64_t* plan, const double2* u, const uint64_t* bits, int count,
                                double2* val, int32_t* ex, int rmax, hipStream_t st) {
  size_t shmem = ((size_t)1 << rmax) * sizeof(S);
  int grid = std::min(count, 256 * 16);
  static bool attr = false;
  if (!attr) {
    // dynamic LDS up to the CU's 160 KiB less the kernel's own static array
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_scan_batch<S, K, W>),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024 - 256);
    if (e != hipSuccess) return e;
    attr = true;
  }
  hipLaunchKernelGGL((k_scan_batch<S, K, W>), dim3(grid), dim3(64 * W), shmem, st, plan, u, bits, count, val, ex);
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

template <typename S>
__global__ void __launch_bounds__(256) k_scan_step(const uint64_t* __restrict__ st, const S* __restrict__ ein,
                                                   S* __restrict__ eout, double2 f0, double2 f1,
                                                   const unsigned long long* __restrict__ mx_prev,
                                                   unsigned long long* __restrict__ mx_out) {
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
  const uint64_t total = 1ull << r_out;
  double lmax = 0.0;
  for (uint64_t cp = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; cp < total;
       cp += (uint64_t)gridDim.x * blockDim.x) {
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
    const uint64_t outs = 1ull << s.r_out;
    const uint64_t blocks = std::min<uint64_t>((outs + 255) / 256, 256ull * 64);
    hipLaunchKernelGGL((k_scan_step<S>), dim3((unsigned)blocks), dim3(256), 0, st, plan + plan_host_offsets[j],
                       b[j & 1], b[(j + 1) & 1], f0, f1, j ? mx + (j - 1) : nullptr, mx + j);
    RGS_LAUNCH_CHECK("k_scan_step");
  }
  hipLaunchKernelGGL((k_scan_finish<S>), dim3(1), dim3(64), 0, st, b[n & 1], mx, n, out_val, out_exp);
  RGS_LAUNCH_CHECK("k_scan_finish");
  return kOk;
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
  // <K, W>: W waves per proposal, K = 2^(rmax-6) / W outputs per lane.  Four
  // waves for complex64 tables and eight for complex128 were the fastest
  // measured at rank 12 (5.5 and 3.9 M proposals/s).
  if (precision == 64) {
    if (rmax <= 6) e = launch_scan<float2, 1, 1>(plan_dev, u, bits_dev, count, val, out_exp, 6, st);
    else if (rmax <= 8) e = launch_scan<float2, 1, 4>(plan_dev, u, bits_dev, count, val, out_exp, 8, st);
    else if (rmax <= 10) e = launch_scan<float2, 4, 4>(plan_dev, u, bits_dev, count, val, out_exp, 10, st);
    else if (rmax <= 12) e = launch_scan<float2, 16, 4>(plan_dev, u, bits_dev, count, val, out_exp, 12, st);
    else if (rmax <= 13) e = launch_scan<float2, 32, 4>(plan_dev, u, bits_dev, count, val, out_exp, 13, st);
    else return fail(kInvalid, "scan_batch: complex64 batch limited to rmax <= 13");
  } else if (precision == 128) {
    if (rmax <= 6) e = launch_scan<double2, 1, 1>(plan_dev, u, bits_dev, count, val, out_exp, 6, st);
    else if (rmax <= 8) e = launch_scan<double2, 1, 4>(plan_dev, u, bits_dev, count, val, out_exp, 8, st);
    else if (rmax <= 10) e = launch_scan<double2, 2, 8>(plan_dev, u, bits_dev, count, val, out_exp, 10, st);
    else if (rmax <= 11) e = launch_scan<double2, 4, 8>(plan_dev, u, bits_dev, count, val, out_exp, 11, st);
    else if (rmax <= 12) e = launch_scan<double2, 8, 8>(plan_dev, u, bits_dev, count, val, out_exp, 12, st);
    else return fail(kInvalid, "scan_batch: complex128 batch limited to rmax <= 12");
  } else {
    return fail(kInvalid, "precision must be 64 (complex64 tables) or 128 (complex128)");
  }
  if (e != hipSuccess) return fail(kHip, std::string("k_scan_batch: ") + hipGetErrorString(e));
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
