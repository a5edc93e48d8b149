// Host planning: error plumbing, GF(2) algebra and chunked lookup tables.
//
// All arithmetic here is exact integer work on bit masks (uint64 rows for
// matrices up to 64 columns, two words up to 128).  The tables follow the
// conventions documented in include/rgs_amd.h.
#include <stdint.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>
#include <string>
#include <vector>

#include "rgs_amd.h"
#include "rgs_status.h"

namespace rgs {

static thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }
int fail(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

static inline int parity(uint64_t x) { return __builtin_popcountll(x) & 1; }

// Rows of an n x n 0/1 matrix as masks (bit j = column j).
static std::vector<uint64_t> pack_rows(const uint8_t* m, int rows, int cols, int stride) {
  std::vector<uint64_t> r(rows, 0);
  for (int i = 0; i < rows; ++i)
    for (int j = 0; j < cols; ++j)
      if (m[i * stride + j] & 1) r[i] |= 1ull << j;
  return r;
}

// Gauss-Jordan over GF(2); returns false if singular.
static bool inverse_rows(std::vector<uint64_t> a, int n, std::vector<uint64_t>& inv) {
  inv.assign(n, 0);
  for (int i = 0; i < n; ++i) inv[i] = 1ull << i;
  for (int col = 0; col < n; ++col) {
    int piv = -1;
    for (int r = col; r < n; ++r)
      if ((a[r] >> col) & 1) { piv = r; break; }
    if (piv < 0) return false;
    std::swap(a[col], a[piv]);
    std::swap(inv[col], inv[piv]);
    for (int r = 0; r < n; ++r)
      if (r != col && ((a[r] >> col) & 1)) {
        a[r] ^= a[col];
        inv[r] ^= inv[col];
      }
  }
  return true;
}

// Phase words for a quadratic form given by strict-upper rows (bit j of
// rows[i] = Q_ij, j > i) and a linear mask: one table of 256 words per chunk.
static void phase_words(const std::vector<uint64_t>& rows, uint64_t linear, int n, uint64_t* out) {
  const int chunks = (n + 7) / 8;
  const uint64_t full = n == 64 ? ~0ull : ((1ull << n) - 1);
  for (int c = 0; c < chunks; ++c) {
    const int lo = 8 * c, hi = std::min(n, 8 * c + 8);
    const uint64_t higher = full & ~((hi == 64) ? ~0ull : ((1ull << hi) - 1));
    for (int code = 0; code < 256; ++code) {
      uint64_t word = 0;
      if (code < (1 << (hi - lo))) {
        const uint64_t z = (uint64_t)code << lo;
        uint64_t cross = 0;
        int inner = parity(z & linear);
        for (int j = lo; j < hi; ++j)
          if ((z >> j) & 1) {
            cross ^= rows[j] & higher;
            inner ^= parity(z & rows[j]);
          }
        word = cross | ((uint64_t)inner << 63);
      }
      out[c * 256 + code] = word;
    }
  }
}

}  // namespace rgs

using namespace rgs;

extern "C" const char* rgs_last_error(void) { return g_last_error.c_str(); }
extern "C" int rgs_abi_version(void) { return RGS_ABI_VERSION; }

extern "C" int rgs_gf2_inverse(const uint8_t* m, int n, uint8_t* inv) {
  if (!m || !inv || n < 1 || n > 64) return fail(kInvalid, "gf2_inverse: need 1 <= n <= 64");
  std::vector<uint64_t> r;
  if (!inverse_rows(pack_rows(m, n, n, n), n, r)) return fail(kInvalid, "singular GF(2) matrix");
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) inv[i * n + j] = (r[i] >> j) & 1;
  return kOk;
}

extern "C" int rgs_gf2_rank(const uint8_t* m, int rows, int cols, int* rank) {
  if (!m || !rank || rows < 0 || cols < 0 || cols > 128) return fail(kInvalid, "gf2_rank: cols <= 128");
  std::vector<std::pair<uint64_t, uint64_t>> a(rows);
  for (int i = 0; i < rows; ++i) {
    uint64_t lo = 0, hi = 0;
    for (int j = 0; j < cols; ++j)
      if (m[i * cols + j] & 1) (j < 64 ? lo : hi) |= 1ull << (j & 63);
    a[i] = {lo, hi};
  }
  int r = 0;
  for (int col = 0; col < cols && r < rows; ++col) {
    auto bit = [&](int i) { return col < 64 ? (a[i].first >> col) & 1 : (a[i].second >> (col - 64)) & 1; };
    int piv = -1;
    for (int i = r; i < rows; ++i)
      if (bit(i)) { piv = i; break; }
    if (piv < 0) continue;
    std::swap(a[r], a[piv]);
    for (int i = r + 1; i < rows; ++i)
      if (bit(i)) {
        a[i].first ^= a[r].first;
        a[i].second ^= a[r].second;
      }
    ++r;
  }
  *rank = r;
  return kOk;
}

extern "C" int rgs_graph_phase_lut(const uint8_t* g, int n, uint64_t* lut) {
  if (!g || !lut || n < 1 || n > 63) return fail(kInvalid, "graph_phase_lut: need 1 <= n <= 63");
  std::vector<uint64_t> rows(n, 0);
  for (int i = 0; i < n; ++i)
    for (int j = i + 1; j < n; ++j)
      if (g[i * n + j] & 1) rows[i] |= 1ull << j;
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j)
      if ((g[i * n + j] & 1) != (g[j * n + i] & 1) || (i == j && (g[i * n + j] & 1)))
        return fail(kInvalid, "graph must be symmetric with zero diagonal");
  phase_words(rows, 0, n, lut);
  return kOk;
}

extern "C" int rgs_cov_luts(const uint8_t* g, int nv, const int32_t* A, const int32_t* B, int m,
                            uint64_t* map_lut, uint64_t* aphase_lut) {
  if (!g || !A || !B || !map_lut || m < 1 || m > 63) return fail(kInvalid, "cov_luts: need 1 <= m <= 63");
  for (int i = 0; i < m; ++i)
    if (A[i] < 0 || A[i] >= nv || B[i] < 0 || B[i] >= nv) return fail(kInvalid, "vertex out of range");
  // M = G[B, A]; N = M^{-T}; b = N z.
  std::vector<uint64_t> M(m, 0);
  for (int i = 0; i < m; ++i)
    for (int j = 0; j < m; ++j)
      if (g[B[i] * nv + A[j]] & 1) M[i] |= 1ull << j;
  std::vector<uint64_t> Minv;
  if (!inverse_rows(M, m, Minv)) return fail(kInvalid, "cut matrix G[B,A] is singular (B marginal not uniform)");
  // N = (M^{-1})^T: column j of N = row j of M^{-1}.
  const std::vector<uint64_t>& colN = Minv;
  // Transformed quadratic: P = N^T U N with U = strict upper of G[B,B].
  std::vector<uint64_t> U(m, 0);
  for (int i = 0; i < m; ++i)
    for (int j = i + 1; j < m; ++j)
      if (g[B[i] * nv + B[j]] & 1) U[i] |= 1ull << j;
  // P_ab = sum_{i<j} N_ia U_ij N_jb = colN[a]^T U colN[b]
  std::vector<uint64_t> Ucol(m);  // U times column b of N: (U n_b)_i = parity(U[i] & colN[b])
  for (int b = 0; b < m; ++b) {
    uint64_t v = 0;
    for (int i = 0; i < m; ++i)
      if (parity(U[i] & colN[b])) v |= 1ull << i;
    Ucol[b] = v;
  }
  std::vector<uint64_t> Qrows(m, 0);
  uint64_t linear = 0;
  for (int a = 0; a < m; ++a)
    for (int b = 0; b < m; ++b) {
      int pab = parity(colN[a] & Ucol[b]);
      int pba = parity(colN[b] & Ucol[a]);
      if (a == b && pab) linear |= 1ull << a;
      if (b > a && (pab ^ pba)) Qrows[a] |= 1ull << b;
    }
  std::vector<uint64_t> words(((m + 7) / 8) * 256);
  phase_words(Qrows, linear, m, words.data());
  const int chunks = (m + 7) / 8;
  for (int c = 0; c < chunks; ++c) {
    const int lo = 8 * c, hi = std::min(m, 8 * c + 8);
    for (int code = 0; code < 256; ++code) {
      uint64_t mapped = 0;
      if (code < (1 << (hi - lo)))
        for (int j = lo; j < hi; ++j)
          if ((code >> (j - lo)) & 1) mapped ^= colN[j];
      map_lut[2 * (c * 256 + code)] = mapped;
      map_lut[2 * (c * 256 + code) + 1] = words[c * 256 + code];
    }
  }
  if (aphase_lut) {
    std::vector<uint64_t> rowsA(m, 0);
    for (int i = 0; i < m; ++i)
      for (int j = i + 1; j < m; ++j)
        if (g[A[i] * nv + A[j]] & 1) rowsA[i] |= 1ull << j;
    phase_words(rowsA, 0, m, aphase_lut);
  }
  return kOk;
}

// u = D1 R D2 with D1 = diag(e^{i a0}, e^{i a1}), R = [[c, -s], [s, c]] real,
// D2 = diag(1, e^{i b}): |u00| = c, |u10| = s, a0 = arg u00, a1 = arg u10,
// b = arg u11 - a1 (= arg(-u01) - a0 by unitarity).  Degenerate columns
// (c = 0 or s = 0) put the free phase into D1.
// Float32 rounding of a unit complex number (xd, yd): rounding each part
// independently leaves x^2 + y^2 - 1 = eps with |eps| up to about 6e-8, which
// scales every probability the phase multiplies; among the floats within one
// ulp of (xd, yd) this picks the pair of smallest |eps| (ties: nearest to
// (xd, yd)).
static void unit_pair(double xd, double yd, float* x, float* y) {
  const float x0 = (float)xd, y0 = (float)yd;
  float bx = x0, by = y0;
  double best = 1e300, best_dist = 1e300;
  for (int i = -1; i <= 1; ++i)
    for (int j = -1; j <= 1; ++j) {
      float cx = x0, cy = y0;
      if (i) cx = std::nextafter(x0, i > 0 ? 2.0f : -2.0f);
      if (j) cy = std::nextafter(y0, j > 0 ? 2.0f : -2.0f);
      const double eps = (double)cx * cx + (double)cy * cy - 1.0;
      const double score = std::abs(eps);
      const double dist = std::abs(cx - xd) + std::abs(cy - yd);
      if (score < best - 1e-18 || (std::abs(score - best) <= 1e-18 && dist < best_dist)) {
        best = score;
        best_dist = dist;
        bx = cx;
        by = cy;
      }
    }
  *x = bx;
  *y = by;
}


extern "C" int rgs_gate_shear(const double* u, int n, int gate_limit, float* tan, float* dphase_lut, double* d1,
                              uint64_t* flip) {
  if (!u || n < 1 || n > 63) return fail(kInvalid, "gate_shear: need 1 <= n <= 63");
  using cd = std::complex<double>;
  std::vector<double> b(n, 0.0);
  uint64_t fl = 0;
  for (int q = 0; q < n; ++q) {
    const double* e = u + 8 * q;
    cd u00(e[0], e[1]), u01(e[2], e[3]), u10(e[4], e[5]), u11(e[6], e[7]);
    double c = std::abs(u00), s = std::abs(u10);
    double nrm = std::hypot(c, s);
    if (std::abs(nrm - 1.0) > 1e-9 || std::abs(std::abs(u01) - s) > 1e-9 || std::abs(std::abs(u11) - c) > 1e-9)
      return fail(kInvalid, "gate_shear: matrix is not unitary");
    c /= nrm;
    s /= nrm;
    // u = D1 R D2 with R = [[c, -s], [s, c]], D1 = diag(e^{i a0}, e^{i a1}),
    // D2 = diag(1, e^{i b})
    double a0, a1, bb;
    if (s < 1e-300) {           // diagonal
      a0 = std::arg(u00);
      a1 = std::arg(u11);
      bb = 0.0;
      c = 1.0;
      s = 0.0;
    } else if (c < 1e-300) {    // anti-diagonal: u01 = -e^{i(a0 + b)}, u10 = e^{i a1}
      a0 = 0.0;
      a1 = std::arg(u10);
      bb = std::arg(-u01);
      c = 0.0;
      s = 1.0;
    } else {
      a0 = std::arg(u00);
      a1 = std::arg(u10);
      bb = std::arg(u11) - a1;
    }
    // c >= s: R = c M(s / c).  s > c: R = X Z s M(-c / s) with X Z =
    // [[0, -1], [1, 0]], and D1 X = X diag(e^{i a1}, e^{i a0}), so
    // u = X diag(e^{i a1}, -e^{i a0}) s M(-c / s) D2.
    const bool swap = q < gate_limit && s > c;
    const double t = swap ? -c / s : (c > 0.0 ? s / c : 0.0);
    if (tan) tan[q] = (float)t;
    if (swap) fl |= 1ull << q;
    if (d1) {
      const double p0 = swap ? a1 : a0, p1 = swap ? a0 : a1, sg1 = swap ? -1.0 : 1.0;
      d1[4 * q + 0] = std::cos(p0);
      d1[4 * q + 1] = std::sin(p0);
      d1[4 * q + 2] = sg1 * std::cos(p1);
      d1[4 * q + 3] = sg1 * std::sin(p1);
    }
    b[q] = q < gate_limit ? bb : 0.0;
  }
  if (flip) *flip = fl;
  if (dphase_lut) {
    const int chunks = (n + 7) / 8;
    for (int ch = 0; ch < chunks; ++ch)
      for (int code = 0; code < 256; ++code) {
        double ang = 0.0;
        for (int t = 0; t < 8; ++t) {
          int q = 8 * ch + t;
          if (q < n && ((code >> t) & 1)) ang += b[q];
        }
        unit_pair(std::cos(ang), std::sin(ang), dphase_lut + 2 * (ch * 256 + code),
                  dphase_lut + 2 * (ch * 256 + code) + 1);
      }
  }
  return kOk;
}

extern "C" double rgs_gate_scale(const float* tan, int gate_limit) {
  double sc = 1.0;
  for (int q = 0; tan && q < gate_limit; ++q) sc /= std::sqrt(1.0 + (double)tan[q] * (double)tan[q]);
  return sc;
}

extern "C" int rgs_cov_weight_lut_bits(const double* u, int nv, const int32_t* B, const uint8_t* outB, int m,
                                       double scale, int chunk_bits, float* wlut) {
  if (!u || !B || !outB || !wlut || m < 1 || m > 63) return fail(kInvalid, "cov_weight_lut: need 1 <= m <= 63");
  if (chunk_bits < 8 || chunk_bits > 9) return fail(kInvalid, "cov_weight_lut: chunk_bits must be 8 or 9");
  const int cb = chunk_bits, size = 1 << cb;
  const int chunks = (m + cb - 1) / cb;
  for (int c = 0; c < chunks; ++c) {
    for (int code = 0; code < size; ++code) {
      std::complex<double> w = c == 0 ? std::complex<double>(scale, 0.0) : std::complex<double>(1.0, 0.0);
      for (int t = 0; t < cb; ++t) {
        int j = cb * c + t;
        if (j >= m) break;
        int site = B[j];
        if (site < 0 || site >= nv || outB[j] > 1) return fail(kInvalid, "bad site or output bit");
        int col = (code >> t) & 1;
        const double* e = u + 2 * (site * 4 + outB[j] * 2 + col);
        w *= std::complex<double>(e[0], e[1]);
      }
      wlut[2 * (c * size + code)] = (float)w.real();
      wlut[2 * (c * size + code) + 1] = (float)w.imag();
    }
  }
  return kOk;
}

extern "C" int rgs_cov_weight_table(const double* u0, const uint8_t* outB, int m, double scale, float* table) {
  if (!u0 || !outB || !table || m < 1 || m > 63) return fail(kInvalid, "cov_weight_table: need 1 <= m <= 63");
  using cd = std::complex<double>;
  int r = 0;
  for (int j = 0; j < m; ++j) {
    if (outB[j] > 1) return fail(kInvalid, "cov_weight_table: outB must hold bits");
    r += outB[j];
  }
  // powers 0..m of each entry, by repeated products in complex128
  std::vector<cd> pw(4 * (m + 1));
  for (int e = 0; e < 4; ++e) {
    const cd a(u0[2 * e], u0[2 * e + 1]);
    pw[e * (m + 1)] = cd(1.0, 0.0);
    for (int k = 1; k <= m; ++k) pw[e * (m + 1) + k] = pw[e * (m + 1) + k - 1] * a;
  }
  const int st = m + 1;
  for (int k = 0; k <= m; ++k)
    for (int h = 0; h <= m; ++h) {
      cd w(0.0, 0.0);
      if (h >= std::max(0, r + k - m) && h <= std::min(r, k)) {
        const int n00 = m - r - k + h, n01 = k - h, n10 = r - h, n11 = h;
        w = scale * pw[n00] * pw[st + n01] * pw[2 * st + n10] * pw[3 * st + n11];
      }
      table[2 * (k * st + h)] = (float)w.real();
      table[2 * (k * st + h) + 1] = (float)w.imag();
    }
  return kOk;
}

extern "C" int rgs_cov_weight_lut(const double* u, int nv, const int32_t* B, const uint8_t* outB, int m,
                                  double scale, float* wlut) {
  return rgs_cov_weight_lut_bits(u, nv, B, outB, m, scale, 8, wlut);
}
