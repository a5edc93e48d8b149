// Host side of the linear rank-width scan: plan compiler, cut-rank profiles,
// the order annealer, and the packed device plan.
//
// Positions are bits of 128-bit masks (n <= 128).  Step j sums the vertex at
// position j; the environment is indexed by coordinates over a basis of the
// syndrome space (parities of processed neighbours of each later position).
// Generators of the next space are the current basis with bit j cleared and
// the vertex's later-neighbour mask; echelon reduction with highest-bit
// pivots gives the new basis, a right inverse (lcols) and a kernel basis of
// the map p = (c, y) -> c'.  The packed device plan stores, per step,
// p0 = L c' split into a 64-entry table for the low six bits of c' and a
// table for the remaining bits, plus all kernel combinations.
#include <stdint.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <random>
#include <string>
#include <thread>
#include <vector>

#include "rgs_amd.h"
#include "rgs_status.h"

namespace rgs {

using u128 = unsigned __int128;

static inline int msb128(u128 x) {
  uint64_t hi = (uint64_t)(x >> 64);
  return hi ? 127 - __builtin_clzll(hi) : 63 - __builtin_clzll((uint64_t)x);
}

static bool check_graph(const uint8_t* g, int n, std::string& err) {
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j)
      if ((g[i * n + j] & 1) != (g[j * n + i] & 1) || (i == j && (g[i * n + j] & 1))) {
        err = "graph must be symmetric with zero diagonal";
        return false;
      }
  return true;
}

static bool check_order(const int32_t* order, int n, std::vector<int>& pos, std::string& err) {
  pos.assign(n, -1);
  for (int j = 0; j < n; ++j) {
    if (order[j] < 0 || order[j] >= n || pos[order[j]] >= 0) {
      err = "order must be a permutation of 0..n-1";
      return false;
    }
    pos[order[j]] = j;
  }
  return true;
}

// fut[j]: later neighbours of the vertex at position j, as position bits.
static std::vector<u128> future_masks(const uint8_t* g, int n, const int32_t* order, const std::vector<int>& pos) {
  std::vector<u128> fut(n, 0);
  for (int j = 0; j < n; ++j)
    for (int w = 0; w < n; ++w)
      if ((g[order[j] * n + w] & 1) && pos[w] > j) fut[j] |= (u128)1 << pos[w];
  return fut;
}

int compile(const uint8_t* g, int n, const int32_t* order, rgs_scan_step* steps, int32_t* profile,
            std::string& err) {
  std::vector<int> pos;
  if (!check_graph(g, n, err) || !check_order(order, n, pos, err)) return kInvalid;
  std::vector<u128> fut = future_masks(g, n, order, pos);
  std::vector<u128> basis;
  for (int j = 0; j < n; ++j) {
    rgs_scan_step& st = steps[j];
    std::memset(&st, 0, sizeof st);
    st.v = order[j];
    st.r_in = (int32_t)basis.size();
    if (st.r_in > RGS_SCAN_MAX_R) {
      err = "cut rank exceeds the scan limit";
      return kInvalid;
    }
    for (size_t i = 0; i < basis.size(); ++i) st.beta |= (uint64_t)((basis[i] >> j) & 1) << i;
    std::vector<u128> gens;
    const u128 clear = ~((u128)1 << j);
    for (auto b : basis) gens.push_back(b & clear);
    gens.push_back(fut[j]);
    std::vector<u128> nb;
    std::vector<int> piv;
    std::vector<uint64_t> coords;
    std::vector<int> creator;
    for (size_t i = 0; i < gens.size(); ++i) {
      u128 w = gens[i];
      uint64_t coord = 0;
      for (size_t k = 0; k < nb.size(); ++k)
        if ((w >> piv[k]) & 1) {
          w ^= nb[k];
          coord |= 1ull << k;
        }
      if (w) {
        piv.push_back(msb128(w));
        nb.push_back(w);
        coord |= 1ull << (nb.size() - 1);
        creator.push_back((int)i);
      }
      coords.push_back(coord);
    }
    const int r_out = (int)nb.size();
    st.r_out = r_out;
    for (int k = 0; k < r_out; ++k) {
      int i = creator[k];
      uint64_t p = 1ull << i;
      uint64_t rest = coords[i] & ~(1ull << k);
      for (int l = 0; l < k; ++l)
        if ((rest >> l) & 1) p ^= st.lcols[l];
      st.lcols[k] = p;
    }
    std::vector<char> is_creator(gens.size(), 0);
    for (int i : creator) is_creator[i] = 1;
    int nk = 0;
    for (size_t i = 0; i < gens.size(); ++i) {
      if (is_creator[i]) continue;
      uint64_t kv = 1ull << i;
      for (int k = 0; k < r_out; ++k)
        if ((coords[i] >> k) & 1) kv ^= st.lcols[k];
      st.kern[nk++] = kv;
    }
    st.nker = nk;
    if (profile) profile[j] = r_out;
    basis.swap(nb);
  }
  return kOk;
}

// Rank over GF(2) of the cut (prefix | rest) by elimination on 128-bit rows.
static int cut_rank(const std::vector<u128>& adj, u128 left, int n) {
  u128 full = n == 128 ? ~(u128)0 : (((u128)1 << n) - 1);
  u128 right = full & ~left;
  std::vector<u128> piv(128, 0);
  int rank = 0;
  u128 todo = left;
  while (todo) {
    int v = msb128(todo);
    todo &= ~((u128)1 << v);
    u128 x = adj[v] & right;
    while (x) {
      int p = msb128(x);
      if (piv[p]) x ^= piv[p];
      else {
        piv[p] = x;
        ++rank;
        break;
      }
    }
  }
  return rank;
}

static std::vector<u128> adjacency(const uint8_t* g, int n) {
  std::vector<u128> adj(n, 0);
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j)
      if (g[i * n + j] & 1) adj[i] |= (u128)1 << j;
  return adj;
}

static void profile_of(const std::vector<u128>& adj, const std::vector<int>& order, int n, std::vector<int>& prof,
                       int lo = 0) {
  prof.resize(n);
  u128 prefix = 0;
  for (int k = 0; k < lo; ++k) prefix |= (u128)1 << order[k];
  for (int k = lo; k < n; ++k) {
    prefix |= (u128)1 << order[k];
    prof[k] = (k == n - 1) ? 0 : cut_rank(adj, prefix, n);
  }
}

static uint64_t cost_of(const std::vector<int>& prof) {
  uint64_t c = 0;
  for (int r : prof) c += 1ull << r;
  return c;
}

}  // namespace rgs

using namespace rgs;

extern "C" int rgs_scan_compile(const uint8_t* g, int n, const int32_t* order, rgs_scan_step* steps,
                                int32_t* profile) {
  if (!g || !order || !steps || n < 1 || n > 128) return fail(kInvalid, "scan_compile: need 1 <= n <= 128");
  std::string err;
  int rc = compile(g, n, order, steps, profile, err);
  return rc ? fail(rc, err) : kOk;
}

extern "C" int rgs_scan_profile(const uint8_t* g, int n, const int32_t* order, int32_t* profile) {
  if (!g || !order || !profile || n < 1 || n > 128) return fail(kInvalid, "scan_profile: need 1 <= n <= 128");
  std::string err;
  std::vector<int> pos;
  if (!check_graph(g, n, err) || !check_order(order, n, pos, err)) return fail(kInvalid, err);
  std::vector<int> ord(order, order + n), prof;
  profile_of(adjacency(g, n), ord, n, prof);
  for (int k = 0; k < n; ++k) profile[k] = prof[k];
  return kOk;
}

// One annealing chain over orders.  A move permutes the positions i..j only,
// so the prefix sets before i and from j on are unchanged and only the ranks
// of the cuts i..j-1 are recomputed.  Moves: swap of neighbours, insertion
// and reversal within a short window, and a swap of two arbitrary positions.
// The chain minimises sum 2^r with the cuts above the rank cap weighted 64x,
// so a start above the cap is driven under it; once under, it stays.  The
// run is split into epochs; each restarts from the chain's best order and
// cools geometrically, with the temperature relative to the current cost (so
// the same schedule serves any graph), and later epochs start colder.  The
// chain is a pure function of (graph, start, iterations, seed, cap).
namespace rgs {

struct Chain {
  std::vector<int> order, rank;
  std::vector<u128> prefix;
  uint64_t cost = 0;   // sum of 2^r
  uint64_t eff = 0;    // the same with cuts above the cap weighted 64x: what the chain minimises
  int peak = 0;
};

static inline uint64_t eff_term(int r, int cap) { return (uint64_t)(r > cap ? 64 : 1) << r; }

static int cut_rank_small_side(const std::vector<u128>& adj, u128 left, int n, int left_count) {
  u128 full = n == 128 ? ~(u128)0 : (((u128)1 << n) - 1);
  // the cut rank is symmetric: eliminate over the smaller side
  return 2 * left_count > n ? cut_rank(adj, full & ~left, n) : cut_rank(adj, left, n);
}

static void chain_init(const std::vector<u128>& adj, const std::vector<int>& order, int n, int cap, Chain& c) {
  c.order = order;
  c.prefix.assign(n, 0);
  c.rank.assign(n, 0);
  u128 p = 0;
  for (int k = 0; k < n; ++k) {
    p |= (u128)1 << order[k];
    c.prefix[k] = p;
    c.rank[k] = k == n - 1 ? 0 : cut_rank_small_side(adj, p, n, k + 1);
  }
  c.cost = cost_of(c.rank);
  c.eff = 0;
  for (int r : c.rank) c.eff += eff_term(r, cap);
  c.peak = *std::max_element(c.rank.begin(), c.rank.end());
}

static void run_chain(const std::vector<u128>& adj, const std::vector<int>& start, int n, uint64_t iterations,
                      uint64_t seed, int rmax_cap, Chain& best) {
  Chain cur;
  chain_init(adj, start, n, rmax_cap, cur);
  best = cur;
  bool best_ok = cur.peak <= rmax_cap;
  std::mt19937_64 rng(seed);
  std::uniform_real_distribution<double> unif(0.0, 1.0);
  const int epochs = 6;
  const uint64_t per_epoch = std::max<uint64_t>(1, iterations / epochs);
  std::vector<int> cand_order(n);
  std::vector<int> new_rank(n);
  std::vector<u128> new_prefix(n);
  for (uint64_t it = 0; it < iterations; ++it) {
    const uint64_t e = it / per_epoch, within = it % per_epoch;
    if (within == 0 && e > 0 && best_ok) cur = best;
    const double phase = (double)within / (double)per_epoch;
    const double t0 = e == 0 ? 0.02 : 0.02 * std::pow(0.4, (double)e);
    const double temp = t0 * std::pow(1e-3, phase);
    int i, j;
    const int kind = (int)(rng() % 8);
    i = (int)(rng() % n);
    if (kind < 3) j = i + 1;                                        // neighbour swap
    else if (kind < 7) j = i + 1 + (int)(rng() % 9);                // short window
    else j = (int)(rng() % n);                                      // anywhere
    if (j >= n) j = n - 1;
    if (i == j) continue;
    if (i > j) std::swap(i, j);
    std::copy(cur.order.begin(), cur.order.end(), cand_order.begin());
    if (kind < 3 || kind == 7) std::swap(cand_order[i], cand_order[j]);
    else if (kind < 5) {
      if (rng() & 1) std::rotate(cand_order.begin() + i, cand_order.begin() + i + 1, cand_order.begin() + j + 1);
      else std::rotate(cand_order.begin() + i, cand_order.begin() + j, cand_order.begin() + j + 1);
    } else std::reverse(cand_order.begin() + i, cand_order.begin() + j + 1);
    // ranks of the cuts i..j-1 (the cut after position j holds the same set)
    u128 p = i ? cur.prefix[i - 1] : 0;
    int64_t diff = 0, ediff = 0;
    for (int k = i; k < j; ++k) {
      p |= (u128)1 << cand_order[k];
      new_prefix[k] = p;
      new_rank[k] = cut_rank_small_side(adj, p, n, k + 1);
      diff += (int64_t)(1ull << new_rank[k]) - (int64_t)(1ull << cur.rank[k]);
      ediff += (int64_t)eff_term(new_rank[k], rmax_cap) - (int64_t)eff_term(cur.rank[k], rmax_cap);
    }
    int peak = 0;
    for (int k = 0; k < n; ++k) peak = std::max(peak, (k >= i && k < j) ? new_rank[k] : cur.rank[k]);
    if (peak > rmax_cap && cur.peak <= rmax_cap) continue;
    bool take = ediff <= 0 || unif(rng) < std::exp(-(double)ediff / (temp * (double)cur.eff));
    if (!take) continue;
    for (int k = i; k <= j; ++k) cur.order[k] = cand_order[k];
    for (int k = i; k < j; ++k) {
      cur.prefix[k] = new_prefix[k];
      cur.rank[k] = new_rank[k];
    }
    cur.cost = (uint64_t)((int64_t)cur.cost + diff);
    cur.eff = (uint64_t)((int64_t)cur.eff + ediff);
    cur.peak = peak;
    if (peak <= rmax_cap && (!best_ok || cur.cost < best.cost)) {
      best = cur;
      best_ok = true;
    }
  }
  if (!best_ok) best.cost = ~0ull;
}

}  // namespace rgs

extern "C" int rgs_scan_order_search(const uint8_t* g, int n, const int32_t* start, uint64_t iterations,
                                     uint64_t seed, int rmax_cap, int threads, int32_t* best_order,
                                     uint64_t* best_cost) {
  if (!g || !start || !best_order || n < 2 || n > 128) return fail(kInvalid, "order_search: need 2 <= n <= 128");
  if (threads < 1 || threads > 256) return fail(kInvalid, "order_search: threads must be 1..256");
  std::string err;
  std::vector<int> pos;
  if (!check_graph(g, n, err) || !check_order(start, n, pos, err)) return fail(kInvalid, err);
  const std::vector<u128> adj = adjacency(g, n);
  const std::vector<int> st(start, start + n);
  std::vector<Chain> best(threads);
  std::vector<std::thread> pool;
  for (int t = 0; t < threads; ++t)
    pool.emplace_back([&, t]() {
      run_chain(adj, st, n, iterations, seed + 0x9E3779B97F4A7C15ull * (uint64_t)t, rmax_cap, best[t]);
    });
  for (auto& th : pool) th.join();
  int win = 0;
  for (int t = 1; t < threads; ++t)
    if (best[t].cost < best[win].cost) win = t;          // ties: the lowest chain
  if (best[win].cost == ~0ull) return fail(kInvalid, "no order within the rank cap found");
  for (int k = 0; k < n; ++k) best_order[k] = best[win].order[k];
  if (best_cost) *best_cost = best[win].cost;
  return kOk;
}

// Packed device plan (uint64 words):
//   [0] n  [1] rmax  [2 .. 2+n) offset of each step
//   step: hdr = v | r_in << 8 | r_out << 16 | nker << 24 | split << 32 |
//               uniform << 33, beta,
//         lo[64]  (p0 of the low six bits of c'),
//         hi[2^max(0, r_out-6)] (p0 of the remaining bits),
//         combos[2^nker]
// The kernel basis is first reduced on the y bit (bit r_in of p): at most one
// basis vector keeps it, and that one (kpiv) goes last.  When one does
// (split = 1), every lo/hi entry with the y bit set is moved by kpiv, so each
// p0 = lo ^ hi is the member of its coset p0 + kernel with y = 0; the
// combinations then hold y = 0 in their first half and y = 1 in their second,
// and every output is f0 * (sum of its first-half entries) + f1 * (signed sum
// of its second half), with no per-lane choice of weight.  When none does
// (split = 0), all terms of an output share y(p0).  The set of terms of each
// output, and so its sum, is unchanged; only the order of summation is.
//
// Table coordinates.  Each environment table may be indexed in any basis of
// its syndrome space: a change of coordinates T at the output of step j
// (c'new = T c') is undone at the input of step j+1, whose p-vectors, kernel
// and beta are rewritten in the new coordinates (p = (c, y) -> (T c, y),
// beta -> T^-T beta).  Two linear forms decide an output's weight and signs:
// y(p0(c')) and the sign parity(p0(c') & mask & beta).  The packer picks, for
// every step with enough outputs, coordinates in which both forms vanish on
// the six lane bits (a basis of their common kernel first, pivots taken at
// the highest coordinates so the low bits keep their meaning), so y and the
// signs depend on the 64-output block only (uniform = 1): the kernel then
// decides them once per block instead of once per output and term.
namespace rgs {

static inline int parity64(uint64_t v) { return __builtin_popcountll(v) & 1; }

// Coordinates for one step's outputs: returns the old coordinates of the new
// basis vectors (columns) and whether the first min(6, r) are in the common
// kernel of the forms f1, f2 (bit masks over old coordinates).
static bool choose_coordinates(int r, uint64_t f1, uint64_t f2, uint64_t* inv_cols) {
  uint64_t v[64];
  bool used[64] = {false};
  for (int k = 0; k < r; ++k) v[k] = 1ull << k;
  for (uint64_t f : {f1, f2}) {
    int q = -1;
    for (int k = r - 1; k >= 0; --k)
      if (!used[k] && parity64(v[k] & f)) {
        q = k;
        break;
      }
    if (q < 0) continue;
    for (int k = 0; k < r; ++k)
      if (k != q && !used[k] && parity64(v[k] & f)) v[k] ^= v[q];
    used[q] = true;
  }
  int m = 0;
  for (int k = 0; k < r; ++k)
    if (!used[k]) inv_cols[m++] = v[k];
  const int free_count = m;
  for (int k = 0; k < r; ++k)
    if (used[k]) inv_cols[m++] = v[k];
  return free_count >= std::min(6, r);
}

// Inverse of an r x r GF(2) matrix given by columns.
static void gf2_invert_cols(int r, const uint64_t* cols, uint64_t* inv) {
  // rows of [A | I] by Gauss-Jordan on columns: solve A x = e_i for each i
  uint64_t a[64], b[64];
  for (int i = 0; i < r; ++i) {
    a[i] = 0;
    for (int k = 0; k < r; ++k) a[i] |= ((cols[k] >> i) & 1ull) << k;   // row i of A
    b[i] = 1ull << i;                                                  // row i of I
  }
  for (int c = 0; c < r; ++c) {
    int p = c;
    while (p < r && !((a[p] >> c) & 1ull)) ++p;
    std::swap(a[c], a[p]);
    std::swap(b[c], b[p]);
    for (int i = 0; i < r; ++i)
      if (i != c && ((a[i] >> c) & 1ull)) {
        a[i] ^= a[c];
        b[i] ^= b[c];
      }
  }
  // rows of A^-1 are b; convert to columns
  for (int k = 0; k < r; ++k) {
    inv[k] = 0;
    for (int i = 0; i < r; ++i) inv[k] |= ((b[i] >> k) & 1ull) << i;
  }
}

// v (an old-coordinate vector of r bits) in new coordinates: sum of the
// columns of T selected by v's bits.
static inline uint64_t apply_cols(const uint64_t* cols, uint64_t v) {
  uint64_t out = 0;
  for (int i = 0; v; ++i, v >>= 1)
    if (v & 1ull) out ^= cols[i];
  return out;
}

}  // namespace rgs

extern "C" int rgs_scan_pack(const rgs_scan_step* steps, int n, uint64_t* blob, uint64_t* words) {
  if (!steps || n < 1 || n > 128 || !words) return fail(kInvalid, "scan_pack: bad arguments");
  uint64_t w = 2 + (uint64_t)n;
  int rmax = 0;
  uint64_t t_in[64];   // input coordinate change of the current step: columns T e_i
  for (int j = 0; j < n; ++j) {
    const rgs_scan_step& st = steps[j];
    // A step clears one coordinate of r_in basis vectors (rank drops by at
    // most one) and adds one vector, so r_out >= r_in - 1 and the kernel of
    // p = (c, y) -> c' has dimension r_in + 1 - r_out <= 2; the batch kernel
    // has straight-line paths for nker = 0, 1, 2 only.
    if (st.r_out > 28 || st.nker > 2) return fail(kInvalid, "scan_pack: rank above 28 or kernel dimension above 2");
    rmax = std::max(rmax, st.r_out);
    uint64_t hi = 1ull << std::max(0, st.r_out - 6);
    uint64_t off = w;
    w += 2 + 64 + hi + (1ull << st.nker);
    const uint64_t ybit = 1ull << st.r_in, cmask = ybit - 1ull;
    // p-space vectors of this step in the input's new coordinates
    auto to_new_p = [&](uint64_t p) { return apply_cols(t_in, p & cmask) | (p & ybit); };
    uint64_t t_in_inv[64];
    if (j == 0) {
      for (int i = 0; i < 64; ++i) t_in[i] = t_in_inv[i] = 1ull << i;
    } else {
      gf2_invert_cols(st.r_in, t_in, t_in_inv);
    }
    // beta in new input coordinates: bit k = <beta, T^-1 e_k>
    uint64_t beta = 0;
    for (int k = 0; k < st.r_in; ++k) beta |= (uint64_t)parity64(st.beta & t_in_inv[k]) << k;
    // the two forms on the step's output coordinates
    uint64_t fy = 0, fs = 0;
    for (int k = 0; k < st.r_out; ++k) {
      fy |= ((st.lcols[k] >> st.r_in) & 1ull) << k;
      fs |= (uint64_t)parity64(st.lcols[k] & cmask & st.beta) << k;
    }
    uint64_t out_inv[64];
    const bool uniform = st.r_out >= 8 && choose_coordinates(st.r_out, fy, fs, out_inv);
    if (!uniform)
      for (int k = 0; k < st.r_out; ++k) out_inv[k] = 1ull << k;
    // lcols in both new coordinates: L^ e_k = P L (T_out^-1 e_k)
    uint64_t lcols[64];
    for (int k = 0; k < st.r_out; ++k) lcols[k] = to_new_p(apply_cols(st.lcols, out_inv[k]));
    if (blob) {
      blob[2 + j] = off;
      uint64_t* p = blob + off;
      p[0] = (uint64_t)st.v | ((uint64_t)st.r_in << 8) | ((uint64_t)st.r_out << 16) | ((uint64_t)st.nker << 24) |
             ((uint64_t)uniform << 33);
      p[1] = beta;
      uint64_t kern[64];
      for (int k = 0; k < st.nker; ++k) kern[k] = to_new_p(st.kern[k]);
      int piv = -1;
      for (int k = 0; k < st.nker; ++k)
        if (kern[k] & ybit) {
          if (piv < 0) piv = k;
          else kern[k] ^= kern[piv];
        }
      if (piv >= 0) std::swap(kern[piv], kern[st.nker - 1]);
      const uint64_t kpiv = piv >= 0 ? kern[st.nker - 1] : 0;
      p[0] |= (uint64_t)(piv >= 0) << 32;
      for (int c = 0; c < 64; ++c) {
        uint64_t x = 0;
        for (int i = 0; i < 6 && i < st.r_out; ++i)
          if ((c >> i) & 1) x ^= lcols[i];
        p[2 + c] = (x & ybit) ? x ^ kpiv : x;
      }
      for (uint64_t k = 0; k < hi; ++k) {
        uint64_t x = 0;
        for (int i = 6; i < st.r_out; ++i)
          if ((k >> (i - 6)) & 1) x ^= lcols[i];
        p[66 + k] = (x & ybit) ? x ^ kpiv : x;
      }
      uint64_t* cb = p + 66 + hi;
      cb[0] = 0;
      for (int k = 0; k < st.nker; ++k)
        for (uint64_t t = 0; t < (1ull << k); ++t) cb[(1ull << k) + t] = cb[t] ^ kern[k];
    }
    // the next step's input coordinates are this step's output coordinates
    gf2_invert_cols(st.r_out, out_inv, t_in);
  }
  if (blob) {
    blob[0] = (uint64_t)n;
    blob[1] = (uint64_t)rmax;
  }
  *words = w;
  return kOk;
}

// Byte-table form of a plan for the HBM-table scan: per step a header
// (v | r_in << 8 | r_out << 16 | nker << 24 | nbytes << 32), beta, nbytes x 256
// words (word [b][k] = XOR of the lcols selected by bits of byte b = k) and
// the 2^nker kernel combinations.
extern "C" int rgs_scan_pack_tables(const rgs_scan_step* steps, int n, uint64_t* blob, uint64_t* words) {
  if (!steps || n < 1 || n > 128 || !words) return fail(kInvalid, "scan_pack_tables: bad arguments");
  uint64_t w = 2 + (uint64_t)n;
  int rmax = 0;
  for (int j = 0; j < n; ++j) {
    const rgs_scan_step& st = steps[j];
    if (st.r_out > 32 || st.r_in > 32 || st.nker > 16 || st.r_out < 0 || st.r_in < 0 || st.nker < 0)
      return fail(kInvalid, "scan_pack_tables: rank above 32 or kernel above 16");
    rmax = std::max(rmax, std::max(st.r_out, st.r_in));
    const int nbytes = (st.r_out + 7) / 8;
    const uint64_t off = w;
    w += 2 + 256ull * nbytes + (1ull << st.nker);
    if (blob) {
      blob[2 + j] = off;
      uint64_t* p = blob + off;
      p[0] = (uint64_t)st.v | ((uint64_t)st.r_in << 8) | ((uint64_t)st.r_out << 16) | ((uint64_t)st.nker << 24) |
             ((uint64_t)nbytes << 32);
      p[1] = st.beta;
      for (int b = 0; b < nbytes; ++b)
        for (int k = 0; k < 256; ++k) {
          uint64_t x = 0;
          for (int i = 0; i < 8 && 8 * b + i < st.r_out; ++i)
            if ((k >> i) & 1) x ^= st.lcols[8 * b + i];
          p[2 + 256 * b + k] = x;
        }
      uint64_t* cb = p + 2 + 256ull * nbytes;
      cb[0] = 0;
      for (int k = 0; k < st.nker; ++k)
        for (uint64_t t = 0; t < (1ull << k); ++t) cb[(1ull << k) + t] = cb[t] ^ st.kern[k];
    }
  }
  if (blob) {
    blob[0] = (uint64_t)n;
    blob[1] = (uint64_t)rmax;
  }
  *words = w;
  return kOk;
}
