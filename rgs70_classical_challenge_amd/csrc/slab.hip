// Slab pipeline on one 2^nbits complex64 state, host side: the pass planner,
// the per-pass kernel arguments and the C ABI.  The tile kernel and its
// layouts are described in slab_kernel.h.
//
// Gates.  Every 2x2 unitary is factored on the host as u = D1 R D2 (real
// rotation R, diagonals D1 and D2 = diag(1, e^{ib})).  Because diagonals on
// different qubits commute with everything on other qubits, the product of all
// gates is (x)D1 . (x)R . (x)D2: the D2 layer is one pointwise phase
// e^{i sum_q b_q x_q} applied with the graph phase, the rotations cost four
// FMAs per amplitude and qubit, and the D1 layer, which leaves every
// probability unchanged, is applied when amplitudes are read back.
//
// Schedule.  The low pass holds global bits 0..low-1; the high bits are split
// into strided groups of at most tile_bits - lane_bits bits.  With the Walsh
// transform: low[init, wht], groups up [wht], the top group [wht, phase,
// gates], groups down [gates], low[gates, mass].  Without it: the first group
// [init, phase, gates], the other groups [gates], low[gates, mass].
#include "slab_kernel.h"
#include "slab_plan.h"

#include <algorithm>
#include <cstdio>
#include <string>
#include <vector>

namespace rgs {

static int chunks_of(int bits) { return (bits + 7) / 8; }

size_t lds_bytes(const PassPlan& p, int nbits, const rgs_slab_desc* d) {
  size_t slots = ((size_t)1 << (p.tb - p.rb)) + ((size_t)1 << (p.tb - p.rb - 4));
  size_t bytes = slots * 8;
  // LDS tables: the phase pass's diagonal-phase table (unless it is read from
  // global memory) and the covariant init's weights; sign and map tables are
  // read from global memory
  if ((p.flags & F_PHASE) && !(p.flags & F_LAYER) && d->dphase_lut && !d->dphase_global)
    bytes += (size_t)chunks_of(nbits) * 256 * 8;
  if (!(p.flags & F_LOAD) && d->init_kind == RGS_INIT_COV) {
    const int wb = d->init_wbits ? d->init_wbits : 8;
    // the popcount table as staged: (|o| + 1) (m - |o| + 1) entries, at most
    // (floor(m/2) + 1) (ceil(m/2) + 1) whatever the outcome bits
    bytes += d->init_wpop ? (size_t)(nbits / 2 + 1) * ((nbits + 1) / 2 + 1) * 8
                          : (size_t)((nbits + wb - 1) / wb) * ((size_t)1 << wb) * 8;
  }
  return bytes;
}

int plan_passes(const rgs_slab_desc* d, std::vector<PassPlan>& out, std::string& err) {
  const int n = d->nbits;
  if (n < 6 || n > 40) { err = "nbits must be in 6..40"; return kInvalid; }
  int T0 = std::min(d->low_bits, n);
  if (T0 < 6 || T0 > 14) { err = "low_bits must be in 6..14"; return kInvalid; }
  const int TS = d->tile_bits;
  const int L = d->lane_bits;
  if (d->round_bits < 0 || d->round_bits > 2) { err = "round_bits must be in 0..2"; return kInvalid; }
  if (d->phase_round_bits < -1 || d->phase_round_bits > 2) { err = "phase_round_bits must be in -1..2"; return kInvalid; }
  if (d->fuse_walsh != 0 && d->fuse_walsh != 1) { err = "fuse_walsh must be 0 or 1"; return kInvalid; }
  if (d->fuse_walsh && (!d->do_walsh || d->phase_lut)) {
    err = "fuse_walsh needs the Walsh transform and no sign table after it (phase_lut)";
    return kInvalid;
  }
  if (d->fuse_walsh && d->dphase_lut && !d->dphase_host) {
    err = "fuse_walsh with dphase_lut needs its host copy (dphase_host)";
    return kInvalid;
  }
  std::vector<PassPlan> groups;
  if (d->group_bits[0] > 0) {
    // explicit groups: each with its own width, row width and LDS rounds
    int q = T0;
    for (int i = 0; i < 8 && d->group_bits[i] > 0; ++i) {
      PassPlan g{};
      g.low = 0;
      g.qstart = q;
      g.count = d->group_bits[i];
      g.c = d->group_lanes[i] ? d->group_lanes[i] : L;
      g.tb = g.c + g.count;
      g.rbits = 4;
      if (g.count < g.rbits) { err = "explicit strided group narrower than a register window"; return kInvalid; }
      if (g.c < 2 || g.c > 7) { err = "explicit group lane bits must be in 2..7"; return kInvalid; }
      if (g.tb > 14) { err = "explicit group lane bits plus width exceed 14 tile bits"; return kInvalid; }
      if (g.c > q) { err = "explicit group lane bits overlap its target bits"; return kInvalid; }
      const int r = d->group_rounds[i];
      if (r < 0 || r > 2) { err = "explicit group round bits must be in 0..2"; return kInvalid; }
      g.rb = std::max(0, std::min(std::min(r, g.c), g.tb - 10));
      groups.push_back(g);
      q += g.count;
    }
    if (q != n) { err = "explicit strided groups must cover bits low_bits..nbits-1 exactly"; return kInvalid; }
  } else if (n > T0) {
    if (TS < 6 || TS > 14) { err = "tile_bits must be in 6..14"; return kInvalid; }
    if (L < 2 || L > 7) { err = "lane_bits must be in 2..7"; return kInvalid; }
    const int RBs = 4;
    if (L < 1 || TS - L < RBs) { err = "lane_bits must leave at least 4 target bits"; return kInvalid; }
    if (n - T0 < RBs) T0 = n - RBs;  // a strided group needs a full register window
    if (T0 < 6) { err = "state too small for a strided pass; raise low_bits"; return kInvalid; }
    const int F = TS - L;
    const int rem = n - T0;
    // group widths: as even as the tile allows, or a phase group of exactly
    // top_bits bits and the rest even
    const int top = d->top_bits;
    if (top < 0 || (top > 0 && (top < RBs || top > F))) {
      err = "top_bits must be 0 or in 4..tile_bits-lane_bits";
      return kInvalid;
    }
    std::vector<int> widths;
    const int even = top > 0 && rem > top ? rem - top : rem;
    const int ng = (even + F - 1) / F;
    for (int i = 0; i < ng; ++i) widths.push_back(even / ng + (i < even % ng ? 1 : 0));
    if (even != rem) {
      // the phase group: the last walked with the Walsh transform (highest
      // bits, or lowest with group_order 1), the first without
      if (d->do_walsh && !d->group_order)
        widths.push_back(top);
      else
        widths.insert(widths.begin(), top);
    }
    int q = T0;
    for (int cnt : widths) {
      PassPlan g{};
      g.low = 0;
      g.qstart = q;
      g.count = cnt;
      g.c = L;
      g.tb = L + cnt;
      g.rbits = 4;
      if (cnt < g.rbits) { err = "strided group narrower than a register window"; return kInvalid; }
      if (g.c > q) { err = "lane bits overlap the target bits; lower lane_bits"; return kInvalid; }
      // LDS rounds: at most the lane bits, and the thread bits below the
      // round bits must span a wavefront (tile bits - 4 - rb >= 6)
      g.rb = std::max(0, std::min(std::min(d->round_bits, g.c), g.tb - 10));
      if (g.tb < 6) { err = "strided tile narrower than 6 bits; raise lane_bits"; return kInvalid; }
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
  } else if (walsh && d->fuse_walsh) {
    // every pass takes the Walsh butterflies, the diagonal phases and the
    // rotations of its own bits (operations on different qubits commute):
    // the first group makes the state, the contiguous pass ends with the masses
    const int own = F_LAYER | F_WHT | (phase ? F_PHASE : 0) | (gates ? F_GATES : 0);
    for (int i = 0; i < k; ++i) out.push_back(with(groups[d->group_order ? k - 1 - i : i], (i ? F_LOAD : src) | own));
    out.push_back(with(low, F_LOAD | own | (mass ? F_MASS : 0)));
  } else if (walsh) {
    // the Walsh groups may run in either order (the transform factors over
    // bits); the last one also applies the phase and its gates, and the
    // others are walked back for their gates
    std::vector<int> ord(k);
    for (int i = 0; i < k; ++i) ord[i] = d->group_order ? k - 1 - i : i;
    out.push_back(with(low, src | F_WHT));
    for (int i = 0; i + 1 < k; ++i) out.push_back(with(groups[ord[i]], F_LOAD | F_WHT));
    out.push_back(with(groups[ord[k - 1]], F_LOAD | F_WHT | (phase ? F_PHASE : 0) | (gates ? F_GATES : 0)));
    if (gates)
      for (int i = k - 2; i >= 0; --i) out.push_back(with(groups[ord[i]], F_LOAD | F_GATES));
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
  for (size_t i = 0; i < out.size(); ++i) out[i].first = i == 0;
  if (d->defer_leaves) {
    // the last pass writes its masses only; it must load the state (the
    // earlier passes' values stay in place for its tiles to be recomputed)
    PassPlan& lp = out.back();
    if (out.size() < 2 || !lp.low || !(lp.flags & F_LOAD) || !(lp.flags & F_MASS)) {
      err = "defer_leaves needs a mass level and a last contiguous pass that loads the state";
      return kInvalid;
    }
    lp.flags |= F_NOSTORE;
  }
  // the strided pass that applies the phase tables holds them in LDS besides
  // the tile: it may take its own number of exchange rounds
  if (d->phase_round_bits >= 0)
    for (auto& p : out)
      if (!p.low && (p.flags & F_PHASE)) p.rb = std::max(0, std::min(std::min(d->phase_round_bits, p.c), p.tb - 10));
  for (auto& p : out) {
    if (lds_bytes(p, n, d) > kLdsBudget) { err = "tile plus tables exceed 160 KiB of LDS; lower the tile bits"; return kInvalid; }
    if ((p.flags & F_MASS) && (d->mass_bits < p.rbits || d->mass_bits > std::min(p.tb, p.rbits + 6))) {
      err = "mass_bits must be in 4..min(low tile bits, 10)";
      return kInvalid;
    }
  }
  return kOk;
}


hipError_t launch_pass(const PassArgs& a, const PassPlan& p, size_t shmem, hipStream_t st) {
  switch (p.c) {
    case 0: return launch_c0(p.tb, p.rb, a, shmem, st);
    case 2: return launch_c2(p.tb, p.rb, a, shmem, st);
    case 3: return launch_c3(p.tb, p.rb, a, shmem, st);
    case 4: return launch_c4(p.tb, p.rb, a, shmem, st);
    case 5: return launch_c5(p.tb, p.rb, a, shmem, st);
    case 6: return launch_c6(p.tb, p.rb, a, shmem, st);
    case 7: return launch_c7(p.tb, p.rb, a, shmem, st);
    default: return hipErrorInvalidValue;
  }
}

static void form_tables(const uint64_t* rows, const int pos[4], uint64_t q[4], uint32_t* freg) {
  for (int i = 0; i < 4; ++i) q[i] = rows[pos[i]] & ~(1ull << pos[i]);
  uint32_t m = 0;
  for (int j = 0; j < 16; ++j) {
    uint32_t f = 0;
    for (int i = 0; i < 4; ++i) {
      if (!((j >> i) & 1)) continue;
      f ^= (uint32_t)(rows[pos[i]] >> pos[i]) & 1u;
      for (int k = i + 1; k < 4; ++k)
        if ((j >> k) & 1) f ^= (uint32_t)(rows[pos[i]] >> pos[k]) & 1u;
    }
    m |= f << j;
  }
  *freg = m;
}

// Global bit of local tile bit L.
static int global_bit(const PassPlan& p, int L) { return L < p.c ? L : (p.low ? L : p.qstart + L - p.c); }

// A sign form (symmetric rows, linear terms on the diagonal) restricted to the
// tile's local bits: up[L] = linear term at bit L and cross terms with the
// local bits above L; grow[L] = the global row of local bit L.
static void local_form(const PassPlan& p, const uint64_t* rows, LocalForm& F) {
  for (int L = 0; L < 16; ++L) {
    F.up[L] = 0;
    F.grow[L] = 0;
  }
  for (int L = 0; L < p.tb; ++L) {
    const int g = global_bit(p, L);
    F.grow[L] = rows[g];
    uint32_t up = (uint32_t)((rows[g] >> g) & 1u) << L;
    for (int L2 = L + 1; L2 < p.tb; ++L2) up |= (uint32_t)((rows[g] >> global_bit(p, L2)) & 1u) << L2;
    F.up[L] = up;
  }
}

// The per-layout diagonal phases of an F_LAYER pass: the layouts walk down
// from max(tb - 4, c) in steps of four bits to c, layout K owning local bits
// [s_K, s_{K-1}) (s_{-1} = tb); lph[K][m] is the product, in double and
// rounded once, of the single-bit phases e^{i b_q} (entry 1 << (q & 7) of
// chunk q >> 3 of the host table) of the owned bits set in m.
static void layer_phases(const rgs_slab_desc* d, const PassPlan& p, PassArgs& a) {
  const int s1 = std::max(p.tb - 4, p.c);
  int s = s1, h = p.tb;
  for (int K = 0; K < 4; ++K) {
    double re[16], im[16];
    re[0] = 1.0;
    im[0] = 0.0;
    for (int i = 0; i < 4; ++i) {
      double fr = 1.0, fi = 0.0;
      if (s + i < h) {
        const int q = global_bit(p, s + i);
        const float* e = d->dphase_host + 2 * (((size_t)(q >> 3) << 8) | (1u << (q & 7)));
        fr = e[0];
        fi = e[1];
      }
      for (int j = 1 << i; j < (2 << i); ++j) {
        const int b = j - (1 << i);
        re[j] = re[b] * fr - im[b] * fi;
        im[j] = re[b] * fi + im[b] * fr;
      }
    }
    for (int j = 0; j < 16; ++j) a.lph[K][j] = make_float2((float)re[j], (float)im[j]);
    if (s <= p.c) break;
    h = s;
    s = std::max(s - 4, p.c);
  }
}

int check_desc(const rgs_slab_desc* d, std::string& err) {
  const int n = d->nbits;
  if (d->init_kind < 0 || d->init_kind > 3) { err = "bad init_kind"; return kInvalid; }
  if ((d->init_kind == RGS_INIT_GRAPH || d->init_kind == RGS_INIT_COV) && !d->init_lut) {
    err = "init tables missing";
    return kInvalid;
  }
  if (d->init_kind == RGS_INIT_COV && !d->init_wlut) { err = "covariant weights missing"; return kInvalid; }
  if (d->init_kind == RGS_INIT_GRAPH && !d->init_rows) { err = "graph init needs its form rows"; return kInvalid; }
  if (d->init_kind == RGS_INIT_COV && (!d->init_rows || !d->init_cols)) {
    err = "covariant init needs its form rows and map columns";
    return kInvalid;
  }
  if (d->init_wbits != 0 && d->init_wbits != 8 && d->init_wbits != 9) { err = "init_wbits must be 8 or 9"; return kInvalid; }
  if (d->init_wpop != 0 && d->init_wpop != 1) { err = "init_wpop must be 0 or 1"; return kInvalid; }
  if (d->flip >> n) { err = "flip has bits at or above nbits"; return kInvalid; }
  if (d->init_wpop && (d->init_omask >> n)) { err = "init_omask has bits at or above nbits"; return kInvalid; }
  if (d->phase_lut && !d->phase_rows) { err = "phase_lut needs its form rows"; return kInvalid; }
  if ((d->phase_lut || d->dphase_lut || d->init_kind == RGS_INIT_GRAPH || d->init_kind == RGS_INIT_COV) && n > 63) {
    err = "lookup tables need nbits <= 63";
    return kInvalid;
  }
  return kOk;
}

void build_args(const rgs_slab_desc* d, const PassPlan& p, PassArgs& a) {
  const int n = d->nbits;
  a = PassArgs{};
  a.state = reinterpret_cast<float2*>(d->state);
  a.init_lut = d->init_lut;
  a.init_w = reinterpret_cast<const float2*>(d->init_wlut);
  a.phase_lut = d->phase_lut;
  a.dphase = reinterpret_cast<const float2*>(d->dphase_lut);
  // rotations of this pass's target bits travel in the kernel arguments;
  // qubits at or above gate_limit get t = 0 (the kernel rotates every bit)
  const int t0 = p.low ? 0 : p.qstart;
  for (int i = 0; i < 16; ++i) {
    const int q = t0 + i;
    a.tng[i] = (d->rot && i < p.count && q < n && q < d->gate_limit) ? d->rot[q] : 0.f;
  }
  // the rotations' common scale, applied once: in the values the state is
  // made from, or as the first pass loads a given state
  const double gsc = d->rot ? rgs_gate_scale(d->rot, std::min(d->gate_limit, n)) : 1.0;
  a.mass = d->mass;
  a.init_scale = (float)((double)d->init_scale * gsc);
  a.wscale = (float)gsc;
  a.lscale = (p.first && d->init_kind == RGS_INIT_LOAD) ? (float)gsc : 1.f;
  a.nbits = n;
  a.qstart = p.low ? 0 : p.qstart;
  a.flags = p.flags;
  // a contiguous pass that makes the state and only transforms it walks up
  // from the lowest layout (slab_kernel.h, F_INIT_UP)
  const bool init_up = p.low && p.flags == F_WHT;
  if (init_up) a.flags |= F_INIT_UP;
  a.init_kind = d->init_kind;
  const bool gen = !(p.flags & F_LOAD);
  a.init_chunks = (gen && (d->init_kind == RGS_INIT_GRAPH || d->init_kind == RGS_INIT_COV)) ? chunks_of(n) : 0;
  a.phase_chunks = ((p.flags & F_PHASE) && d->phase_lut) ? chunks_of(n) : 0;
  a.dphase_chunks = ((p.flags & F_PHASE) && !(p.flags & F_LAYER) && d->dphase_lut) ? chunks_of(n) : 0;
  if ((p.flags & F_LAYER) && (p.flags & F_PHASE) && d->dphase_lut) layer_phases(d, p, a);
  a.gate_limit = d->rot ? std::min(d->gate_limit, n) : 0;
  {
    const int s1 = std::max(p.tb - 4, p.c);
    const int sp = (p.flags & F_WHT) ? p.c : s1;  // layout where the phases are applied
    const int si = init_up ? p.c : s1;            // layout where the state is made
    int pi[4], pp[4];
    for (int i = 0; i < 4; ++i) {
      pi[i] = global_bit(p, si + i);
      pp[i] = global_bit(p, sp + i);
    }
    if (a.init_chunks) form_tables(d->init_rows, pi, a.iq, &a.ifreg);
    if (a.init_chunks && d->init_kind == RGS_INIT_GRAPH) local_form(p, d->init_rows, a.iform);
    if (a.init_chunks && d->init_kind == RGS_INIT_COV)
      for (int i = 0; i < 4; ++i) a.icol[i] = d->init_cols[pi[i]];
    if (a.phase_chunks) {
      form_tables(d->phase_rows, pp, a.pq, &a.pfreg);
      local_form(p, d->phase_rows, a.pform);
    }
  }
  a.mass_bits = d->mass_bits;
  a.mflip = d->mass_bits < 64 ? d->flip >> d->mass_bits : 0;
  a.dphase_global = d->dphase_global ? 1 : 0;
  a.wbits = d->init_wbits ? d->init_wbits : 8;
  a.wchunks = (n + a.wbits - 1) / a.wbits;
  a.wpop = d->init_wpop ? n + 1 : 0;
  const int opop = __builtin_popcountll(d->init_omask);
  a.wstride = n - opop + 1;
  a.wentries = a.wpop ? (opop + 1) * a.wstride : a.wchunks << a.wbits;
  a.womask = d->init_omask;
  a.wnmask = ~d->init_omask & ((n < 64 ? 1ull << n : 0ull) - 1ull);
}

hipError_t launch_pass_multi(const PassArgs* args, int shift, int count, const PassPlan& p, size_t shmem,
                             hipStream_t st) {
  switch (p.c) {
    case 0: return launch_c0_multi(p.tb, p.rb, args, shift, count, shmem, st);
    case 2: return launch_c2_multi(p.tb, p.rb, args, shift, count, shmem, st);
    case 3: return launch_c3_multi(p.tb, p.rb, args, shift, count, shmem, st);
    case 4: return launch_c4_multi(p.tb, p.rb, args, shift, count, shmem, st);
    case 5: return launch_c5_multi(p.tb, p.rb, args, shift, count, shmem, st);
    case 6: return launch_c6_multi(p.tb, p.rb, args, shift, count, shmem, st);
    case 7: return launch_c7_multi(p.tb, p.rb, args, shift, count, shmem, st);
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
  rc = check_desc(d, err);
  if (rc) return fail(rc, err);
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  for (int pi = first; pi <= last; ++pi) {
    const PassPlan& p = plan[pi];
    PassArgs a;
    build_args(d, p, a);
    hipError_t e = launch_pass(a, p, lds_bytes(p, d->nbits, d), st);
    if (e != hipSuccess) return fail(kHip, std::string("k_pass launch: ") + hipGetErrorString(e));
  }
  return kOk;
}

// The deferred last pass of d's schedule, its arguments with the masses off
// (the tile values only), and its LDS bytes.
static int leaf_pass(const rgs_slab_desc* d, PassPlan& p, PassArgs& a, size_t& shmem) {
  if (!d || !d->state) return fail(kInvalid, "null descriptor or state");
  if (!d->defer_leaves) return fail(kInvalid, "the descriptor does not defer its leaves");
  std::vector<PassPlan> plan;
  std::string err;
  int rc = plan_passes(d, plan, err);
  if (rc) return fail(rc, err);
  rc = check_desc(d, err);
  if (rc) return fail(rc, err);
  p = plan.back();
  build_args(d, p, a);
  a.flags &= ~(F_MASS | F_NOSTORE);
  shmem = lds_bytes(p, d->nbits, d);
  return kOk;
}

int leaf_tiles(const rgs_slab_desc* d, const int64_t* tiles, int count, void* out, hipStream_t st) {
  if (!tiles || !out) return fail(kInvalid, "null tile list or output");
  if (count < 0 || count > (1 << 24)) return fail(kInvalid, "leaf tile count must be in 0..2^24");
  PassPlan p;
  PassArgs a;
  size_t shmem;
  int rc = leaf_pass(d, p, a, shmem);
  if (rc) return rc;
  if (count == 0) return kOk;
  hipError_t e = launch_c0_leaf(p.tb, p.rb, a, tiles, count, reinterpret_cast<float2*>(out), shmem, st);
  if (e != hipSuccess) return fail(kHip, std::string("k_tile_leaf launch: ") + hipGetErrorString(e));
  return kOk;
}

// (the schedule only: the descriptor's tables are not read)
int leaf_bits(const rgs_slab_desc* d, int* bits) {
  if (!d) return fail(kInvalid, "null descriptor");
  if (!d->defer_leaves) return fail(kInvalid, "the descriptor does not defer its leaves");
  std::vector<PassPlan> plan;
  std::string err;
  int rc = plan_passes(d, plan, err);
  if (rc) return fail(rc, err);
  *bits = plan.back().tb;
  return kOk;
}

}  // namespace rgs

using namespace rgs;

extern "C" int rgs_slab_leaf_bits(const rgs_slab_desc* d, int* bits) {
  if (!bits) return fail(kInvalid, "null output");
  return leaf_bits(d, bits);
}

extern "C" int rgs_slab_leaf_tiles(const rgs_slab_desc* d, const int64_t* tiles, int count, void* out,
                                   void* stream) {
  return leaf_tiles(d, tiles, count, out, reinterpret_cast<hipStream_t>(stream));
}

extern "C" int rgs_slab_finish(const rgs_slab_desc* d, void* stream) {
  PassPlan p;
  PassArgs a;
  size_t shmem;
  int rc = leaf_pass(d, p, a, shmem);
  if (rc) return rc;
  // a mass-only last pass left the state final already
  if (!(a.flags & ~F_LOAD)) return kOk;
  hipError_t e = launch_pass(a, p, shmem, reinterpret_cast<hipStream_t>(stream));
  if (e != hipSuccess) return fail(kHip, std::string("k_pass launch: ") + hipGetErrorString(e));
  return kOk;
}

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
      std::snprintf(buf, sizeof buf, "%s%s%s%s%s%s%s:%d+%d/tb%d/c%d/rb%d ", p.low ? "low" : "strided",
                    (p.flags & F_LOAD) ? "" : "[init]", (p.flags & F_WHT) ? "[wht]" : "",
                    (p.flags & F_PHASE) ? "[phase]" : "", (p.flags & F_GATES) ? "[gates]" : "",
                    (p.flags & F_MASS) ? "[mass]" : "", (p.flags & F_NOSTORE) ? "[leaf]" : "", p.qstart,
                    p.count, p.tb, p.c, p.rb);
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
