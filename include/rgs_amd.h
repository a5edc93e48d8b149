/*
 * rgs_amd.h - C ABI of the MI355X (gfx950) graph-state Born sampler.
 *
 * Plain pointers, sizes and a HIP stream (passed as void*); no PyTorch or C++
 * types.  Every function returns RGS_OK or an error code; rgs_last_error()
 * gives the message of the last failure on the calling thread.  Device
 * pointers are marked "dev"; everything else is host memory.  All device work
 * is enqueued on `stream` (NULL = the default stream) and is asynchronous
 * unless stated otherwise.
 *
 * Complex numbers are interleaved (re, im) pairs: complex64 = 2 floats,
 * complex128 = 2 doubles.  Qubit k is bit k of a state index (qubit 0 = LSB),
 * as in the reference (work/rgs_attack/graph_sampler.py:1).
 *
 * Each declaration cites the reference interface it replaces.  The reference
 * launches CUDA kernels from Python strings through CuPy RawModule
 * (work/rgs_attack/distributed_sampler.py:80,88-90); this library is the
 * native boundary those Python engines would call instead.
 */
#ifndef RGS_AMD_H
#define RGS_AMD_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RGS_OK 0
#define RGS_EINVAL 1 /* ValueError in the Python mirror      */
#define RGS_EARITH 2 /* ArithmeticError                      */
#define RGS_ENOMEM 3 /* MemoryError                          */
#define RGS_EHIP 4   /* RuntimeError (HIP runtime failure)   */

#define RGS_ABI_VERSION 4

/* Thread-local message of the last failure ("" if none). */
const char *rgs_last_error(void);
int rgs_abi_version(void);

/* ------------------------------------------------------------------ *
 * Host planning: GF(2) algebra and chunked lookup tables (8-bit chunks).
 * ------------------------------------------------------------------ */

/* GF(2) inverse of an n x n 0/1 matrix (row-major uint8), n <= 64.
 * Replaces gf2_inverse (work/rgs_attack/fourier_covariant_sampler.py:23-36).
 * RGS_EINVAL when singular. */
int rgs_gf2_inverse(const uint8_t *m, int n, uint8_t *inv);

/* GF(2) rank of a rows x cols 0/1 matrix (cols <= 128).
 * Replaces gf2rank (work/rgs_attack/graph_sampler.py:59-69). */
int rgs_gf2_rank(const uint8_t *m, int rows, int cols, int *rank);

/* Phase words of the graph phase (-1)^{q(x)}, q(x) = sum_{a<b} G_ab x_a x_b,
 * for n <= 63 vertices: lut[c*256 + k] = (cross mask of chunk c code k onto
 * higher bits) | (in-chunk parity << 63); the parity of
 * XOR_c (lut[c*256+byte_c(x)] & (x | 1<<63)) is q(x).  lut has
 * ceil(n/8)*256 words.  Replaces the per-element popcount loop of
 * quadratic_phase / qmasks (work/rgs_attack/graph_sampler.py:76-85,
 * distributed_sampler.py:57) and the phase half of build_specialization
 * (work/rgs_attack/init_specialization.py:160-189). */
int rgs_graph_phase_lut(const uint8_t *g, int n, uint64_t *lut);

/* Fourier-covariant init tables for a balanced cut (|A| = |B| = m <= 63):
 * map_lut[(c*256 + k)*2 + 0] = XOR of the columns of N = (G[B,A])^{-T}
 * selected by code k of chunk c (z -> b = N z), map_lut[...*2 + 1] = phase
 * word of the transformed quadratic q_B(N z) (same convention as above);
 * aphase_lut = rgs_graph_phase_lut of G[A,A].  Replaces build_specialization
 * (work/rgs_attack/init_specialization.py:160-189) without its identical-U
 * restriction.  RGS_EINVAL if G[B,A] is singular. */
int rgs_cov_luts(const uint8_t *g, int nv, const int32_t *A, const int32_t *B, int m,
                 uint64_t *map_lut, uint64_t *aphase_lut);

/* Per-chunk products for arbitrary per-site unitaries: wlut[c*256 + k] =
 * prod_{t<8, c*8+t<m} u[B[c*8+t]][outB[c*8+t]][bit t of k] (complex64,
 * computed in complex128), `scale` folded into chunk 0.  u is complex128
 * [nv][2][2].  Replaces weight_table (init_specialization.py:98-115), which
 * needs identical B unitaries. */
int rgs_cov_weight_lut(const double *u, int nv, const int32_t *B, const uint8_t *outB, int m,
                       double scale, float *wlut);

/* The same products over chunks of chunk_bits (8 or 9) B bits:
 * wlut[c * 2^chunk_bits + k], ceil(m / chunk_bits) chunks.  Nine-bit chunks
 * make m = 35 four chunks instead of five: one table read and one complex
 * product fewer per amplitude in the covariant init (same reference
 * function as above). */
int rgs_cov_weight_lut_bits(const double *u, int nv, const int32_t *B, const uint8_t *outB, int m,
                            double scale, int chunk_bits, float *wlut);

/* Popcount weights for identical B unitaries u0 (complex128 [2][2]): with
 * o = outB and r = popcount(o), table[k * (m + 1) + h] =
 * scale * u00^(m-r-k+h) u01^(k-h) u10^(r-h) u11^h for max(0, r+k-m) <= h <=
 * min(r, k), else 0 (complex64 [(m+1)^2], built in complex128 and rounded
 * once), so that prod_j u0[o_j][b_j] = table[popcount(b)][popcount(b & o)].
 * The covariant init then takes one table read per amplitude.  Replaces
 * InitSpecialization.weight_table (init_specialization.py:98-115). */
int rgs_cov_weight_table(const double *u0, const uint8_t *outB, int m, double scale, float *table);

/* ------------------------------------------------------------------ *
 * Slab pipeline: init -> [Walsh] -> [diagonal phase] -> [2x2 gates] -> masses
 * on one 2^nbits complex64 state, fused into few HBM passes.
 * ------------------------------------------------------------------ */

#define RGS_INIT_LOAD 0     /* the state already holds the input          */
#define RGS_INIT_GRAPH 1    /* scale * (-1)^{q(x)} from graph_lut         */
#define RGS_INIT_COV 2      /* (-1)^{q_B(Nz)} prod_c wlut[c][byte_c(Nz)] */
#define RGS_INIT_DELTA 3    /* scale at index 0, zero elsewhere           */

typedef struct rgs_slab_desc {
  void *state;                 /* dev complex64[2^nbits], overwritten       */
  int nbits;                   /* 6..40                                     */
  int init_kind;               /* RGS_INIT_*                                */
  float init_scale;            /* GRAPH / DELTA                             */
  const uint64_t *init_lut;    /* dev GRAPH: phase words; COV: map_lut      */
  const uint64_t *init_rows;   /* HOST uint64[nbits], required for GRAPH and
                                  COV: row a has bit b set iff the sign form
                                  holds x_a x_b (symmetric), bit a iff it
                                  holds the linear term x_a (the adjacency
                                  rows for a graph phase; for COV the form
                                  of the sign words of init_lut)             */
  const uint64_t *init_cols;   /* HOST uint64[nbits], required for COV:
                                  column a of the linear map of init_lut
                                  (its map word for x = e_a)                 */
  const float *init_wlut;      /* dev COV weights (complex64), chunks of
                                  init_wbits bits                            */
  int do_walsh;                /* unnormalised Walsh transform on all bits  */
  const uint64_t *phase_lut;   /* dev sign words applied after the Walsh,
                                  or NULL                                    */
  const uint64_t *phase_rows;  /* HOST uint64[nbits] form rows of phase_lut
                                  (as init_rows); required with phase_lut   */
  const float *dphase_lut;     /* dev complex64 [ceil(nbits/8)][256] diagonal
                                  phase applied with phase_lut, or NULL      */
  const float *rot;            /* HOST float32 [nbits] tangents t of the
                                  qubits' rotations M(t) = [[1, -t], [t, 1]]
                                  (rgs_gate_shear), or NULL; read when a pass
                                  is enqueued (the values travel in the kernel
                                  arguments).  The state is multiplied once
                                  by rgs_gate_scale(rot, gate_limit): in the
                                  init (GRAPH, DELTA, and COV through its
                                  staged weights) or, for LOAD, as the first
                                  pass loads it                              */
  int gate_limit;              /* rotations on qubits >= gate_limit skipped */
  double *mass;                /* dev float64[2^(nbits-mass_bits)] or NULL  */
  int mass_bits;               /* granularity of the mass level (4..10)     */
  int low_bits;                /* tile bits of the contiguous pass (6..14)  */
  int tile_bits;               /* tile bits of strided passes (6..14)       */
  int lane_bits;               /* contiguous bits per strided row (2..7)    */
  int round_bits;              /* strided passes stage 2^-round_bits of the
                                  tile in LDS per round (0..2, applied where
                                  <= lane_bits and <= tile_bits - 10; fewer
                                  LDS bytes, more workgroups per CU)        */
  int phase_round_bits;        /* the same for the strided pass that applies
                                  the phase tables (0..2), or -1 to use
                                  round_bits                                 */
  int dphase_global;           /* 1: the phase pass reads the diagonal-phase
                                  table from global memory; 0: it stages the
                                  table in LDS per tile (a tuning choice)    */
  int init_wbits;              /* B bits per COV weight chunk: 8 or 9 (0 is
                                  read as 8)                                 */
  int group_order;             /* with the Walsh transform: 0 walks the
                                  strided groups from the low bits up (the
                                  phase in the highest group), 1 from the
                                  high bits down (the phase in the lowest)   */
  int top_bits;                /* 0: strided groups as even as the tile
                                  allows; else the group that applies the
                                  phase tables has exactly this many bits
                                  (4..tile_bits-lane_bits; with four it needs
                                  no LDS exchange) and the others are even   */
  int group_bits[8];           /* all 0: strided groups from tile_bits,
                                  lane_bits, round_bits and top_bits as
                                  above; else the widths of the strided
                                  groups from the low bits up, 0-terminated,
                                  summing to nbits - low_bits (each >= 4),
                                  with their own row widths and rounds:      */
  int group_lanes[8];          /*   lane bits of explicit group i (2..7),
                                  its tile bits group_lanes[i] +
                                  group_bits[i] <= 14                        */
  int group_rounds[8];         /*   LDS round bits of explicit group i (0..2,
                                  applied where <= its lane bits and <= its
                                  tile bits - 10)                            */
  uint64_t flip;               /* outcome x is held at state index x ^ flip
                                  (rgs_gate_shear's flips of the gates in
                                  rot, below gate_limit): the mass level is
                                  written in outcome order                   */
  uint64_t init_omask;         /* COV with init_wpop: the B outcome bits o   */
  int init_wpop;               /* COV: 0: init_wlut holds chunk products
                                  (init_wbits); 1: it holds the (nbits+1)^2
                                  popcount table of rgs_cov_weight_table,
                                  read at [popc(b)][popc(b & init_omask)]    */
  int defer_leaves;            /* 1: the last pass (contiguous, loading the
                                  state, with mass != NULL) writes the mass
                                  level but not the state, which keeps the
                                  values before that pass; draws and reads
                                  recompute the tiles they touch
                                  (rgs_slab_leaf_tiles, rgs_draw_leaves) and
                                  rgs_slab_finish writes the whole state.
                                  0: every pass writes the state             */
  int fuse_walsh;              /* with do_walsh and no phase_lut: 1 gives each
                                  pass the Walsh butterflies, the diagonal
                                  phases and the rotations of its own target
                                  bits, so the first strided group makes the
                                  state and one pass per group follows (four
                                  at n = 34 instead of seven); the per-qubit
                                  order Walsh -> phase -> rotation is kept,
                                  register layout by register layout.
                                  0: the Walsh transform on all bits first   */
  const float *dphase_host;    /* HOST copy of dphase_lut (same layout),
                                  required with fuse_walsh and dphase_lut: the
                                  fused passes take each register layout's
                                  phases from it as kernel arguments         */
} rgs_slab_desc;

/* Factor 2x2 unitaries u (complex128 [n][2][2]) as
 *   u = X^f D1 m M(t) D2,   M(t) = [[1, -t], [t, 1]], m = (1 + t^2)^(-1/2),
 * from u = D1 R D2 with R = [[c, -s], [s, c]] real, D1 diagonal and
 * D2 = diag(1, e^{i b}): t = s / c (f = 0) when c >= s, else t = -c / s with
 * the bit flip X (f = 1) and D1 replaced by diag(e^{i a1}, -e^{i a0}), so
 * |t| <= 1.  The slab pipeline applies D2 with the pointwise phase, M(t) per
 * qubit (two fused multiply-adds per amplitude and qubit) and the product of
 * the m's once (rgs_gate_scale); D1 changes no probability and the flips only
 * relabel the outcome (index x of the state is outcome x ^ flip), both applied
 * to amplitudes and indices read back (graph_sampler.apply_one semantics,
 * work/rgs_attack/graph_sampler.py:71-74).  Outputs (each optional):
 * tan float32 [n] (t, rounded once); dphase_lut complex64 [ceil(n/8)][256],
 * the product of the D2 phases of the qubits < gate_limit selected by each
 * byte; d1 complex128 [n][2]; flip: bit q set iff f_q = 1 (q < gate_limit;
 * the other qubits keep f = 0). */
int rgs_gate_shear(const double *u, int n, int gate_limit, float *tan, float *dphase_lut, double *d1,
                   uint64_t *flip);

/* prod_{q < gate_limit} (1 + tan[q]^2)^(-1/2) in FP64: the factor the slab
 * pipeline applies once for the shear rotations (no reference counterpart). */
double rgs_gate_scale(const float *tan, int gate_limit);

/* Enqueue the fused pipeline.  Replaces SpecializedDeferredSampler.compute
 * and its kernels initcov_specialized_global, lowfwht, fusedfwht,
 * lowprodphase, fusedprod (work/rgs_attack/specialized_deferred_sampler.py:58-74,
 * fourier_fused_sampler.py:100-112, distributed_sampler.py:14-49,
 * fourier_covariant_sampler.py:78-93).  When mass != NULL the last pass also
 * writes the pairwise-tree FP64 masses of every 2^mass_bits block. */
int rgs_slab_pipeline(const rgs_slab_desc *d, void *stream);

/* Enqueue only pass `pass` (0-based) of the same schedule, so a caller can
 * time each HBM pass; running every pass in order equals rgs_slab_pipeline.
 * No reference counterpart beyond the per-stage timers of compute
 * (work/rgs_attack/distributed_sampler.py:127-143). */
int rgs_slab_pass(const rgs_slab_desc *d, int pass, void *stream);

/* Deferred leaves (defer_leaves = 1).  The reference defers its terminal
 * gate to the draw (DeferredGlobalMixin, work/rgs_attack/deferred_global_sampler.py:41-78:
 * draw from the pre-gate masses, transform the two amplitudes drawn);
 * here the whole last contiguous pass is deferred: its kernel computes the
 * masses from the transformed values without storing them, and the values
 * of a tile are recomputed when asked for by the same kernel code, so they
 * are the bits the pass would have stored.
 *
 * rgs_slab_leaf_bits: the tile bits T of the deferred pass (tile t holds
 * state indices t * 2^T .. (t + 1) * 2^T - 1).
 * rgs_slab_leaf_tiles: the final values of `count` tiles (dev int64 tile
 * indices; an index >= 2^(nbits - T) leaves its slot unwritten) into out,
 * dev complex64 [count][2^T] in state order; count <= 2^24.
 * rgs_slab_finish: write the final values of every tile into the state
 * (after which the state equals the one rgs_slab_pipeline writes with
 * defer_leaves = 0; run it once per pipeline run).
 * Each needs the descriptor the pipeline ran with and the state it left. */
int rgs_slab_leaf_bits(const rgs_slab_desc *d, int *bits);
int rgs_slab_leaf_tiles(const rgs_slab_desc *d, const int64_t *tiles, int count, void *out, void *stream);
int rgs_slab_finish(const rgs_slab_desc *d, void *stream);

/* Number of HBM passes and their text description ("<kind>:<first bit>+<count>
 * ..."), for logs and the autotuner. */
int rgs_slab_describe(const rgs_slab_desc *d, int *passes, char *text, int text_len);

/* Many independent states in few launches.  Each item names a state's slab
 * descriptor (its `mass` level is required: the pyramid buffer must hold
 * 2^(nbits-mass_bits+1) - 1 doubles, as for rgs_mass_pyramid) and its draws
 * (dev uniforms, samples, dev outputs as rgs_draw).  Items whose descriptors
 * give the same pass schedule share launches: one per pass for the whole
 * group (items of 2^11 or more tiles per pass, which fill the device alone,
 * keep one launch per item and pass), then one per chunk of pyramid levels
 * and one for every draw of every item.  The results equal rgs_slab_pipeline + rgs_mass_pyramid + rgs_draw
 * per item bit for bit (the same kernels).  create plans the batch and
 * uploads its kernel arguments (one device allocation, synchronous); run only
 * enqueues kernel launches on `stream` (it can be captured in a graph).
 * The descriptors are copied at create; the device buffers they and the items
 * name must stay valid while the batch is used.  create returns RGS_EINVAL
 * when one launch would need 2^31 work-items or more (a pass of a group, a
 * pyramid round, or 2^25 draws): split such a batch.  No
 * counterpart in the reference, which runs one state per process
 * (work/rgs_attack/born_production.py:44-58). */
typedef struct rgs_slab_batch_item {
  const rgs_slab_desc *desc;   /* HOST descriptor */
  const double *uniforms;      /* dev float64[samples] */
  int samples;
  int64_t *out_index;          /* dev int64[samples] */
  double *out_p;               /* dev float64[samples] */
  uint64_t flip;               /* the descriptor's flip (rgs_draw)          */
} rgs_slab_batch_item;
int rgs_slab_batch_create(const rgs_slab_batch_item *items, int count, void **batch);
int rgs_slab_batch_run(void *batch, void *stream);
int rgs_slab_batch_info(void *batch, int *launches, int *groups);
int rgs_slab_batch_destroy(void *batch);

/* Build the upper levels of the FP64 mass pyramid in place.  pyr holds level 0
 * (2^base_bits masses) followed by levels of 2^(base_bits-1), ..., 1 entries:
 * level[k][i] = level[k-1][2i] + level[k-1][2i+1].  Replaces the host
 * np.concatenate + sum of blockmass output
 * (work/rgs_attack/distributed_sampler.py:145-152). */
int rgs_mass_pyramid(double *pyr, int base_bits, void *stream);

/* Qubit-by-qubit draws: for each uniform u, descend the pyramid from the root
 * with t = u * root (go to child 1 iff t >= mass(child 0) and mass(child 1) > 0,
 * subtracting mass(child 0)), then inside the 2^mass_bits leaf block with masses
 * recomputed by the same pairwise tree from the complex64 amplitudes.
 * The state holds outcome x at index x ^ flip (the rgs_gate_shear flips of
 * its gates; 0 for none) and the pyramid is in outcome order (as the slab
 * pipeline writes it), so draws descend in outcome order.  Outputs the
 * outcome and p = re^2 + im^2 (float64) of each draw.  Replaces
 * DistributedSampler.draw's host searchsorted / tile copy
 * (work/rgs_attack/distributed_sampler.py:153-157). */
int rgs_draw(const void *state, int nbits, const double *pyr, int mass_bits,
             const double *uniforms, int count, uint64_t flip, int64_t *out_index, double *out_p,
             void *stream);

/* rgs_draw on a state left by a pipeline with defer_leaves = 1: the pyramid
 * descent (one thread per draw), the recomputation of the tile holding each
 * draw's leaf block (rgs_slab_leaf_tiles), then the leaf-block draw; the
 * outcomes and p are those rgs_draw gives on the finished state, bit for
 * bit.  nbits, mass_bits and flip come from the descriptor.  work: dev
 * memory of at least rgs_draw_leaves_workspace(d, count) bytes (256-byte
 * aligned). */
int rgs_draw_leaves_workspace(const rgs_slab_desc *d, int count, size_t *bytes);
int rgs_draw_leaves(const rgs_slab_desc *d, const double *pyr, const double *uniforms, int count, void *work,
                    size_t work_bytes, int64_t *out_index, double *out_p, void *stream);

/* float64 p = re^2 + im^2 at `count` indices (probability_at,
 * distributed_sampler.py:158-161). */
int rgs_gather_prob(const void *state, const int64_t *index, int count, double *out_p,
                    void *stream);


/* ------------------------------------------------------------------ *
 * Linear rank-width scan: p(x) of bitstrings by a fixed vertex order.
 * ------------------------------------------------------------------ */

#define RGS_SCAN_MAX_R 62

/* One compiled step: sum the vertex v.  p-vectors pack (c, y) as
 * c | y << r_in.  out[c'] = sum_{p in L c' + span(kern)} E[p & (2^r_in - 1)]
 * * f_v(y) * (-1)^{y * parity(c & beta)}, with L the columns lcols. */
typedef struct rgs_scan_step {
  int32_t v;
  int32_t r_in;
  int32_t r_out;
  int32_t nker;
  uint64_t beta;
  uint64_t lcols[64];
  uint64_t kern[64];
} rgs_scan_step;

/* Compile the scan for an ordering of all n <= 128 vertices; profile[j] =
 * rank after step j.  Replaces make_plan and future_masks
 * (work/rgs_attack/upstream_scan/scan_plan.py:45-105). */
int rgs_scan_compile(const uint8_t *g, int n, const int32_t *order, rgs_scan_step *steps,
                     int32_t *profile);

/* Cut-rank profile by direct GF(2) elimination of every prefix cut.
 * Replaces search_width.profile (work/rgs_attack/search_width.py:5-15) and
 * rank_cut (work/rgs_attack/order_optimization/search.cpp:16-23). */
int rgs_scan_profile(const uint8_t *g, int n, const int32_t *order, int32_t *profile);

/* Annealed order search minimising sum_j 2^{r_j} with max r <= rmax_cap.
 * `threads` independent chains (seeds derived from `seed`) each run
 * `iterations` moves (neighbour swap, short-window insertion or reversal,
 * long swap; epochs that restart from the chain's best order and cool
 * geometrically); the best order over the chains is returned (ties: the
 * lowest chain).  Deterministic for given arguments: no wall-clock budget.
 * Replaces the annealer of work/rgs_attack/order_optimization/search.cpp:
 * 24-31 (incremental cut ranks and cost) and 34-67 (epochs, moves, the
 * rank-30 cap, Metropolis acceptance; it runs for a wall-clock budget), and
 * the swap annealer of work/rgs_attack/search_width.py:17-27. */
int rgs_scan_order_search(const uint8_t *g, int n, const int32_t *start, uint64_t iterations, uint64_t seed,
                          int rmax_cap, int threads, int32_t *best_order, uint64_t *best_cost);

/* Pack compiled steps into the device plan (uint64 words): per step a header
 * (v | r_in << 8 | r_out << 16 | nker << 24 | split << 32), beta, p0 over the
 * low six output bits (64 words) and over the rest (2^(r_out-6) words), and
 * the 2^nker kernel combinations.  The kernel basis is reduced on the y bit
 * (bit r_in); when one basis vector keeps it (split = 1) that vector goes
 * last and every p0 is stored as the y = 0 member of its coset, so the
 * combinations' first half has y = 0 and the second y = 1.  Call with
 * blob == NULL to get the size in *words.  Ranks must be <= 28.  Replaces the
 * per-step byte_tables / kernel_combos uploads of GPUMixin.alloc
 * (upstream_scan/scan_plan.py:133-152, gpu_kernel.py:111-123). */
int rgs_scan_pack(const rgs_scan_step *steps, int n, uint64_t *blob, uint64_t *words);

/* p(x) of `count` proposals in one persistent launch.  plan_dev: packed plan;
 * rmax: its largest rank, max(r_in, r_out) over the steps (the LDS tables and
 * the kernel, one wavefront per four proposals at rmax <= 6, are chosen by
 * it; a larger value is allowed; when the plan's own rank, word 1 of the
 * blob, exceeds what the chosen kernel holds, the kernel reads no table and
 * marks every output: value NaN, exponent INT32_MIN);
 * u_dev: complex128 [n][2][2] measurement unitaries (f_v = u[v][x_v][:]);
 * bits_dev: uint64 [count][2] bitstrings (qubit q = bit q).  precision 64
 * stores the tables as complex64 (rmax <= 13), 128 as complex128
 * (rmax <= 12); sums are FP64 in both.  The plan must come from
 * rgs_scan_pack (its split flag and y = 0 coset representatives are what the
 * kernel's one-weight-per-output sums rely on).  Outputs: out_val complex128
 * [count] and out_exp int32 [count] with amplitude
 * = out_val * 2^out_exp * 2^{-n/2}.  Replaces ScanOracle.amplitude /
 * GPUMixin._sweep / scan_step_scaled (work/rgs_attack/scan_oracle.py:25-29,
 * upstream_scan/gpu_kernel.py:59-102,125-170, gpu_kernel_fp64.py:24-67,90-135),
 * which run one proposal per 70 launches with a host read per step. */
int rgs_scan_batch(const uint64_t *plan_dev, int rmax, const double *u_dev, const uint64_t *bits_dev,
                   int count, int precision, double *out_val, int32_t *out_exp, void *stream);

/* Capped-rejection decisions of `count` scored proposals, on the device:
 * p = (re^2 + im^2) 2^(2 out_exp - n) from rgs_scan_batch's outputs (val
 * complex128 [count], ex int32 [count]), then keep[i] = u[i] < min(p * scale,
 * 1) for uniform proposals (logq == NULL, scale = 2^n / cap) or
 * u[i] < min(p / (scale * exp(logq[i])), 1) for proposals of density q
 * (scale = cap).  Each product, sum and quotient is one FP64 rounding, in
 * that order.  keep: uint8 [count] (0 or 1); the number kept is added to
 * *accepted (device uint64, not cleared here).  Replaces the accept test of
 * rejection_sampler.py:18-24, one proposal and one host comparison per call. */
int rgs_capped_decide(const double *val, const int32_t *ex, const double *u, const double *logq, int count, int n,
                      double scale, uint8_t *keep, unsigned long long *accepted, void *stream);

/* Byte-table form of a plan for rgs_scan_large (ranks <= 32): per step a
 * header, beta, ceil(r_out/8) x 256 words of XORs of the lcols and the kernel
 * combinations; blob[2 + j] is the word offset of step j.  Call with
 * blob == NULL to get the size in *words.  Replaces byte_tables /
 * kernel_combos (upstream_scan/scan_plan.py:133-152). */
int rgs_scan_pack_tables(const rgs_scan_step *steps, int n, uint64_t *blob, uint64_t *words);

/* p(x) of ONE proposal with the environment tables in HBM, for ranks up to
 * 32: one launch per step over 2^r_out outputs, no host round trip between
 * steps (the per-step power-of-two scale is a device maximum applied by the
 * next step).  steps: the compiled plan (host); plan_dev: the packed tables
 * (rgs_scan_pack_tables) on the device; plan_offsets: host copy of blob[2..]
 * (the step offsets); u_host: complex128 [n][2][2]; bits_host: uint64[2];
 * buf0 / buf1: device tables of buf_entries >= 2^rmax entries each
 * (complex64 for precision 64, complex128 for 128); mx_dev: device uint64[n]
 * scratch.  Outputs on the device: out_val complex128[1], out_exp int32[1]
 * with amplitude = out_val * 2^out_exp * 2^{-n/2}.  Replaces GPUMixin._sweep /
 * scan_step_scaled at large rank (upstream_scan/gpu_kernel.py:59-102,125-170,
 * gpu_kernel_fp64.py:24-67,90-135), which read 4 bytes back to the host after
 * every step. */
int rgs_scan_large(const rgs_scan_step *steps, int n, const uint64_t *plan_dev, const uint64_t *plan_offsets,
                   const double *u_host, const uint64_t *bits_host, int precision, void *buf0, void *buf1,
                   uint64_t buf_entries, uint64_t *mx_dev, double *out_val, int32_t *out_exp, void *stream);

/* rgs_scan_large without caller buffers, for callers that cannot allocate
 * device memory (code under HIP graph capture, serving loops, harnesses that
 * forbid hipMalloc).  No counterpart in the reference.  The two tables and
 * the step maxima live in a fixed array of this library, one copy per device
 * (32 MiB + 4 KiB, held until the process ends), so plans are limited to
 * rank 21 with complex64 tables and rank 20 with complex128 tables (larger
 * ranks return RGS_EINVAL before any launch).  The call allocates nothing.
 * Every call on a device shares that device's array: calls that overlap on
 * one device (other streams, other threads) race.  Outputs as rgs_scan_large. */
int rgs_scan_large_static(const rgs_scan_step *steps, int n, const uint64_t *plan_dev, const uint64_t *plan_offsets,
                          const double *u_host, const uint64_t *bits_host, int precision, double *out_val,
                          int32_t *out_exp, void *stream);

/* ---- learned proposal (MADE) -------------------------------------------
 * The reference has no counterpart for any entry point below: the learned
 * proposal is new work.  All matrices are row-major; bf16 values are raw
 * uint16 bit patterns (round-to-nearest-even from FP32). */

/* C[M][N] (+)= alpha * op(A) op(B) with op = transpose when ta / tb is 1
 * (K may be 0: the product is empty and only the epilogue applies);
 * A is [M][K] (or [K][M]), B is [K][N] (or [N][K]), bf16, FP32 MFMA
 * accumulation.  Epilogue in order: + bias[N] (may be NULL), ReLU if relu,
 * multiply by (dmask[i][j] != 0) if dmask is non-NULL (ReLU backward), then
 * store FP32 to C (may be NULL; accumulate adds to C) and/or bf16 to Cb
 * (may be NULL). */
int rgs_gemm_bf16(int ta, int tb, int M, int N, int K, const uint16_t *A, int lda, const uint16_t *B, int ldb,
                  float alpha, float *C, int ldc, uint16_t *Cb, int ldcb, const float *bias, int relu,
                  const uint16_t *dmask, int ldm, int accumulate, void *stream);

/* Bernoulli negative log-likelihood of bits x[B][n] (bf16 0/1) under
 * logits[B][n]: nll[b] = sum_i softplus(l) - x l.  If dlogits is non-NULL it
 * receives scale * (sigmoid(l) - x) in FP32, and dlogits_bf16 (may be NULL)
 * the same in bf16. */
int rgs_made_bce(const float *logits, const uint16_t *x, int B, int n, float scale, float *nll, float *dlogits,
                 uint16_t *dlogits_bf16, void *stream);

/* out[j] (+)= sum_r in[r][j] over rows (bias gradients), FP32 / bf16 input. */
int rgs_colsum(const float *in, int rows, int cols, int ld, float *out, int accumulate, void *stream);
int rgs_colsum_bf16(const uint16_t *in, int rows, int cols, int ld, float *out, void *stream);

/* One masked Adam step on FP32 master weights w[count] (bias-corrected with
 * the device step counter *t_dev, which this call increments first);
 * entries with mask 0 stay untouched; wb receives the bf16 copy. */
int rgs_adam(float *w, const float *grad, const float *mask, float *m, float *v, uint16_t *wb, long long count,
             int *t_dev, float lr, float b1, float b2, float eps, void *stream);

/* Ancestral sampling of column o: x[b][o] = (u[b][o] < sigmoid(logits[b][o]))
 * with u an FP64 [B][n] array of uniforms; x holds bf16 1.0 / 0.0. */
int rgs_sample_col(const float *logits, const double *u, int B, int n, int o, uint16_t *x, void *stream);

/* One step of incremental autoregressive sampling: x[b][o] = (u[b][o] <
 * sigmoid(logit[b * ldl])) for b < B (u FP64 [B][n], x bf16 [B][n]), and
 * logq[b] = (first ? 0 : logq[b]) + log sigmoid(+-logit) of the chosen bit
 * (FP64). */
int rgs_made_sample_step(const float *logit, int ldl, const double *u, int B, int n, int o, uint16_t *x,
                         double *logq, int first, void *stream);

/* All n steps of incremental sampling in one launch (n <= 64, H a multiple
 * of 16), with the same arithmetic as the sequence of rgs_gemm_bf16 /
 * rgs_made_sample_step calls (bit-identical results), for widths up to 2048
 * whose degrees own whole aligned 16-unit blocks (H = 16 m (n-1), or
 * H = 16 k with k < n-1 and degrees 1..k, as hidden_degrees assigns them):
 * one wavefront per 16 proposals with transposed MFMA products and no
 * workgroup barrier.  Other widths return an invalid-argument error (the step
 * sequence serves them).  kpref is checked for presence only: the degree
 * blocks follow from n and H.  W1 bf16 [H][n], W2 [H][H], W3 [n][H] (rows masked),
 * biases FP32, kpref dev int32[n] = number of hidden units of degree <= o
 * (degrees sorted), u FP64 [B][n].  Outputs (words and/or xout required):
 * words uint64 [B][2] (qubit q = bit q % 64 of word q / 64), xout bf16 0/1
 * [B][n], logq FP64 [B]. */
int rgs_made_sample(const uint16_t *W1, const float *b1, const uint16_t *W2, const float *b2, const uint16_t *W3,
                    const float *b3, const int *kpref, const double *u, int B, int n, int H, uint64_t *words,
                    uint16_t *xout, double *logq, void *stream);

/* Fragment-major copy of the sampler's weights for rgs_made_sample_packed
 * (degree-blocked widths only): first layer [H/16][ceil(n/16)] tiles with
 * inputs at or past each block's degree zeroed, second layer [H/16][H/16]
 * tiles, logit rows [ceil(n/16)][H/16] tiles (rows past n zeroed); each
 * 16 x 16 tile is 64 lanes x 4 bf16 in the MFMA A-operand order, so the
 * sampler's fragment loads are contiguous whatever the row stride.  Call
 * with packed == NULL to get the size in *bytes.  No reference counterpart
 * (the learned proposal is new). */
int rgs_made_sample_pack(const uint16_t *W1, const uint16_t *W2, const uint16_t *W3, int n, int H, void *packed,
                         uint64_t *bytes, void *stream);

/* rgs_made_sample reading the packed weights of rgs_made_sample_pack:
 * bit-identical outputs. */
int rgs_made_sample_packed(const void *packed, const float *b1, const float *b2, const float *b3, const double *u,
                           int B, int n, int H, uint64_t *words, uint16_t *xout, double *logq, void *stream);

/* bf16 copy (round to nearest even) of an FP32 array. */
int rgs_f32_to_bf16(const float *in, uint16_t *out, long long count, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* RGS_AMD_H */
