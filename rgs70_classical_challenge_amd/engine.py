"""Sampler engines: the drop-in Python boundary over the HIP slab pipeline.

Two engines share the kernels:

* ``GraphStateSampler`` - the plain method on a full n-qubit state
  (x)_k u[k] |G> (graph_sampler.graph_vector + apply_one,
  work/rgs_attack/graph_sampler.py:87-88,150-151), computed by the slab
  pipeline init -> Walsh -> gates (``pipeline='slab'``: g(y) = 2^{-n}(-1)^{q(y)},
  unnormalised Walsh, gates u_k H), the literal circuit
  (``'literal'``: delta -> Walsh -> graph phase -> u), or with H^n|0> folded
  into the init as the reference's graph_vector does (``'folded'``).
  Draws descend the FP64 mass pyramid qubit by qubit on the device.

* ``ConditionalSampler`` - the reference's production algorithm for a
  balanced full-rank cut A|B (SpecializedDeferredSampler,
  work/rgs_attack/specialized_deferred_sampler.py:40-74, with
  DeferredGlobalMixin, deferred_global_sampler.py:40-88): the Fourier-covariant
  init from chunk tables (any per-site unitaries), Walsh, the A-side graph
  phase and the A gates, with the terminal gate on bit m-1 deferred to the
  draw.  Method names, return tuples, exceptions and the 3-uniform draw
  schedule follow the reference (compute / masses / draw / probability_at /
  host / sync).

Device memory and streams come from PyTorch (plumbing only); every kernel is
in librgs_amd.so.  There is no CPU fallback.
"""
import time

import numpy as np

from . import _lib
from ._lib import call, ptr

HADAMARD = np.array([[1, 1], [1, -1]], np.complex128) / np.sqrt(2.0)


def _torch():
    import torch
    return torch


def default_tuning(nbits):
    """Tile choices used when neither the caller nor the autotuner table gives
    one (the fastest measured full pipeline at n = 34: a 12-bit contiguous
    pass, strided passes of 14-bit tiles with 64-amplitude rows and their LDS
    exchanges in four rounds)."""
    return dict(low_bits=min(12, nbits), tile_bits=14, lane_bits=6, round_bits=2, phase_round_bits=-1,
                dphase_global=0, top_bits=0, group_order=0, fuse_walsh=0)


def resolve_tuning(nbits, props, tuning=None, kind=''):
    """Caller's tuning, else the autotuner's entry for this device (by its
    properties), pipeline kind ('' plain, 'c' conditional) and size - a
    conditional engine falls back to the plain entry - else default_tuning;
    missing keys are filled from default_tuning."""
    if tuning is None:
        from .autotune import lookup_props
        entry = lookup_props(f'{kind}{nbits}', props) if kind else None
        if entry is None:
            entry = lookup_props(nbits, props)
        if entry is not None:
            tuning = {k: entry[k] for k in ('low_bits', 'tile_bits', 'lane_bits', 'round_bits', 'phase_round_bits',
                                            'dphase_global', 'top_bits', 'group_order', 'groups', 'fuse_walsh')
                      if k in entry}
    return dict(default_tuning(nbits), **(tuning or {}))


def effective_low_bits(nbits, low_bits):
    """Bits of the contiguous pass as the pass planner sets them: at most
    nbits, and narrowed so that a strided group keeps a full register window
    (4 bits)."""
    low = min(low_bits, nbits)
    if nbits > low and nbits - low < 4:
        low = nbits - 4
    return low


class _Slab:
    """A 2^nbits complex64 state, its lookup tables and its mass pyramid on one device."""

    def __init__(self, nbits, device, mass_bits=8, tuning=None, kind='', arena=None):
        torch = _torch()
        if not torch.cuda.is_available():
            raise RuntimeError('no HIP device visible')
        _lib.lib()
        self.torch = torch
        self.nbits = nbits
        self.device = torch.device('cuda', device)
        if arena is None:
            free, _total = torch.cuda.mem_get_info(self.device)
            need = (8 << nbits) + (16 << max(0, nbits - mass_bits)) + (64 << 20)
            if free < need:
                raise MemoryError(f'device {device}: need {need} bytes, free {free}')
        self.tuning = resolve_tuning(nbits, torch.cuda.get_device_properties(self.device), tuning, kind)
        self.mass_bits = max(4, min(mass_bits, 8, nbits, effective_low_bits(nbits, self.tuning['low_bits'])))
        self.base_bits = nbits - self.mass_bits
        # state and pyramid: own tensors, or views into a caller's arena (take(count, dtype))
        take = arena.take if arena is not None else (
            lambda count, dtype: torch.empty(count, dtype=dtype, device=self.device))
        self.state = take(1 << nbits, torch.complex64)
        self.pyramid = take((2 << self.base_bits) - 1, torch.float64)
        self.stream = torch.cuda.current_stream(self.device)

    def to_device(self, array, dtype=None):
        t = self.torch.as_tensor(np.ascontiguousarray(array))
        if dtype is not None:
            t = t.to(dtype)
        return t.to(self.device, non_blocking=False)

    def desc(self, init_kind, init_scale=0.0, init_lut=None, init_wlut=None, walsh=False, phase_lut=None,
             gates=None, gate_limit=None, mass=True, init_rows=None, phase_rows=None, init_cols=None,
             init_wbits=8, defer_leaves=False):
        """gates: a Gates object (rotations + diagonal phase) or None;
        init_rows / phase_rows: host form rows (form_rows, cov_form) of the
        sign tables init_lut (GRAPH and COV init) and phase_lut; init_cols:
        host map columns of a COV init_lut (cov_form)."""
        d = _lib.SlabDesc()
        d.state = ptr(self.state)
        d.nbits = self.nbits
        d.init_kind = init_kind
        d.init_scale = init_scale
        d.init_lut = ptr(init_lut)
        d.init_rows = ptr(init_rows)
        d.init_cols = ptr(init_cols)
        d.init_wbits = init_wbits
        d.init_wlut = ptr(init_wlut)
        d.do_walsh = int(walsh)
        d.phase_lut = ptr(phase_lut)
        d.phase_rows = ptr(phase_rows)
        d.dphase_lut = ptr(gates.dphase) if gates is not None else None
        d.dphase_host = gates.dphase_host.ctypes.data if gates is not None else None
        d.rot = ptr(gates.rot) if gates is not None else None
        d.gate_limit = self.nbits if gate_limit is None else gate_limit
        d.mass = ptr(self.pyramid) if mass else None
        d.mass_bits = self.mass_bits
        d.low_bits = self.tuning['low_bits']
        d.tile_bits = self.tuning['tile_bits']
        d.lane_bits = self.tuning['lane_bits']
        d.round_bits = self.tuning.get('round_bits', 0)
        d.phase_round_bits = self.tuning.get('phase_round_bits', -1)
        d.dphase_global = int(self.tuning.get('dphase_global', 0))
        d.top_bits = int(self.tuning.get('top_bits', 0))
        d.group_order = int(self.tuning.get('group_order', 0))
        d.flip = gates.flip if gates is not None else 0
        d.defer_leaves = int(defer_leaves)
        # explicit strided groups, low bits up: (width, lane bits, LDS round bits)
        groups = self.tuning.get('groups') or ()
        if len(groups) > 8:
            raise ValueError('at most 8 explicit strided groups')
        for i, (bits, lane, rounds) in enumerate(groups):
            d.group_bits[i] = int(bits)
            d.group_lanes[i] = int(lane)
            d.group_rounds[i] = int(rounds)
        # the descriptor holds raw pointers: keep their owners alive with it
        d._owners = (self.state, self.pyramid, init_lut, init_wlut, phase_lut, gates, init_rows, phase_rows,
                     init_cols)
        return d

    def run(self, d, mass=True):
        s = self.stream.cuda_stream
        call('rgs_slab_pipeline', ctypes_ref(d), s)
        if mass:
            call('rgs_mass_pyramid', ptr(self.pyramid), self.base_bits, s)

    def describe(self, d):
        import ctypes
        n = ctypes.c_int()
        buf = ctypes.create_string_buffer(2048)
        call('rgs_slab_describe', ctypes_ref(d), ctypes.byref(n), buf, 2048)
        return n.value, buf.value.decode()

    def level(self, bits):
        """Masses of aligned 2^bits blocks (bits >= mass_bits) as a host float64 array."""
        lvl = bits - self.mass_bits
        if lvl < 0:
            raise ValueError('mass level finer than the pyramid base')
        B = self.base_bits
        off = (2 << B) - (2 << (B - lvl))
        return self.pyramid[off:off + (1 << (B - lvl))].cpu().numpy()

    def root(self):
        return float(self.pyramid[-1].item())

    def draw_many(self, uniforms, flip=0):
        """Device draws (the state holds outcome x at index x ^ flip, the
        pyramid is in outcome order): (outcomes, p)."""
        torch = self.torch
        u = torch.as_tensor(np.ascontiguousarray(uniforms, np.float64)).to(self.device)
        idx = torch.empty(len(u), dtype=torch.int64, device=self.device)
        p = torch.empty(len(u), dtype=torch.float64, device=self.device)
        call('rgs_draw', ptr(self.state), self.nbits, ptr(self.pyramid), self.mass_bits, ptr(u), len(u), int(flip),
             ptr(idx), ptr(p), self.stream.cuda_stream)
        return idx.cpu().numpy(), p.cpu().numpy()

    # ---- deferred leaves: the last pass wrote masses only (defer_leaves)
    def leaf_bits(self, d):
        import ctypes
        b = ctypes.c_int()
        call('rgs_slab_leaf_bits', ctypes_ref(d), ctypes.byref(b))
        return b.value

    def leaf_tiles(self, d, tiles):
        """Final values of the listed tiles (state order), a device
        complex64 tensor [len(tiles), 2^leaf_bits]."""
        torch = self.torch
        lb = self.leaf_bits(d)
        t = torch.as_tensor(np.ascontiguousarray(tiles, np.int64)).to(self.device)
        out = torch.empty((len(t), 1 << lb), dtype=torch.complex64, device=self.device)
        call('rgs_slab_leaf_tiles', ctypes_ref(d), ptr(t), len(t), ptr(out), self.stream.cuda_stream)
        return out

    def leaf_values(self, d, indices):
        """Final complex64 values at state indices (host array)."""
        ix = np.asarray(indices, np.int64)
        lb = self.leaf_bits(d)
        tiles, inv = np.unique(ix >> lb, return_inverse=True)
        vals = self.leaf_tiles(d, tiles).cpu().numpy()
        return vals[inv.reshape(ix.shape), ix & ((1 << lb) - 1)]

    def leaf_block(self, d, start, stop):
        """Final complex64 values of state indices start..stop-1 (host array)."""
        lb = self.leaf_bits(d)
        t0, t1 = start >> lb, ((stop - 1) >> lb) + 1
        vals = self.leaf_tiles(d, np.arange(t0, t1, dtype=np.int64)).reshape(-1).cpu().numpy()
        return vals[start - (t0 << lb):stop - (t0 << lb)]

    def draw_leaves(self, d, uniforms):
        """rgs_draw on a deferred state (rgs_draw_leaves): (outcomes, p)."""
        import ctypes
        torch = self.torch
        u = torch.as_tensor(np.ascontiguousarray(uniforms, np.float64)).to(self.device)
        nb = ctypes.c_size_t()
        call('rgs_draw_leaves_workspace', ctypes_ref(d), len(u), ctypes.byref(nb))
        work = torch.empty(max(nb.value, 1), dtype=torch.uint8, device=self.device)
        idx = torch.empty(len(u), dtype=torch.int64, device=self.device)
        p = torch.empty(len(u), dtype=torch.float64, device=self.device)
        call('rgs_draw_leaves', ctypes_ref(d), ptr(self.pyramid), ptr(u), len(u), ptr(work), nb.value, ptr(idx),
             ptr(p), self.stream.cuda_stream)
        return idx.cpu().numpy(), p.cpu().numpy()

    def gather_prob(self, indices):
        torch = self.torch
        ix = torch.as_tensor(np.ascontiguousarray(indices, np.int64)).to(self.device)
        out = torch.empty(len(ix), dtype=torch.float64, device=self.device)
        call('rgs_gather_prob', ptr(self.state), ptr(ix), len(ix), ptr(out), self.stream.cuda_stream)
        return out.cpu().numpy()

    def sync(self):
        self.torch.cuda.synchronize(self.device)


def _deferred(slab, d, defer):
    """d with its last pass deferred (defer_leaves) when asked for and the
    schedule has one to defer (a single pass makes the state itself)."""
    if defer:
        passes, _ = slab.describe(d)
        d.defer_leaves = int(passes >= 2)
    return d


def ctypes_ref(d):
    import ctypes
    return ctypes.byref(d)


class Gates:
    """Per-qubit unitaries u_q = X^f D1 m M(t) D2 (rgs_gate_shear): host
    rotation tangents t (read when passes are enqueued; the pipeline applies
    the product of the m's once), the device chunk table of the D2 layer
    (qubits < gate_limit), the host D1 phases that read-back applies, and the
    flip mask: outcome x is held at state index x ^ flip."""

    def __init__(self, slab, mats, gate_limit=None):
        import ctypes
        u = np.ascontiguousarray(np.asarray(mats, np.complex128))
        n = len(u)
        self.gate_limit = n if gate_limit is None else gate_limit
        tan = np.zeros(n, np.float32)
        dph = np.zeros(((n + 7) // 8) * 256 * 2, np.float32)
        d1 = np.zeros(4 * n, np.float64)
        flip = ctypes.c_uint64()
        call('rgs_gate_shear', ptr(u), n, self.gate_limit, ptr(tan), ptr(dph), ptr(d1), ctypes.byref(flip))
        self.rot = tan                      # host: read when passes are enqueued
        self.flip = int(flip.value)
        self.dphase = slab.to_device(dph)
        self.dphase_host = dph              # host copy: per-layout phases of fused passes
        self.d1 = d1.view(np.complex128).reshape(n, 2)

    def apply_d1(self, v):
        """In place: multiply amplitude x (state order) by prod_{q < gate_limit} D1_q[x_q]."""
        ids = np.arange(len(v), dtype=np.uint64)
        for q in range(self.gate_limit):
            bit = ((ids >> np.uint64(q)) & np.uint64(1)).astype(np.int64)
            v *= self.d1[q][bit]
        return v

    def to_outcomes(self, v):
        """A state-ordered vector in outcome order: out[x] = v[x ^ flip]."""
        if not self.flip:
            return v
        return v[np.arange(len(v), dtype=np.int64) ^ self.flip]


def form_rows(g):
    """Host rows of the graph phase form: bit b of row a = g[a][b] (uint64)."""
    g = np.asarray(g, np.uint8)
    w = (np.uint64(1) << np.arange(len(g), dtype=np.uint64))
    return np.ascontiguousarray((g.astype(np.uint64) * w[None, :]).sum(axis=1, dtype=np.uint64))


def cov_form(map_lut, m):
    """Rows and columns of a covariant init table (rgs_cov_luts layout
    [chunk][256][map word, sign word]), read off the table itself: column a
    of the linear map is the map word of z = e_a; the sign form
    Q(z) = parity((XOR of the sign words) & (z | 1 << 63)) is quadratic, so
    row a holds bit b (b != a) iff Q(e_a ^ e_b) ^ Q(e_a) ^ Q(e_b) ^ Q(0) = 1
    and bit a iff Q(e_a) ^ Q(0) = 1 (the form_rows convention)."""
    lut = np.asarray(map_lut, np.uint64).reshape(-1, 256, 2)
    chunks = lut.shape[0]
    if m > 63 or chunks != (m + 7) // 8:
        raise ValueError('table does not match the bit count')
    top = 1 << 63

    def entry(z, k):
        acc = 0
        for c in range(chunks):
            acc ^= int(lut[c, (z >> (8 * c)) & 255, k])
        return acc

    def q(z):
        return bin(entry(z, 1) & (z | top)).count('1') & 1

    q0 = q(0)
    qa = [q(1 << a) for a in range(m)]
    rows = np.zeros(m, np.uint64)
    cols = np.zeros(m, np.uint64)
    for a in range(m):
        cols[a] = entry(1 << a, 0)
        r = (qa[a] ^ q0) << a
        for b in range(m):
            if b != a and q((1 << a) | (1 << b)) ^ qa[a] ^ qa[b] ^ q0:
                r |= 1 << b
        rows[a] = r
    return rows, cols


def graph_phase_lut(g):
    g = np.ascontiguousarray(np.asarray(g, np.uint8))
    n = len(g)
    lut = np.zeros(((n + 7) // 8) * 256, np.uint64)
    call('rgs_graph_phase_lut', ptr(g), n, ptr(lut))
    return lut


def gf2_rank(m):
    import ctypes
    m = np.ascontiguousarray(np.asarray(m, np.uint8))
    r = ctypes.c_int()
    call('rgs_gf2_rank', ptr(m), m.shape[0], m.shape[1], ctypes.byref(r))
    return r.value


def gf2_inverse(m):
    m = np.ascontiguousarray(np.asarray(m, np.uint8))
    out = np.zeros_like(m)
    call('rgs_gf2_inverse', ptr(m), len(m), ptr(out))
    return out


class GraphStateSampler:
    """Plain exact sampler on the full state (x)_k u[k] |G> of one graph."""

    PIPELINES = ('slab', 'literal', 'folded')

    def __init__(self, g, u, device=0, pipeline='slab', mass_bits=8, tuning=None, arena=None, defer_leaves=False):
        g = np.asarray(g, np.uint8)
        u = np.asarray(u, np.complex128)
        n = len(g)
        if g.shape != (n, n) or u.shape != (n, 2, 2):
            raise ValueError('graph must be n x n and unitaries n x 2 x 2')
        if np.max(np.abs(u @ np.conj(np.transpose(u, (0, 2, 1))) - np.eye(2))) > 1e-12:
            raise ValueError('measurement matrices are not unitary')
        if pipeline not in self.PIPELINES:
            raise ValueError(f'pipeline must be one of {self.PIPELINES}')
        self.g, self.u, self.n, self.pipeline = g, u, n, pipeline
        # defer_leaves: the last pass writes the masses only; draws and reads
        # recompute the tiles they touch, host() finishes the state
        self.defer_leaves = bool(defer_leaves)
        self._pending = False
        self.slab = _Slab(n, device, mass_bits, tuning, arena=arena)
        sl = self.slab
        self._lut = sl.to_device(graph_phase_lut(g))
        self._rows = form_rows(g)
        self._gates = Gates(sl, u @ HADAMARD if pipeline == 'slab' else u)
        self._desc = self._make_desc()
        self.passes, self.plan_text = sl.describe(self._desc)
        self._norm = None

    def _make_desc(self):
        """Descriptor of this pipeline with the slab's current tuning."""
        sl, n = self.slab, self.n
        if self.pipeline == 'slab':
            d = sl.desc(_lib.INIT_GRAPH, 2.0 ** (-n), init_lut=self._lut, init_rows=self._rows, walsh=True,
                        gates=self._gates)
            # each pass: Walsh, D2 phase and rotations of its own bits (the
            # tuning's choice; no sign table follows the Walsh transform here)
            d.fuse_walsh = int(sl.tuning.get('fuse_walsh', 0))
        elif self.pipeline == 'literal':
            # H^n = WHT / 2^{n/2}: the delta carries the normalisation
            d = sl.desc(_lib.INIT_DELTA, 2.0 ** (-n / 2), walsh=True, phase_lut=self._lut,
                        phase_rows=self._rows, gates=self._gates)
        else:
            d = sl.desc(_lib.INIT_GRAPH, 2.0 ** (-n / 2), init_lut=self._lut, init_rows=self._rows,
                        gates=self._gates)
        return _deferred(sl, d, self.defer_leaves)

    def compute(self):
        """Build the state and its mass pyramid; returns stage seconds."""
        t = time.perf_counter()
        self.slab.run(self._desc)
        self.slab.sync()
        self._norm = None
        self._pending = bool(self._desc.defer_leaves)
        return {'state_and_masses': time.perf_counter() - t}

    def finish(self):
        """Write the deferred last pass's values into the state (once per compute)."""
        if self._pending:
            call('rgs_slab_finish', ctypes_ref(self._desc), self.slab.stream.cuda_stream)
            self.slab.sync()
            self._pending = False

    def masses(self, tilebits=None):
        """(mass float64[nmass], norm, tilebits, nmass) as DistributedSampler.masses
        (distributed_sampler.py:145-152), from the device pyramid."""
        tb = min(20, self.n) if tilebits is None else tilebits
        tb = max(tb, self.slab.mass_bits)
        mass = self.slab.level(tb)
        norm = float(mass.sum())
        if not np.isfinite(norm) or abs(norm - 1) > 1e-3:
            raise ArithmeticError(f'Conditional norm check failed: {norm}')
        return mass, norm, tb, len(mass)

    def norm(self):
        r = self.slab.root()
        if not np.isfinite(r) or abs(r - 1) > 1e-3:
            raise ArithmeticError(f'norm check failed: {r}')
        return r

    @property
    def flip(self):
        """Outcome x is held at state index x ^ flip (rgs_gate_shear)."""
        return self._gates.flip

    def sample(self, uniforms):
        """Device qubit-by-qubit draws, one uniform each: (outcomes, p)."""
        if self._pending:
            return self.slab.draw_leaves(self._desc, uniforms)
        return self.slab.draw_many(uniforms, self.flip)

    def draw(self, rng):
        """One draw consuming one uniform: (outcome, p, norm)."""
        norm = self.norm()
        idx, p = self.sample([rng.random()])
        return int(idx[0]), float(p[0]), norm

    def probability_at(self, index):
        if not 0 <= index < (1 << self.n):
            raise ValueError('index outside the state')
        if self._pending:
            v = self.slab.leaf_values(self._desc, [index ^ self.flip])[0]
            return float(np.float64(v.real) ** 2 + np.float64(v.imag) ** 2)
        return float(self.slab.gather_prob([index ^ self.flip])[0])

    def device_amplitudes(self):
        """The device complex64 values before the D1 phases, as complex128, in
        outcome order: the numbers the masses, the pyramid and the draws are
        computed from."""
        self.finish()
        return self._gates.to_outcomes(self.slab.state.cpu().numpy().astype(np.complex128))

    def host(self):
        """The state as complex128 in outcome order (device complex64 values
        times the D1 phases)."""
        self.finish()
        return self._gates.to_outcomes(self._gates.apply_d1(self.slab.state.cpu().numpy().astype(np.complex128)))

    def sync(self):
        self.slab.sync()


def _edge_margin(cum, t, k):
    """Distance of t to the nearer edge of bin k of an inclusive cumulative sum."""
    lo = cum[k - 1] if k > 0 else 0.0
    hi = cum[k] if k < len(cum) else cum[-1]
    return float(min(abs(t - lo), abs(hi - t)))


def _check_unitary(m):
    m = np.asarray(m, np.complex128)
    if np.max(np.abs(m.conj().T @ m - np.eye(2))) > 1e-12:
        raise ValueError('Last product gate is not unitary')
    return m


class ConditionalSampler:
    """Exact conditional Born sampler for a balanced full-rank cut (one device).

    Drop-in for SpecializedDeferredSampler / the deferred fused engines: same
    constructor arguments (chunk_mib, fusion_bits, lane_bits, lookup_bits are
    accepted; tiles come from ``tuning``), same methods.  ``defer_top`` keeps
    the terminal gate on bit m-1 out of the state and applies it in complex128
    at draw time with a third uniform, as DeferredGlobalMixin does.
    """

    def __init__(self, g, u, A, B, devices=(0,), chunk_mib=1, fusion_bits=5, lane_bits=5, lookup_bits=8,
                 defer_top=True, mass_bits=8, tuning=None, weight_bits=None, weight_table=None,
                 defer_leaves=True):
        g = np.ascontiguousarray(np.asarray(g, np.uint8))
        u = np.ascontiguousarray(np.asarray(u, np.complex128))
        A, B = [int(x) for x in A], [int(x) for x in B]
        if len(A) != len(B):
            raise ValueError('balanced partition required')
        if lookup_bits != 8:
            raise ValueError('device lookup tables use 8-bit chunks')
        if len(list(devices)) != 1:
            raise ValueError('one device per engine; shard by running one engine per GPU')
        m = len(A)
        self.g, self.u, self.A, self.B, self.m = g, u, A, B, m
        self.ng = 1
        self.defer_top = defer_top
        # the last pass writes the masses only; the draw recomputes the tiles
        # it reads (as the terminal gate is applied at draw time)
        self.defer_leaves = bool(defer_leaves)
        self._pending = False
        self.slab = _Slab(m, list(devices)[0], mass_bits, tuning, kind='c')
        sl = self.slab
        chunks = (m + 7) // 8
        map_lut = np.zeros(chunks * 256 * 2, np.uint64)
        aphase = np.zeros(chunks * 256, np.uint64)
        Aa = np.asarray(A, np.int32)
        Ba = np.asarray(B, np.int32)
        call('rgs_cov_luts', ptr(g), len(g), ptr(Aa), ptr(Ba), m, ptr(map_lut), ptr(aphase))
        self._map = sl.to_device(map_lut)
        self._cov_rows, self._cov_cols = cov_form(map_lut, m)
        self._aphase = sl.to_device(aphase)
        self._arows = form_rows(g[np.ix_(A, A)])
        self._Ba = Ba
        self._u = np.ascontiguousarray(u)
        self._gates = Gates(sl, u[A], m - 1 if defer_top else m)
        if weight_bits not in (None, 8, 9):
            raise ValueError('weight_bits must be 8 or 9')
        self._weight_bits = weight_bits
        # identical B unitaries: the weight of b depends on two popcounts
        # (InitSpecialization.weight_table), one table read per amplitude
        same = bool(np.all(u[Ba] == u[Ba[0]]))
        if weight_table is None:
            weight_table = same and weight_bits is None
        if weight_table and not same:
            raise ValueError('the popcount weight table needs identical B unitaries')
        self._weight_table = bool(weight_table)
        # room for either chunking or the popcount table (the descriptor names the one in use)
        self._wlut_host = np.zeros(max(((m + 7) // 8) << 8, ((m + 8) // 9) << 9, (m + 1) ** 2) * 2, np.float32)
        self._wlut = sl.to_device(self._wlut_host)
        self.deferred_global_unitary = _check_unitary(u[A[-1]]) if defer_top else None
        self.n = 1 << (m - 1)          # shard size of the reference's two-device layout
        self._desc = self._make_desc()
        self.passes, self.plan_text = sl.describe(self._desc)

    def _make_desc(self):
        """Descriptor of the conditional pipeline with the slab's current tuning.
        Weight chunks are 9 bits wide where that saves a chunk and the
        contiguous tile leaves LDS room for the larger table, else 8."""
        m = self.m
        wbits = self._weight_bits
        if wbits is None:
            low = effective_low_bits(m, self.slab.tuning['low_bits'])
            wbits = 9 if (m + 8) // 9 < (m + 7) // 8 and low <= 12 else 8
        d = self.slab.desc(_lib.INIT_COV, 0.0, init_lut=self._map, init_wlut=self._wlut, walsh=True,
                           init_rows=self._cov_rows, init_cols=self._cov_cols, phase_lut=self._aphase,
                           phase_rows=self._arows, gates=self._gates,
                           gate_limit=m - 1 if self.defer_top else m, init_wbits=wbits)
        d.init_wpop = int(self._weight_table)
        return _deferred(self.slab, d, self.defer_leaves)

    def compute(self, outB):
        if len(outB) != self.m:
            raise ValueError('wrong output condition length')
        ob = np.asarray(outB, np.uint8)
        if np.any(ob > 1):
            raise ValueError('outB must contain bits')
        stage = {}
        t = time.perf_counter()
        if self._weight_table:
            call('rgs_cov_weight_table', ptr(self._u[self._Ba[0]]), ptr(ob), self.m, 2.0 ** (-self.m / 2),
                 ptr(self._wlut_host))
            self._desc.init_omask = int(sum(1 << j for j in range(self.m) if ob[j]))
        else:
            call('rgs_cov_weight_lut_bits', ptr(self._u), len(self.g), ptr(self._Ba), ptr(ob), self.m,
                 2.0 ** (-self.m / 2), self._desc.init_wbits, ptr(self._wlut_host))
        self._wlut.copy_(self.slab.torch.as_tensor(self._wlut_host))
        stage['weights'] = time.perf_counter() - t
        t = time.perf_counter()
        self.slab.run(self._desc)
        self.slab.sync()
        self._pending = bool(self._desc.defer_leaves)
        stage['state_and_masses'] = time.perf_counter() - t
        stage['linear_permutation_phase'] = 0.0
        return stage

    def finish(self):
        """Write the deferred last pass's values into the state (once per compute)."""
        if self._pending:
            call('rgs_slab_finish', ctypes_ref(self._desc), self.slab.stream.cuda_stream)
            self.slab.sync()
            self._pending = False

    def masses(self):
        """Tile masses (tiles of 2^min(20, m-1) inside each half), norm, tilebits, nmass per half."""
        tb = min(20, self.m - 1)
        tb = max(tb, self.slab.mass_bits)
        mass = self.slab.level(tb)
        norm = float(mass.sum())
        if not np.isfinite(norm) or abs(norm - 1) > 1e-3:
            raise ArithmeticError(f'Conditional norm check failed: {norm}')
        return mass, norm, tb, len(mass) // 2

    def _slice(self, start, stop):
        if self._pending:
            return self.slab.leaf_block(self._desc, start, stop)
        return self.slab.state[start:stop].cpu().numpy()

    def _outcome_slice(self, start, stop):
        """Outcomes start..stop-1 (an aligned power-of-two block) of the device
        state: the block of state indices x ^ flip, reordered."""
        flip = self._gates.flip
        if not flip:
            return self._slice(start, stop)
        size = stop - start
        lo = flip & (size - 1)
        s0 = start ^ (flip & ~(size - 1))
        blk = self._slice(s0, s0 + size)
        return blk[np.arange(size, dtype=np.int64) ^ lo]

    def _pair(self, local):
        """The terminal gate applied to the two amplitudes at state index local."""
        raw = np.array([complex(self._slice(local, local + 1)[0]),
                        complex(self._slice(self.n + local, self.n + local + 1)[0])], np.complex128)
        return self.deferred_global_unitary @ raw

    def transformed_pair(self, local_index):
        """The two amplitudes of local outcome local_index after the terminal gate."""
        if not 0 <= local_index < self.n:
            raise ValueError('Local index outside shard')
        return self._pair(local_index ^ self._gates.flip)

    def draw(self, rng):
        """Reference draw: tile by searchsorted over FP64 tile masses, index in
        the tile by searchsorted over float64 p, then (deferred) the terminal
        bit; exactly 3 uniforms when deferred (distributed_sampler.py:153-157,
        deferred_global_sampler.py:72-78)."""
        mass, norm, tb, nmass = self.masses()
        cm = np.cumsum(mass)
        t1 = rng.random() * norm
        block = int(np.searchsorted(cm, t1, side='right'))
        half, chunk = block // nmass, block % nmass
        start = half * self.n + (chunk << tb)
        tile = self._outcome_slice(start, min(start + (1 << tb), (half + 1) * self.n))
        p = tile.real.astype(np.float64) ** 2 + tile.imag.astype(np.float64) ** 2
        cp = np.cumsum(p)
        t2 = rng.random() * p.sum()
        within = int(np.searchsorted(cp, t2, side='right'))
        index = start + within
        # distance of each target to the nearer edge of its bin, relative to
        # the level's total (how far the FP32 state may move before the draw
        # would pick another bin)
        self.last_margins = [_edge_margin(cm, t1, block) / norm, _edge_margin(cp, t2, within) / p.sum()]
        if not self.defer_top:
            return index, float(p[within]), norm
        local = index % self.n
        pair = self.transformed_pair(local)
        pp = np.abs(pair) ** 2
        total = float(pp.sum())
        if not np.isfinite(total) or total <= 0:
            raise ArithmeticError('Invalid transformed local mass')
        t3 = rng.random() * total
        bit = int(t3 >= pp[0])
        self.last_margins.append(abs(t3 - pp[0]) / total)
        return bit * self.n + local, float(pp[bit]), norm

    def sample(self, uniforms, bit_uniforms=None):
        """Batched draws: the device descent of the mass pyramid picks each
        index of the pre-terminal state (distribution of the local index
        exactly as the reference's tile-then-index draw), then, when the top
        gate is deferred, the two amplitudes at that local index go through
        the terminal gate in complex128 and one more uniform picks the top bit
        (DeferredGlobalMixin.draw, deferred_global_sampler.py:72-78).
        Returns (indices, unnormalised p)."""
        flip = self._gates.flip
        if self._pending:
            idx, p = self.slab.draw_leaves(self._desc, uniforms)
        else:
            idx, p = self.slab.draw_many(uniforms, flip)
        if not self.defer_top:
            return idx, p
        if bit_uniforms is None or len(bit_uniforms) != len(idx):
            raise ValueError('one bit uniform per draw is needed with the deferred top gate')
        local = (idx % self.n) ^ flip     # state indices
        if self._pending:
            lo = self.slab.leaf_values(self._desc, local).astype(np.complex128)
            hi = self.slab.leaf_values(self._desc, local + self.n).astype(np.complex128)
        else:
            torch = self.slab.torch
            lt = torch.as_tensor(local).to(self.slab.device)
            st = self.slab.state
            lo = st.index_select(0, lt).cpu().numpy().astype(np.complex128)
            hi = st.index_select(0, lt + self.n).cpu().numpy().astype(np.complex128)
        U = self.deferred_global_unitary
        a = U[0, 0] * lo + U[0, 1] * hi
        b = U[1, 0] * lo + U[1, 1] * hi
        pa, pb = np.abs(a) ** 2, np.abs(b) ** 2
        bit = (np.asarray(bit_uniforms) * (pa + pb) >= pa).astype(np.int64)
        return bit * self.n + (idx % self.n), np.where(bit == 1, pb, pa)

    def probability_at(self, index):
        if self.defer_top:
            bit = int(index // self.n)
            if bit not in (0, 1):
                raise ValueError('Bad global index')
            return float(abs(self.transformed_pair(index % self.n)[bit]) ** 2)
        if self._pending:
            v = self.slab.leaf_values(self._desc, [index ^ self._gates.flip])[0]
            return float(np.float64(v.real) ** 2 + np.float64(v.imag) ** 2)
        return float(self.slab.gather_prob([index ^ self._gates.flip])[0])

    def host(self):
        """The conditional state as complex128 in outcome order."""
        self.finish()
        raw = self._gates.apply_d1(self.slab.state.cpu().numpy().astype(np.complex128))
        if self.defer_top:
            blk = raw.reshape(2, -1)
            a, b = blk[0].copy(), blk[1].copy()
            U = self.deferred_global_unitary
            blk[0] = U[0, 0] * a + U[0, 1] * b
            blk[1] = U[1, 0] * a + U[1, 1] * b
        return self._gates.to_outcomes(raw)

    def sync(self):
        self.slab.sync()
