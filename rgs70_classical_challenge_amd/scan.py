"""Batched rank-scan scorer: the drop-in for the reference's ScanOracle.

``ScanOracle(g, u, order, device=0, fp64=False)`` keeps the reference's
constructor and ``amplitude(bits)`` (work/rgs_attack/scan_oracle.py:11-29)
and adds ``amplitudes(bits_batch)`` / ``probabilities(bits_batch)``, which
score thousands of proposals in one persistent launch (librgs_amd
rgs_scan_batch).  The plan is compiled by the library's C++ planner;
``order_search`` is its annealer (the reference's order_optimization/search.cpp).
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import call, ptr


class ScanStep(ctypes.Structure):
    _fields_ = [('v', ctypes.c_int32), ('r_in', ctypes.c_int32), ('r_out', ctypes.c_int32),
                ('nker', ctypes.c_int32), ('beta', ctypes.c_uint64),
                ('lcols', ctypes.c_uint64 * 64), ('kern', ctypes.c_uint64 * 64)]


_V = ctypes.c_void_p
_lib.register('rgs_scan_compile', ctypes.c_int, [_V, ctypes.c_int, _V, _V, _V])
_lib.register('rgs_scan_profile', ctypes.c_int, [_V, ctypes.c_int, _V, _V])
_lib.register('rgs_scan_order_search', ctypes.c_int,
              [_V, ctypes.c_int, _V, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int, ctypes.c_int, _V,
               ctypes.POINTER(ctypes.c_uint64)])
_lib.register('rgs_scan_pack', ctypes.c_int, [_V, ctypes.c_int, _V, ctypes.POINTER(ctypes.c_uint64)])
_lib.register('rgs_scan_batch', ctypes.c_int,
              [_V, ctypes.c_int, _V, _V, ctypes.c_int, ctypes.c_int, _V, _V, _V])
_lib.register('rgs_capped_decide', ctypes.c_int,
              [_V, _V, _V, _V, ctypes.c_int, ctypes.c_int, ctypes.c_double, _V, _V, _V])
_lib.register('rgs_scan_pack_tables', ctypes.c_int, [_V, ctypes.c_int, _V, ctypes.POINTER(ctypes.c_uint64)])
_lib.register('rgs_scan_large', ctypes.c_int,
              [_V, ctypes.c_int, _V, _V, _V, _V, ctypes.c_int, _V, _V, ctypes.c_uint64, _V, _V, _V, _V])
_lib.register('rgs_scan_large_static', ctypes.c_int,
              [_V, ctypes.c_int, _V, _V, _V, _V, ctypes.c_int, _V, _V, _V])

# Largest rank the one-launch batch kernels hold in LDS, per table precision.
BATCH_RMAX = {64: 13, 128: 12}


def _graph(g):
    return np.ascontiguousarray(np.asarray(g, np.uint8))


def profile(g, order):
    """Cut-rank profile of an order (rank after each prefix)."""
    g = _graph(g)
    o = np.ascontiguousarray(np.asarray(order, np.int32))
    out = np.zeros(len(g), np.int32)
    call('rgs_scan_profile', ptr(g), len(g), ptr(o), ptr(out))
    return out.tolist()


def order_search(g, start=None, iterations=200000, seed=20261030, rmax_cap=62, threads=1):
    """Annealed order minimising sum 2^r (max r <= rmax_cap): `threads`
    chains of `iterations` moves each, deterministic for given arguments.
    Returns (order, cost)."""
    g = _graph(g)
    n = len(g)
    s = np.ascontiguousarray(np.arange(n, dtype=np.int32) if start is None else np.asarray(start, np.int32))
    best = np.zeros(n, np.int32)
    cost = ctypes.c_uint64()
    call('rgs_scan_order_search', ptr(g), n, ptr(s), int(iterations), int(seed), int(rmax_cap), int(threads),
         ptr(best), ctypes.byref(cost))
    return best.tolist(), int(cost.value)


def probabilities_from_scores(val, ex, n):
    """p = (re re + im im) 2^(2 ex - n) of the batch kernel's device outputs.
    The power of two is assembled from its exponent bits, so it is exact
    whatever a device exp2 or pow would round to (normal range: p of every
    proposal of a normalised state lies far above 2^-1022)."""
    import torch
    pw =((2 * ex.to(torch.int64) - n + 1023) << 52).view(torch.float64)
    return (val[:, 0] * val[:, 0] + val[:, 1] * val[:, 1]) * pw


def bits_to_words(bits):
    """uint8 [count][n] (qubit q in column q) -> uint64 [count][2]."""
    b = np.asarray(bits, np.uint8)
    if b.ndim == 1:
        b = b[None, :]
    count, n = b.shape
    if n > 128:
        raise ValueError('at most 128 qubits')
    words = np.zeros((count, 2), np.uint64)
    for q in range(n):
        words[:, q >> 6] |= b[:, q].astype(np.uint64) << np.uint64(q & 63)
    return words


class ScanPlan:
    def __init__(self, g, order):
        g = _graph(g)
        n = len(g)
        self.n = n
        self.order = [int(x) for x in order]
        o = np.ascontiguousarray(np.asarray(self.order, np.int32))
        self.steps = (ScanStep * n)()
        prof = np.zeros(n, np.int32)
        call('rgs_scan_compile', ptr(g), n, ptr(o), ctypes.addressof(self.steps), ptr(prof))
        self.ranks = prof.tolist()
        self.rmax = max(self.ranks)
        self.sum_states = sum(1 << r for r in self.ranks)

    def packed(self):
        words = ctypes.c_uint64()
        call('rgs_scan_pack', ctypes.addressof(self.steps), self.n, None, ctypes.byref(words))
        blob = np.zeros(words.value, np.uint64)
        call('rgs_scan_pack', ctypes.addressof(self.steps), self.n, ptr(blob), ctypes.byref(words))
        return blob

    def packed_tables(self):
        """Byte-table plan for the HBM-table scan (ranks up to 32)."""
        words = ctypes.c_uint64()
        call('rgs_scan_pack_tables', ctypes.addressof(self.steps), self.n, None, ctypes.byref(words))
        blob = np.zeros(words.value, np.uint64)
        call('rgs_scan_pack_tables', ctypes.addressof(self.steps), self.n, ptr(blob), ctypes.byref(words))
        return blob

    def table_bytes(self, precision=64):
        """Algorithmic table traffic of one proposal: every step reads its input
        table once and writes its output table once."""
        size = 8 if precision == 64 else 16
        r_in = [0] + self.ranks[:-1]
        return sum(size * ((1 << a) + (1 << b)) for a, b in zip(r_in, self.ranks))


class ScanOracle:
    """p(x) by the linear rank-width scan on one device (drop-in for ScanOracle)."""

    def __init__(self, g, u, order, source=None, device=0, gpu=True, fp64=False, batch_rmax=None):
        """order: the vertex order, or 'auto' for the autotuner's order of
        this graph on this device (tune_scan_order; an annealed order when the
        table has none).  batch_rmax: largest rank scored by the one-launch
        batch kernels (default BATCH_RMAX, their LDS limit); above it
        proposals are scored one at a time with HBM tables."""
        import torch
        if not gpu:
            raise ValueError('this scorer runs on the device only')
        if not torch.cuda.is_available():
            raise RuntimeError('no HIP device visible')
        if isinstance(order, str):
            if order != 'auto':
                raise ValueError("order must be a vertex order or 'auto'")
            from .autotune import scan_order
            order = scan_order(g, torch.cuda.get_device_properties(device)) or \
                order_search(g, iterations=1_000_000, rmax_cap=28, threads=8)[0]
        self.torch = torch
        self.n = len(g)
        self.u = np.ascontiguousarray(np.asarray(u, np.complex128))
        self.order = list(order)
        self.plan = ScanPlan(g, order)
        self.ranks = self.plan.ranks
        self.rmax = self.plan.rmax
        self.steps = self.plan.steps
        self.precision = 128 if fp64 else 64
        self.device = torch.device('cuda', device)
        self.batch_rmax = dict(BATCH_RMAX) if batch_rmax is None else {k: min(v, batch_rmax)
                                                                         for k, v in BATCH_RMAX.items()}
        self._plan = None
        if self.rmax <= max(self.batch_rmax.values()):
            self._plan = torch.as_tensor(self.plan.packed().view(np.int64)).to(self.device)
        self._u = torch.as_tensor(self.u.view(np.float64).reshape(-1)).to(self.device)
        self.stream = torch.cuda.current_stream(self.device)
        self._large = {}

    def _large_state(self, precision):
        """Tables plan on the device, its step offsets on the host, two HBM
        tables of 2^rmax entries and the per-step maxima."""
        torch = self.torch
        if 'plan' not in self._large:
            blob = self.plan.packed_tables()
            self._large['plan'] = torch.as_tensor(blob.view(np.int64)).to(self.device)
            self._large['offsets'] = np.ascontiguousarray(blob[2:2 + self.n])
            self._large['mx'] = torch.zeros(self.n, dtype=torch.int64, device=self.device)
        if precision not in self._large:
            entries = 1 << self.rmax
            dt = torch.complex64 if precision == 64 else torch.complex128
            self._large[precision] = (torch.empty(entries, dtype=dt, device=self.device),
                                      torch.empty(entries, dtype=dt, device=self.device))
        return self._large

    def raw(self, bits, precision=None):
        """Enqueue a batch; returns device tensors (value complex128 pairs,
        exponent).  Ranks within the LDS limit run as one launch for the whole
        batch (rgs_scan_batch); larger ranks run one proposal at a time with
        HBM tables (rgs_scan_large)."""
        torch = self.torch
        prec = precision or self.precision
        words = bits_to_words(bits)
        count = len(words)
        val = torch.empty((count, 2), dtype=torch.float64, device=self.device)
        ex = torch.empty(count, dtype=torch.int32, device=self.device)
        if self.rmax <= self.batch_rmax[prec]:
            dev_bits = torch.as_tensor(words.view(np.int64)).to(self.device)
            call('rgs_scan_batch', ptr(self._plan), self.rmax, ptr(self._u), ptr(dev_bits), count, prec, ptr(val),
                 ptr(ex), self.stream.cuda_stream)
            return val, ex, dev_bits
        st = self._large_state(prec)
        b0, b1 = st[prec]
        for k in range(count):
            w = np.ascontiguousarray(words[k])
            call('rgs_scan_large', ctypes.addressof(self.steps), self.n, ptr(st['plan']), ptr(st['offsets']),
                 ptr(self.u), ptr(w), prec, ptr(b0), ptr(b1), b0.numel(), ptr(st['mx']),
                 ptr(val) + 16 * k, ptr(ex) + 4 * k, self.stream.cuda_stream)
        return val, ex, None

    def scores_device(self, words_dev, precision=None):
        """The batch kernel's outputs for proposals already on the device as
        int64 [count, 2] words (bits_to_words layout): device tensors val
        float64 [count, 2] (re, im) and ex int32 [count], amplitude =
        (re + i im) 2^ex 2^(-n/2); nothing is copied to the host (batch
        kernels only: rank within their LDS limit)."""
        torch = self.torch
        prec = precision or self.precision
        if self.rmax > self.batch_rmax[prec]:
            raise ValueError('device-resident scoring needs the batch kernels (rank within their limit)')
        count = words_dev.shape[0]
        val = torch.empty((count, 2), dtype=torch.float64, device=self.device)
        ex = torch.empty(count, dtype=torch.int32, device=self.device)
        call('rgs_scan_batch', ptr(self._plan), self.rmax, ptr(self._u), ptr(words_dev), count, prec, ptr(val),
             ptr(ex), self.stream.cuda_stream)
        return val, ex

    def probabilities_device(self, words_dev, precision=None):
        """p(x) of proposals already on the device (scores_device) as a
        device float64 tensor, with rgs_capped_decide's arithmetic: each
        square and the sum rounded once, the power of two exact."""
        val, ex = self.scores_device(words_dev, precision)
        return probabilities_from_scores(val, ex, self.n)

    def probabilities_checked(self, words_dev, z_min=0.1):
        """p(x) to 1e-5 relative for device-resident proposals: every proposal
        is scored with complex64 tables (FP64 sums), and those whose
        2^n p(x) falls below ``z_min`` are rescored with complex128 tables in
        one more launch.  Complex64 table rounding is absolute in the
        amplitude (about 2^-24 of the step's largest entry), so the relative
        error of p grows as 1 / sqrt(2^n p): the small-p tail is where the
        complex64 result can miss 1e-5.  Returns device tensors (p float64,
        rescored bool)."""
        torch = self.torch
        if self.rmax > self.batch_rmax[128]:
            raise ValueError('the complex128 rescoring needs the batch kernels (rank within their limit)')
        p = self.probabilities_device(words_dev, 64)
        flag = p < z_min * 2.0 ** -self.n
        idx = torch.nonzero(flag).squeeze(1)
        if idx.numel():
            p[idx] = self.probabilities_device(words_dev.index_select(0, idx).contiguous(), 128)
        return p, flag

    def amplitudes(self, bits, precision=None):
        val, ex, _ = self.raw(bits, precision)
        v = val.cpu().numpy()
        e = ex.cpu().numpy().astype(np.float64)
        scale = np.exp2(e - self.n / 2)
        return (v[:, 0] + 1j * v[:, 1]) * scale

    def probabilities(self, bits, precision=None):
        val, ex, _ = self.raw(bits, precision)
        v = val.cpu().numpy()
        e = ex.cpu().numpy().astype(np.float64)
        return (v[:, 0] ** 2 + v[:, 1] ** 2) * np.exp2(2 * e - self.n)

    def amplitude(self, bits):
        return complex(self.amplitudes(np.asarray(bits, np.uint8)[None, :])[0])


def main(argv):
    """``python -m rgs70_classical_challenge_amd.scan [count [precision [seed]]]``:
    p(x) of ``count`` uniform proposals of the 70-vertex 3-regular instance
    (seeded QR unitaries, the autotuner's order) in one launch each, best of five;
    prints one JSON line with proposals/s."""
    import json
    import torch
    from .instances import instance
    count = int(argv[0]) if argv else 65536
    precision = int(argv[1]) if len(argv) > 1 else 64
    seed = int(argv[2]) if len(argv) > 2 else 70
    g, u = instance(70, seed)
    so = ScanOracle(g, u, 'auto')            # the autotuner's order of this graph
    bits = np.random.default_rng(seed).integers(0, 2, (count, 70), dtype=np.uint8)
    words = torch.as_tensor(bits_to_words(bits).view(np.int64)).to(so.device)
    so.probabilities_device(words, precision)
    torch.cuda.synchronize()
    best = float('inf')
    for _ in range(5):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(so.stream)
        so.probabilities_device(words, precision)
        e1.record(so.stream)
        torch.cuda.synchronize()
        best = min(best, e0.elapsed_time(e1))
    print(json.dumps({'workload': f'70-vertex 3-regular, seed {seed}, {count} uniform proposals per launch',
                      'precision': precision, 'rmax': so.rmax, 'sum_states': so.plan.sum_states,
                      'ms_per_launch': round(best, 3), 'proposals_per_s': round(count / best * 1e3, 1)}), flush=True)


if __name__ == '__main__':
    import sys
    main(sys.argv[1:])
