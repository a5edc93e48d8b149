"""Learned proposal: a MADE network trained and sampled with this build's kernels.

Architecture (BF16 weights and activations, FP32 accumulation, sigmoid
conditionals): n binary inputs -> H -> H -> n logits (H = 512 in the
learned-proposal runs: 32 degrees of one aligned 16-unit block each), masked so that
logit o depends only on inputs before o in qubit order (MADE, Germain et al.,
arXiv:1502.03509).  Training is Adam on the mean binary cross-entropy of
bitstrings drawn by the exact sampler; every GEMM, the loss, the bias
gradients and the masked Adam update (FP32 master weights, BF16 working
copy) run in librgs_amd.  Sampling is autoregressive (n forward passes) with
host-drawn uniforms, and returns log q(x) for the acceptance test.
No reference counterpart: the reference proposes uniformly
(work/rgs_attack/rejection_sampler.py:21).
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import call, ptr

_V = ctypes.c_void_p
_I = ctypes.c_int
_F = ctypes.c_float
_lib.register('rgs_gemm_bf16', _I, [_I, _I, _I, _I, _I, _V, _I, _V, _I, _F, _V, _I, _V, _I, _V, _I, _V, _I, _I, _V])
_lib.register('rgs_made_bce', _I, [_V, _V, _I, _I, _F, _V, _V, _V, _V])
_lib.register('rgs_colsum', _I, [_V, _I, _I, _I, _V, _I, _V])
_lib.register('rgs_colsum_bf16', _I, [_V, _I, _I, _I, _V, _V])
_lib.register('rgs_adam', _I, [_V, _V, _V, _V, _V, _V, ctypes.c_longlong, _V, _F, _F, _F, _F, _V])
_lib.register('rgs_sample_col', _I, [_V, _V, _I, _I, _I, _V, _V])
_lib.register('rgs_f32_to_bf16', _I, [_V, _V, ctypes.c_longlong, _V])
_lib.register('rgs_made_sample_step', _I, [_V, _I, _V, _I, _I, _I, _V, _V, _I, _V])
_lib.register('rgs_made_sample', _I, [_V, _V, _V, _V, _V, _V, _V, _V, _I, _I, _I, _V, _V, _V, _V])
_lib.register('rgs_made_sample_pack', _I, [_V, _V, _V, _I, _I, _V, ctypes.POINTER(ctypes.c_uint64), _V])
_lib.register('rgs_made_sample_packed', _I, [_V, _V, _V, _V, _V, _I, _I, _I, _V, _V, _V, _V])


def hidden_degrees(n, hidden):
    """Sorted hidden-unit degrees in 1 .. n-1, so the units of degree <= d form
    a prefix (incremental sampling relies on it).  When the width allows,
    every degree owns whole aligned 16-unit blocks (the one-wave sampler
    needs it): with hidden = 16 m (n-1), unit i has degree i // (16 m) + 1;
    with hidden = 16 k and k < n-1 blocks, unit i has degree i // 16 + 1
    (degrees above k get no units; output n-1 still sees every unit).
    Otherwise unit i has degree floor(i (n-1) / hidden) + 1."""
    i = np.arange(hidden)
    d = max(1, n - 1)
    if hidden % 16 == 0:
        blocks = hidden // 16
        if blocks % d == 0:
            return i // (16 * (blocks // d)) + 1
        if blocks < d:
            return i // 16 + 1
    return (i * d) // hidden + 1


def made_masks(n, hidden, seed=0):
    """Degrees: inputs 1..n; hidden degrees from hidden_degrees, in sorted
    order (``seed`` is kept for the signature; the assignment is deterministic).
    Returns float masks [H,n], [H,H], [n,H]."""
    deg_in = np.arange(1, n + 1)
    d1 = hidden_degrees(n, hidden)
    d2 = hidden_degrees(n, hidden)
    m1 = (d1[:, None] >= deg_in[None, :]).astype(np.float32)
    m2 = (d2[:, None] >= d1[None, :]).astype(np.float32)
    m3 = (deg_in[:, None] > d2[None, :]).astype(np.float32)
    return m1, m2, m3


def bits_to_bf16(bits):
    """uint8 0/1 -> bfloat16 bit patterns (0x0000 / 0x3f80) as int16."""
    b = np.asarray(bits, np.uint8)
    return (b.astype(np.uint16) * np.uint16(0x3F80)).view(np.int16)


class MADE:
    def __init__(self, n, hidden=512, batch=512, device=0, seed=0, lr=1e-3):
        import torch
        if not torch.cuda.is_available():
            raise RuntimeError('no HIP device visible')
        _lib.lib()
        self.torch = torch
        self.n, self.H, self.B = n, hidden, batch
        self.lr = lr
        self.dev = torch.device('cuda', device)
        self.stream = torch.cuda.current_stream(self.dev)
        n_, H = n, hidden
        self.shapes = [(H, n_), (H,), (H, H), (H,), (n_, H), (n_,)]
        sizes = [int(np.prod(s)) for s in self.shapes]
        self.offsets = np.concatenate([[0], np.cumsum(sizes)]).astype(int)
        P = int(self.offsets[-1])
        self.P = P
        rng = np.random.default_rng(seed)
        m1, m2, m3 = made_masks(n, hidden, seed)
        self.masks_np = (m1, m2, m3)
        w = np.zeros(P, np.float32)
        mask = np.ones(P, np.float32)
        for k, (shape, m) in enumerate(zip(self.shapes[0::2], (m1, m2, m3))):
            fan_in = shape[1]
            lo = self.offsets[2 * k]
            W = rng.uniform(-1, 1, size=shape).astype(np.float32) / np.sqrt(fan_in)
            w[lo:lo + W.size] = (W * m).reshape(-1)
            mask[lo:lo + W.size] = m.reshape(-1)
        f32, bf = torch.float32, torch.int16
        self.w = torch.as_tensor(w).to(self.dev)
        self.mask = torch.as_tensor(mask).to(self.dev)
        self.wb = torch.empty(P, dtype=bf, device=self.dev)
        call('rgs_f32_to_bf16', ptr(self.w), ptr(self.wb), P, self.stream.cuda_stream)
        self.g = torch.zeros(P, dtype=f32, device=self.dev)
        self.m = torch.zeros(P, dtype=f32, device=self.dev)
        self.v = torch.zeros(P, dtype=f32, device=self.dev)
        self.t = torch.zeros(1, dtype=torch.int32, device=self.dev)
        self._buffers(batch)

    def _buffers(self, B):
        torch = self.torch
        n, H = self.n, self.H
        bf, f32 = torch.int16, torch.float32
        self.bufB = B
        self.x = torch.zeros((B, n), dtype=bf, device=self.dev)
        self.h1 = torch.empty((B, H), dtype=bf, device=self.dev)
        self.h2 = torch.empty((B, H), dtype=bf, device=self.dev)
        self.logits = torch.empty((B, n), dtype=f32, device=self.dev)
        self.nll = torch.empty(B, dtype=f32, device=self.dev)
        self.dl = torch.empty((B, n), dtype=f32, device=self.dev)
        self.dlb = torch.empty((B, n), dtype=bf, device=self.dev)
        self.da2 = torch.empty((B, H), dtype=f32, device=self.dev)
        self.da2b = torch.empty((B, H), dtype=bf, device=self.dev)
        self.da1 = torch.empty((B, H), dtype=f32, device=self.dev)
        self.da1b = torch.empty((B, H), dtype=bf, device=self.dev)

    # views into the flat vectors (element offsets)
    def _p(self, tensor, k):
        return ptr(tensor) + int(self.offsets[k]) * tensor.element_size()

    def _gemm(self, ta, tb, M, N, K, A, lda, Bm, ldb, C=None, ldc=0, Cb=None, ldcb=0, bias=None, relu=0,
              dmask=None, ldm=0, acc=0, alpha=1.0, stream=None):
        call('rgs_gemm_bf16', ta, tb, M, N, K, A, lda, Bm, ldb, alpha, C, ldc, Cb, ldcb, bias, relu, dmask, ldm,
             acc, self.stream.cuda_stream if stream is None else stream)

    def forward(self, B=None):
        """Forward on self.x[:B]: h1, h2, logits."""
        B = B or self.bufB
        n, H = self.n, self.H
        wb = self.wb
        self._gemm(0, 1, B, H, n, ptr(self.x), n, self._p(wb, 0), n, Cb=ptr(self.h1), ldcb=H,
                   bias=self._p(self.w, 1), relu=1)
        self._gemm(0, 1, B, H, H, ptr(self.h1), H, self._p(wb, 2), H, Cb=ptr(self.h2), ldcb=H,
                   bias=self._p(self.w, 3), relu=1)
        self._gemm(0, 1, B, n, H, ptr(self.h2), H, self._p(wb, 4), H, C=ptr(self.logits), ldc=n,
                   bias=self._p(self.w, 5))

    def loss(self, B=None, grad=True):
        B = B or self.bufB
        call('rgs_made_bce', ptr(self.logits), ptr(self.x), B, self.n, 1.0 / B, ptr(self.nll),
             ptr(self.dl) if grad else None, ptr(self.dlb) if grad else None, self.stream.cuda_stream)

    def backward(self, B=None):
        B = B or self.bufB
        n, H = self.n, self.H
        g, wb, s = self.g, self.wb, self.stream.cuda_stream
        self._gemm(1, 0, n, H, B, ptr(self.dlb), n, ptr(self.h2), H, C=self._p(g, 4), ldc=H)
        call('rgs_colsum', ptr(self.dl), B, n, n, self._p(g, 5), 0, s)
        self._gemm(0, 0, B, H, n, ptr(self.dlb), n, self._p(wb, 4), H, C=ptr(self.da2), ldc=H, Cb=ptr(self.da2b),
                   ldcb=H, dmask=ptr(self.h2), ldm=H)
        call('rgs_colsum', ptr(self.da2), B, H, H, self._p(g, 3), 0, s)
        self._gemm(1, 0, H, H, B, ptr(self.da2b), H, ptr(self.h1), H, C=self._p(g, 2), ldc=H)
        self._gemm(0, 0, B, H, H, ptr(self.da2b), H, self._p(wb, 2), H, C=ptr(self.da1), ldc=H, Cb=ptr(self.da1b),
                   ldcb=H, dmask=ptr(self.h1), ldm=H)
        call('rgs_colsum', ptr(self.da1), B, H, H, self._p(g, 1), 0, s)
        self._gemm(1, 0, H, n, B, ptr(self.da1b), H, ptr(self.x), n, C=self._p(g, 0), ldc=n)

    def adam(self):
        call('rgs_adam', ptr(self.w), ptr(self.g), ptr(self.mask), ptr(self.m), ptr(self.v), ptr(self.wb), self.P,
             ptr(self.t), self.lr, 0.9, 0.999, 1e-8, self.stream.cuda_stream)
        self._packed_stale = True

    def train_step(self):
        self.forward()
        self.loss()
        self.backward()
        self.adam()

    def load_batch(self, bits_bf16_dev):
        """Copy a [B, n] int16 (bf16 0/1) device tensor into the input buffer."""
        self.x.copy_(bits_bf16_dev)

    def train(self, data_bits, steps, seed=0, log_every=0):
        """data_bits: uint8 [N, n] exact samples.  Epochs are reshuffled on the
        host and uploaded once; each step copies its slice into the static
        input buffer (device to device)."""
        torch = self.torch
        N = len(data_bits)
        B = self.bufB
        rng = np.random.default_rng(seed)
        per_epoch = N // B
        epochs = -(-steps // per_epoch)
        order = np.concatenate([rng.permutation(N)[:per_epoch * B] for _ in range(epochs)])[:steps * B]
        dev = torch.as_tensor(bits_to_bf16(data_bits[order])).to(self.dev)
        losses = []
        for k in range(steps):
            self.x.copy_(dev[k * B:(k + 1) * B])
            self.train_step()
            if log_every and (k + 1) % log_every == 0:
                losses.append(float(self.nll.mean().item()))
        return losses

    def log_q(self, bits, chunk=None):
        """log q(x) (natural log) of uint8 bitstrings [count, n]."""
        torch = self.torch
        bits = np.asarray(bits, np.uint8)
        out = np.empty(len(bits))
        B = self.bufB
        for s0 in range(0, len(bits), B):
            part = bits[s0:s0 + B]
            b = len(part)
            self.x[:b].copy_(torch.as_tensor(bits_to_bf16(part)))
            self.forward(b)
            self.loss(b, grad=False)
            out[s0:s0 + b] = -self.nll[:b].double().cpu().numpy()
        return out

    def _sampling_buffers(self, B):
        sb = getattr(self, '_sbuf', None)
        if sb is not None and sb['B'] >= B:
            return sb
        torch = self.torch
        n, H = self.n, self.H
        self._sbuf = sb = dict(B=B, x=torch.zeros((B, n), dtype=torch.int16, device=self.dev),
                               h1=torch.empty((B, H), dtype=torch.int16, device=self.dev),
                               h2=torch.empty((B, H), dtype=torch.int16, device=self.dev),
                               logit=torch.empty(B, dtype=torch.float32, device=self.dev),
                               logq=torch.empty(B, dtype=torch.float64, device=self.dev),
                               u=torch.empty((B, n), dtype=torch.float64, device=self.dev),
                               words=torch.empty((B, 2), dtype=torch.int64, device=self.dev))
        return sb

    def sample_fused(self, b, bits=False):
        """Draw b proposals from the uniforms in sb['u'][:b] with the fused
        sampling kernel in one launch, with the same arithmetic as the
        incremental GEMM sequence, for widths whose degrees own whole aligned
        16-unit blocks (H = 16 m (n - 1), or H = 16 k with k < n - 1; see
        hidden_degrees): one wavefront per 16 proposals, no workgroup
        barrier.  The weights are read from a fragment-major copy, made again
        after any training step.  Returns device views (packed words int64
        [b, 2] in the scan's layout, log q float64 [b]) and, with bits=True,
        also writes the bf16 0/1 bits to sb['x'][:b]."""
        sb = self._sbuf
        wb, w = self.wb, self.w
        s = self.stream.cuda_stream
        if getattr(self, '_packed', None) is None or getattr(self, '_packed_stale', True):
            size = ctypes.c_uint64()
            call('rgs_made_sample_pack', None, None, None, self.n, self.H, None, ctypes.byref(size), s)
            if getattr(self, '_packed', None) is None:
                self._packed = self.torch.empty(int(size.value), dtype=self.torch.uint8, device=self.dev)
            call('rgs_made_sample_pack', self._p(wb, 0), self._p(wb, 2), self._p(wb, 4), self.n, self.H,
                 ptr(self._packed), ctypes.byref(size), s)
            self._packed_stale = False
        call('rgs_made_sample_packed', ptr(self._packed), self._p(w, 1), self._p(w, 3), self._p(w, 5), ptr(sb['u']),
             b, self.n, self.H, ptr(sb['words']), ptr(sb['x']) if bits else None, ptr(sb['logq']), s)
        return sb['words'][:b], sb['logq'][:b]

    def sample_fused_raw(self, b, bits=False):
        """sample_fused reading the row-major weights in place
        (rgs_made_sample); the same results."""
        sb = self._sbuf
        if not hasattr(self, '_kpref'):
            k = np.searchsorted(hidden_degrees(self.n, self.H), np.arange(self.n), side='right').astype(np.int32)
            self._kpref = self.torch.as_tensor(k).to(self.dev)
        wb, w = self.wb, self.w
        call('rgs_made_sample', self._p(wb, 0), self._p(w, 1), self._p(wb, 2), self._p(w, 3), self._p(wb, 4),
             self._p(w, 5), ptr(self._kpref), ptr(sb['u']), b, self.n, self.H, ptr(sb['words']),
             ptr(sb['x']) if bits else None, ptr(sb['logq']), self.stream.cuda_stream)
        return sb['words'][:b], sb['logq'][:b]

    def _sample_steps(self, b, stream):
        """Enqueue the n incremental steps for the first b rows of the
        sampling buffers (uniforms in sb['u']) on `stream`.

        Hidden degrees are sorted, so at step o the units of degree <= o are a
        prefix of each hidden layer and only the units of degree exactly o are
        new.  Step o computes those units of both layers (their inputs are
        final) and logit o alone: about H^2 + H n multiply-adds per proposal
        in total instead of n full forward passes.  log q accumulates
        log sigmoid(+-l_o) of the drawn bits from the same logits."""
        n, H = self.n, self.H
        sb = self._sbuf
        k = np.searchsorted(hidden_degrees(n, H), np.arange(n), side='right')   # units of degree <= o
        wb, w = self.wb, self.w
        W1, W2, W3 = self._p(wb, 0), self._p(wb, 2), self._p(wb, 4)
        b1, b2, b3 = self._p(w, 1), self._p(w, 3), self._p(w, 5)
        x, h1, h2 = ptr(sb['x']), ptr(sb['h1']), ptr(sb['h2'])
        logit, logq, u = ptr(sb['logit']), ptr(sb['logq']), ptr(sb['u'])
        for o in range(n):
            k0, k1 = (int(k[o - 1]) if o else 0), int(k[o])
            if k1 > k0:
                # new units of degree o: inputs 0..o-1, then the first-layer prefix
                self._gemm(0, 1, b, k1 - k0, o, x, n, W1 + 2 * k0 * n, n, Cb=h1 + 2 * k0, ldcb=H,
                           bias=b1 + 4 * k0, relu=1, stream=stream)
                self._gemm(0, 1, b, k1 - k0, k1, h1, H, W2 + 2 * k0 * H, H, Cb=h2 + 2 * k0, ldcb=H,
                           bias=b2 + 4 * k0, relu=1, stream=stream)
            self._gemm(0, 1, b, 1, k1, h2, H, W3 + 2 * o * H, H, C=logit, ldc=1, bias=b3 + 4 * o, stream=stream)
            call('rgs_made_sample_step', logit, 1, u, b, n, o, x, logq, int(o == 0), stream)

    def sample_device(self, b, graph=True):
        """Draw b proposals from the uniforms already in the sampling buffer
        sb['u'][:b] (device float64 [b, n]); returns device views (x as bf16
        0/1 [b, n], log q float64 [b]).  With graph=True the n steps are
        captured once per b and replayed."""
        torch = self.torch
        sb = self._sbuf
        if not graph:
            self._sample_steps(b, self.stream.cuda_stream)
            return sb['x'][:b], sb['logq'][:b]
        graphs = self.__dict__.setdefault('_graphs', {})
        if b not in graphs:
            side = torch.cuda.Stream(self.dev)
            side.wait_stream(torch.cuda.current_stream(self.dev))
            with torch.cuda.stream(side):
                self._sample_steps(b, side.cuda_stream)            # warm-up outside capture
            torch.cuda.current_stream(self.dev).wait_stream(side)
            torch.cuda.synchronize(self.dev)
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g):
                self._sample_steps(b, torch.cuda.current_stream(self.dev).cuda_stream)
            graphs[b] = g
        graphs[b].replay()
        return sb['x'][:b], sb['logq'][:b]

    def sample(self, uniforms, batch=16384):
        """Autoregressive samples from q with uniforms [count, n] (float64):
        x_o = [u_o < sigmoid(l_o)].  Returns (bits uint8 [count, n], log q).
        Incremental steps (see _sample_steps), in batches of `batch`."""
        torch = self.torch
        u_all = np.ascontiguousarray(uniforms, np.float64)
        count = len(u_all)
        sb = self._sampling_buffers(min(batch, max(count, 1)))
        B = sb['B']
        bits = np.empty((count, self.n), np.uint8)
        out_logq = np.empty(count)
        for s0 in range(0, count, B):
            b = min(B, count - s0)
            sb['u'][:b].copy_(torch.as_tensor(u_all[s0:s0 + b]))
            self._sample_steps(b, self.stream.cuda_stream)
            bits[s0:s0 + b] = (sb['x'][:b].cpu().numpy() != 0).astype(np.uint8)
            out_logq[s0:s0 + b] = sb['logq'][:b].cpu().numpy()
        return bits, out_logq

    def sample_full(self, uniforms):
        """Autoregressive samples with a full forward pass per column (the
        plain method; kept to check the incremental sampler)."""
        torch = self.torch
        u_all = np.ascontiguousarray(uniforms, np.float64)
        count = len(u_all)
        B = self.bufB
        bits = np.empty((count, self.n), np.uint8)
        logq = np.empty(count)
        for s0 in range(0, count, B):
            b = min(B, count - s0)
            u = torch.as_tensor(u_all[s0:s0 + b]).to(self.dev)
            self.x.zero_()
            for o in range(self.n):
                self.forward(b)
                call('rgs_sample_col', ptr(self.logits), ptr(u), b, self.n, o, ptr(self.x), self.stream.cuda_stream)
            self.forward(b)
            self.loss(b, grad=False)
            bits[s0:s0 + b] = (self.x[:b].cpu().numpy() != 0).astype(np.uint8)
            logq[s0:s0 + b] = -self.nll[:b].double().cpu().numpy()
        return bits, logq

    def params_numpy(self):
        """FP32 master weights as float64 arrays in (W1, b1, W2, b2, W3, b3) order."""
        w = self.w.cpu().numpy().astype(np.float64)
        return [w[self.offsets[k]:self.offsets[k + 1]].reshape(self.shapes[k]) for k in range(6)]

    def bf16_params_numpy(self):
        """The BF16 working weights as float64 arrays (what the GEMMs use)."""
        wb = self.wb.cpu().numpy().view(np.uint16).astype(np.uint32) << 16
        w = wb.view(np.float32).astype(np.float64)
        return [w[self.offsets[k]:self.offsets[k + 1]].reshape(self.shapes[k]) for k in range(6)]

    def grads_numpy(self):
        g = self.g.cpu().numpy().astype(np.float64)
        return [g[self.offsets[k]:self.offsets[k + 1]].reshape(self.shapes[k]) for k in range(6)]
