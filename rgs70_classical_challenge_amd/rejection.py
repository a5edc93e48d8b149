"""Capped rejection sampling, every acceptance decided by the exact p(x).

Uniform proposals follow the reference's rule and RNG schedule
(work/rgs_attack/rejection_sampler.py:19-24): per proposal
``bits = rng.integers(0, 2, n)`` then ``u = rng.random()``, accept iff
``u < min(2^n p(x) / cap, 1)``; the distribution of accepted outputs is
proportional to min(p, cap 2^-n).  Learned proposals use the same rule with
q(x) in place of 2^-n: accept iff ``u < min(p(x) / (cap q(x)), 1)``, output
proportional to min(p, cap q).  Proposals are scored in batches of thousands
by the device scan (complex64 tables); every decision is then re-scored with
complex128 tables and its FP64 margin |u - acceptance| recorded
(audit_all_proposals_fp64.py:43-50).  A decision that differs from the
complex128 one is a flip; flips whose margin exceeds FLIP_TOLERANCE of the
acceptance (the 1e-5 relative p(x) tolerance) are counted apart, since only
those are errors of the scorer rather than ties within its precision.
"""
import math
import time

import numpy as np

from ._lib import call, ptr

FLIP_TOLERANCE = 1e-5


def uniform_proposals(rng, n, count, schedule='reference'):
    """(bits uint8 [count, n], acceptance uniforms [count]).

    'reference' interleaves per proposal exactly as the reference does, so the
    stream replays its recorded proposals; 'batched' draws all bits then all
    uniforms (faster, a different stream)."""
    if schedule == 'reference':
        bits = np.empty((count, n), np.uint8)
        u = np.empty(count)
        for k in range(count):
            bits[k] = rng.integers(0, 2, n)
            u[k] = rng.random()
        return bits, u
    return rng.integers(0, 2, (count, n), dtype=np.uint8), rng.random(count)


def decide(p, logq, u, cap, n):
    """Acceptance probabilities and decisions.  logq None means uniform q."""
    if logq is None:
        acc = np.minimum(2.0 ** n * p / cap, 1.0)
    else:
        acc = np.minimum(p / (cap * np.exp(logq)), 1.0)
    return acc, u < acc


class CappedRejection:
    """Batched capped rejection with uniform or learned proposals."""

    def __init__(self, scorer, auditor, n, cap, proposal=None):
        self.scorer, self.auditor, self.n, self.cap = scorer, auditor, n, cap
        self.proposal = proposal          # a MADE (learned) or None (uniform)

    def run(self, rng, accept_target, batch=8192, max_batches=1000, schedule='batched', record_all=False):
        """Draw until `accept_target` outputs are accepted.  Returns a record
        with accepted bitstrings, p (complex64 scan), p_fp64 (complex128),
        decisions, margins and timing; with ``record_all`` also every
        proposal's bits, uniform, p, p_fp64, acceptance and decision (the
        reference records all of them, rejection_sampler.py:20-23)."""
        import torch
        n = self.n
        acc_bits, acc_p, acc_p64 = [], [], []
        flips = beyond = 0
        flip_margin = 0.0
        margins = []
        every = []
        proposals = 0
        gpu_s = 0.0
        t0 = time.perf_counter()
        for _ in range(max_batches):
            t = time.perf_counter()
            if self.proposal is None:
                bits, u = uniform_proposals(rng, n, batch, schedule)
                logq = None
            else:
                bits, logq = self.proposal.sample(rng.random((batch, n)))
                u = rng.random(batch)
            p = self.scorer.probabilities(bits)
            torch.cuda.synchronize()
            gpu_s += time.perf_counter() - t
            acc, keep = decide(p, logq, u, self.cap, n)
            p64 = self.auditor.probabilities(bits)
            acc64, keep64 = decide(p64, logq, u, self.cap, n)
            flip = keep != keep64
            rel = np.abs(u - acc64) / acc64
            flips += int(np.sum(flip))
            beyond += int(np.sum(flip & (rel > FLIP_TOLERANCE)))
            if flip.any():
                flip_margin = max(flip_margin, float(rel[flip].max()))
            margins.append(np.abs(u - acc64))
            if record_all:
                every.append((bits, u, p, p64, acc, keep, acc64, keep64))
            proposals += batch
            acc_bits.append(bits[keep])
            acc_p.append(p[keep])
            acc_p64.append(p64[keep])
            if sum(len(a) for a in acc_bits) >= accept_target:
                break
        wall = time.perf_counter() - t0
        bits = np.concatenate(acc_bits)[:accept_target]
        p = np.concatenate(acc_p)[:accept_target]
        p64 = np.concatenate(acc_p64)[:accept_target]
        z = 2.0 ** n * p64
        m = np.concatenate(margins)
        extra = {}
        if record_all:
            names = ('bits', 'u', 'p', 'p_fp64', 'acceptance', 'accepted', 'acceptance_fp64', 'accepted_fp64')
            extra['all'] = {k: np.concatenate([e[i] for e in every]) for i, k in enumerate(names)}
        return dict(**extra,accepted=len(bits), proposals=proposals, acceptance=len(bits) / proposals,
                    sample_seconds=gpu_s, wall_seconds=wall, accepted_per_s=len(bits) / gpu_s,
                    xeb=float(z.mean() - 1), xeb_se=float(z.std(ddof=1) / np.sqrt(len(z))),
                    decision_flips_fp64=flips, min_margin_fp64=float(m.min()),
                    decision_flips_beyond_tolerance=beyond, max_flip_margin_rel=flip_margin,
                    rel_err_c64_vs_c128={'median': float(np.median(np.abs(p - p64) / p64)),
                                         'max': float(np.max(np.abs(p - p64) / p64))},
                    bits=bits, p=p, p_fp64=p64)


def pack_words_device(x, n):
    """Device 0/1 tensor [count, n] (any integer type, nonzero = 1) -> int64
    [count, 2] words with qubit q at bit q of word q // 64 (bits_to_words)."""
    import torch
    b = (x != 0).to(torch.int64)
    words = torch.zeros((x.shape[0], 2), dtype=torch.int64, device=x.device)
    for w in range(2):
        cols = range(64 * w, min(n, 64 * w + 64))
        if len(cols):
            sh = torch.arange(len(cols), device=x.device, dtype=torch.int64)
            words[:, w] = (b[:, cols.start:cols.stop] << sh).sum(1)
    return words


def uniform_words_device(count, n, gen, dev):
    """count uniform n-bit proposals drawn directly as int64 [count, 2] words
    (qubit q at bit q of word q // 64): each word is one uniform integer
    below 2^bits, so every bit is an independent fair coin."""
    import torch
    words = torch.zeros((count, 2), dtype=torch.int64, device=dev)
    for w in range(2):
        bits = min(64, n - 64 * w)
        if bits <= 0:
            continue
        if bits <= 62:
            words[:, w].random_(0, 1 << bits, generator=gen)
        else:
            lo = torch.randint(0, 1 << 32, (count,), generator=gen, device=dev, dtype=torch.int64)
            hi = torch.randint(0, 1 << (bits - 32), (count,), generator=gen, device=dev, dtype=torch.int64)
            words[:, w] = lo | (hi << 32)
    return words


class DeviceCappedRejection:
    """Capped rejection with every proposal, uniform and decision on the
    device (seeded torch generator): uniform proposals (packed words drawn as
    uniform integers, and acceptance uniforms) or MADE proposals (the fused
    sampler).  The
    timed loop draws, scores (complex64-table scan) and decides; the
    complex128 audit of every decision runs afterwards, untimed."""

    def __init__(self, scorer, auditor, n, cap, proposal=None):
        self.scorer, self.auditor, self.n, self.cap = scorer, auditor, n, cap
        self.proposal = proposal

    def run(self, seed, accept_target, batch=65536, max_batches=10000, max_ahead=64):
        import torch
        n, cap = self.n, self.cap
        dev = self.scorer.device
        gen = torch.Generator(device=dev)
        gen.manual_seed(int(seed))
        net = self.proposal
        if net is not None:
            sb = net._sampling_buffers(batch)
        kept_words, kept_u, kept_logq, kept_keep = [], [], [], []
        accepted = 0
        proposals = 0
        batches = 0
        torch.cuda.synchronize(dev)
        t0 = time.perf_counter()
        # The host reads the accepted count after the first batch, then
        # enqueues as many batches as that rate says the target needs (at most
        # max_ahead) before reading again, so the device is not left idle
        # between batches.
        ahead = 1
        # each batch is a few launches: proposals, acceptance uniforms, the
        # scan, and one fused decision kernel (p, the capped acceptance
        # ratio, u < ratio and the kept count), so the host keeps ahead of
        # the device
        scale = 2.0 ** n / cap if net is None else cap
        count_dev = torch.zeros(1, dtype=torch.int64, device=dev)
        while batches < max_batches:
            count_dev.zero_()
            for _ in range(min(ahead, max_batches - batches)):
                if net is None:
                    words = uniform_words_device(batch, n, gen, dev)
                    logq = None
                else:
                    sb['u'][:batch].uniform_(generator=gen)
                    words, logq = net.sample_fused(batch)
                    words = words.clone()
                    logq = logq.clone()
                u = torch.rand(batch, generator=gen, device=dev, dtype=torch.float64)
                val, ex = self.scorer.scores_device(words)
                keep = torch.empty(batch, dtype=torch.uint8, device=dev)
                call('rgs_capped_decide', ptr(val), ptr(ex), ptr(u), None if logq is None else ptr(logq), batch, n,
                     scale, ptr(keep), ptr(count_dev), self.scorer.stream.cuda_stream)
                proposals += batch
                batches += 1
                kept_words.append(words)
                kept_u.append(u)
                kept_logq.append(logq)
                kept_keep.append(keep.view(torch.bool))
            accepted += int(count_dev.item())
            if accepted >= accept_target:
                break
            rate = max(accepted / proposals, 1e-9)
            ahead = min(max_ahead, max(1, math.ceil((accept_target - accepted) / (rate * batch))))
        torch.cuda.synchronize(dev)
        seconds = time.perf_counter() - t0
        # audit: complex128 tables for every proposal, decisions recomputed in FP64
        flips = beyond = 0
        flip_margin = 0.0
        margins = []
        z = []
        for words, u, logq, keep in zip(kept_words, kept_u, kept_logq, kept_keep):
            p64 = self.auditor.probabilities_device(words)
            if logq is None:
                acc64 = torch.clamp(p64 * (2.0 ** n / cap), max=1.0)
            else:
                acc64 = torch.clamp(p64 / (cap * torch.exp(logq)), max=1.0)
            keep64 = u < acc64
            flip = keep64 != keep
            rel = (u - acc64).abs() / acc64
            flips += int(flip.sum().item())
            beyond += int((flip & (rel > FLIP_TOLERANCE)).sum().item())
            if bool(flip.any()):
                flip_margin = max(flip_margin, rel[flip].max().item())
            margins.append((u - acc64).abs().min().item())
            z.append((p64[keep] * 2.0 ** n).cpu().numpy())
        z = np.concatenate(z)
        return dict(accepted=int(accepted), proposals=int(proposals), acceptance=accepted / proposals,
                    seconds=seconds, accepted_per_s=accepted / seconds, proposals_per_s=proposals / seconds,
                    xeb=float(z.mean() - 1), xeb_se=float(z.std(ddof=1) / np.sqrt(len(z))),
                    decision_flips_fp64=flips, min_margin_fp64=float(min(margins)),
                    decision_flips_beyond_tolerance=beyond, max_flip_margin_rel=flip_margin, z=z)
