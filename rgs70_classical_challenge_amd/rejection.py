"""Capped rejection sampling with exact p(x) as the judge.

Uniform proposals follow the reference's rule and RNG schedule
(work/rgs_attack/rejection_sampler.py:19-24): per proposal
``bits = rng.integers(0, 2, n)`` then ``u = rng.random()``, accept iff
``u < min(2^n p(x) / cap, 1)``; the distribution of accepted outputs is
proportional to min(p, cap 2^-n).  Learned proposals use the same rule with
q(x) in place of 2^-n: accept iff ``u < min(p(x) / (cap q(x)), 1)``, output
proportional to min(p, cap q).  Proposals are scored in batches of thousands
by the device scan (complex64 tables); every decision is then re-scored with
complex128 tables and its FP64 margin |u - acceptance| recorded
(audit_all_proposals_fp64.py:43-50).
"""
import time

import numpy as np


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

    def run(self, rng, accept_target, batch=8192, max_batches=1000, schedule='batched'):
        """Draw until `accept_target` outputs are accepted.  Returns a record
        with accepted bitstrings, p (complex64 scan), p_fp64 (complex128),
        decisions, margins and timing."""
        import torch
        n = self.n
        acc_bits, acc_p, acc_p64 = [], [], []
        flips = 0
        margins = []
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
            flips += int(np.sum(keep != keep64))
            margins.append(np.abs(u - acc64))
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
        return dict(accepted=len(bits), proposals=proposals, acceptance=len(bits) / proposals,
                    sample_seconds=gpu_s, wall_seconds=wall, accepted_per_s=len(bits) / gpu_s,
                    xeb=float(z.mean() - 1), xeb_se=float(z.std(ddof=1) / np.sqrt(len(z))),
                    decision_flips_fp64=flips, min_margin_fp64=float(m.min()),
                    rel_err_c64_vs_c128={'median': float(np.median(np.abs(p - p64) / p64)),
                                         'max': float(np.max(np.abs(p - p64) / p64))},
                    bits=bits, p=p, p_fp64=p64)
