"""Learned proposal versus uniform proposal for capped rejection sampling.

1. Draw training bitstrings with the exact sampler (full state, device
   qubit-by-qubit draws) of a seeded n-vertex 3-regular instance.
2. Train a MADE (n -> 512 -> 512 -> n, BF16 weights, FP32 accumulation) with
   Adam on them, entirely with this build's kernels.
3. Run capped rejection with uniform proposals (the reference's rule,
   work/rgs_attack/rejection_sampler.py:19-24) and with the MADE as proposal,
   both judged by the exact device scan p(x) (complex64 tables, FP64 sums) and
   audited by the complex128 scan; report accepted samples/s, acceptance and
   FP64-audited linear XEB for each, plus the held-out log-likelihood gain of
   the network over the uniform distribution.
"""
import math
import time

import numpy as np


def index_to_bits(idx, n):
    idx = np.asarray(idx, np.int64)
    return ((idx[:, None] >> np.arange(n)) & 1).astype(np.uint8)


def run(n=34, seed=20261205, train_samples=200_000, steps=20_000, batch=512, hidden=512, cap=8.0,
        caps_learned=(1.0, 1.2, 1.4, 1.5, 1.6, 1.7, 1.8, 1.9, 2.0, 2.2, 2.5, 4.0, 8.0), accept_target=300_000,
        proposal_batch=65536, lr=1e-3, order_iterations=1000000, analysis_resamples=500, log=print):
    import torch
    from .engine import GraphStateSampler
    from .records import capped_statistics
    from .instances import instance
    from .made import MADE
    from .rejection import DeviceCappedRejection
    from .scan import ScanOracle, order_search

    out = {'n': n, 'seed': seed, 'cap': cap, 'train_samples': train_samples, 'steps': steps, 'batch': batch,
           'hidden': hidden}
    g, u = instance(n, seed)
    rng = np.random.Generator(np.random.PCG64(seed + 1))

    # 1. exact training and held-out samples from the full state
    t = time.perf_counter()
    s = GraphStateSampler(g, u, pipeline='slab')
    s.compute()
    held = 20_000
    idx, p = s.sample(rng.random(train_samples + held))
    out['exact_sampling_s'] = round(time.perf_counter() - t, 3)
    bits = index_to_bits(idx, n)
    train_bits, held_bits = bits[:train_samples], bits[train_samples:]
    held_p = p[train_samples:] / s.norm()
    del s
    torch.cuda.empty_cache()
    log(f'[learned] {train_samples + held} exact samples drawn')

    # 2. train
    net = MADE(n, hidden=hidden, batch=batch, seed=seed % 1000, lr=lr)
    torch.cuda.synchronize()
    t = time.perf_counter()
    losses = net.train(train_bits, steps, seed=seed, log_every=max(1, steps // 10))
    torch.cuda.synchronize()
    out['train_s'] = round(time.perf_counter() - t, 3)
    out['train_loss_trace'] = [round(x, 5) for x in losses]
    log(f'[learned] trained {steps} steps in {out["train_s"]} s')

    # held-out gain over uniform: E_p[log q] - log 2^-n, against the ceiling E_p[log(2^n p)]
    logq = net.log_q(held_bits)
    out['heldout_gain_nats'] = float(np.mean(logq) + n * math.log(2))
    out['heldout_ceiling_nats'] = float(np.mean(np.log(held_p * 2.0 ** n)))

    # 3. rejection, judged by the exact scan and audited in complex128
    from .autotune import scan_order
    order = scan_order(g, torch.cuda.get_device_properties(0)) or \
        order_search(g, iterations=order_iterations, rmax_cap=12)[0]
    scorer = ScanOracle(g, u, order)
    auditor = ScanOracle(g, u, order, fp64=True)
    out['scan_rmax'] = scorer.rmax
    out['scan_sum_states'] = scorer.plan.sum_states   # the annealed order's cost, sum of 2^r

    def summary(r):
        return {k: (round(v, 6) if isinstance(v, float) else v) for k, v in r.items()
                if k not in ('bits', 'p', 'p_fp64', 'z')}

    # warm both paths (graph capture, kernel attributes) outside the measurement
    DeviceCappedRejection(scorer, auditor, n, cap).run(seed + 1, 1, batch=proposal_batch)
    DeviceCappedRejection(scorer, auditor, n, cap, proposal=net).run(seed + 1, 1, batch=proposal_batch)
    uni = DeviceCappedRejection(scorer, auditor, n, cap).run(seed + 2, accept_target, batch=proposal_batch)
    out['uniform'] = summary(uni)
    log(f'[learned] uniform: {out["uniform"]["accepted_per_s"]:.1f} accepted/s, XEB {uni["xeb"]:.4f} '
        f'+- {uni["xeb_se"]:.4f}')
    out['learned'] = {}
    zs = {}

    def learned_at(c):
        r = DeviceCappedRejection(scorer, auditor, n, c, proposal=net).run(seed + 3, accept_target,
                                                                           batch=proposal_batch)
        out['learned'][str(c)] = summary(r)
        zs[float(c)] = r['z']
        log(f'[learned] cap {c}: {r["accepted_per_s"]:.1f} accepted/s, XEB {r["xeb"]:.4f} +- {r["xeb_se"]:.4f}')
        return r

    def matches(r):
        return abs(r['xeb'] - uni['xeb']) <= math.hypot(r['xeb_se'], uni['xeb_se'])

    for c in caps_learned:
        learned_at(c)
    # XEB rises with the cap: when no grid cap matches the uniform XEB within
    # one standard error, widen the grid until it brackets that XEB, then
    # interpolate between the bracketing caps
    for _ in range(8):
        if any(matches(r) for r in out['learned'].values()):
            break
        below = [(float(c), r['xeb']) for c, r in out['learned'].items() if r['xeb'] < uni['xeb']]
        above = [(float(c), r['xeb']) for c, r in out['learned'].items() if r['xeb'] > uni['xeb']]
        if not above:
            c = round(max(below)[0] * 1.5, 3)
        elif not below:
            c = round(max(1.0, min(above)[0] / 1.5), 3)
        else:
            (c0, x0), (c1, x1) = max(below), min(above)
            c = round(c0 + (uni['xeb'] - x0) * (c1 - c0) / (x1 - x0), 3)
            if not c0 < c < c1:
                break
        if str(c) in out['learned']:
            break
        learned_at(c)
    # learned caps whose XEB matches the uniform run within one standard error;
    # the fastest of them is the comparison
    match = [(r['accepted_per_s'], float(c)) for c, r in out['learned'].items() if matches(r)]
    if match:
        best = max(match)
        lr_ = out['learned'][str(best[1])]
        out['equal_xeb_comparison'] = {'learned_cap': best[1], 'learned_accepted_per_s': best[0],
                                       'learned_xeb': lr_['xeb'], 'learned_xeb_se': lr_['xeb_se'],
                                       'uniform_cap': cap, 'uniform_accepted_per_s': uni['accepted_per_s'],
                                       'uniform_xeb': uni['xeb'], 'uniform_xeb_se': uni['xeb_se'],
                                       'learned_acceptance': lr_['acceptance'],
                                       'uniform_acceptance': uni['acceptance'],
                                       'speedup': best[0] / uni['accepted_per_s'],
                                       'matching_caps': sorted(c for _, c in match)}
        # capped-run analysis of both accepted sets (analyze_challenge.py:11-15 semantics,
        # fewer bootstrap resamples for 3 x 10^5 strings)
        out['analysis'] = {'uniform': capped_statistics(uni['z'] - 1, resamples=analysis_resamples),
                           'learned': capped_statistics(zs[best[1]] - 1, resamples=analysis_resamples)}
    else:
        out['equal_xeb_comparison'] = None
    return out


def main(argv):
    """python -m rgs70_classical_challenge_amd.learned [n] [steps] [out.json]"""
    import json
    import sys
    n = int(argv[0]) if argv else 34
    steps = int(argv[1]) if len(argv) > 1 else 20_000
    res = run(n=n, steps=steps, log=lambda *a: print(*a, file=sys.stderr, flush=True))
    line = json.dumps(res)
    if len(argv) > 2:
        with open(argv[2], 'w') as f:
            f.write(line + '\n')
    print(line, flush=True)


if __name__ == '__main__':
    import sys
    main(sys.argv[1:])
