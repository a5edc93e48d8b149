"""Device-side capped rejection against the exact output distribution: the
accepted strings follow min(p, cap q) (q uniform or the MADE's q), so their
linear XEB and the acceptance rate have closed forms at small n."""
import numpy as np
import pytest

from oracle import graph_state as gs

pytestmark = pytest.mark.gpu


def _torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip('no HIP device')
    return torch


def _expected(p, q, cap, n):
    w = np.minimum(p, cap * q)                       # accepted density (unnormalised)
    acc_rate = w.sum() / cap                         # E_q[min(p / (cap q), 1)]
    xeb = (w * p).sum() / w.sum() * 2 ** n - 1
    return acc_rate, xeb


def test_uniform_and_learned_rejection_match_closed_forms():
    _torch()
    from rgs70_classical_challenge_amd.instances import instance
    from rgs70_classical_challenge_amd.made import MADE
    from rgs70_classical_challenge_amd.rejection import DeviceCappedRejection
    from rgs70_classical_challenge_amd.scan import ScanOracle, order_search
    n = 12
    g, u = instance(n, 4242)
    psi = gs.full_state(g, u)
    p = np.abs(psi) ** 2
    order, _ = order_search(g, iterations=100000)
    scorer, auditor = ScanOracle(g, u, order), ScanOracle(g, u, order, fp64=True)
    # uniform proposals
    cap = 4.0
    r = DeviceCappedRejection(scorer, auditor, n, cap).run(7, 40000, batch=8192)
    acc, xeb = _expected(p, np.full(1 << n, 2.0 ** -n), cap, n)
    print('uniform', r['acceptance'], acc, r['xeb'], xeb, r['xeb_se'], flush=True)
    assert abs(r['acceptance'] - acc) < 5 * np.sqrt(acc * (1 - acc) / r['proposals'])
    assert abs(r['xeb'] - xeb) < 5 * r['xeb_se']
    assert r['decision_flips_fp64'] == 0 and r['decision_flips_beyond_tolerance'] == 0
    # learned proposals: a MADE trained briefly on exact samples
    rng = np.random.default_rng(1)
    data_idx = rng.choice(1 << n, size=20000, p=p / p.sum())
    data = ((data_idx[:, None] >> np.arange(n)) & 1).astype(np.uint8)
    net = MADE(n, hidden=64, batch=256, seed=4, lr=3e-3)
    net.train(data, 400, seed=2)
    all_bits = ((np.arange(1 << n)[:, None] >> np.arange(n)) & 1).astype(np.uint8)
    q = np.exp(net.log_q(all_bits))
    for cap in (1.0, 2.0):
        r = DeviceCappedRejection(scorer, auditor, n, cap, proposal=net).run(9, 40000, batch=8192)
        acc, xeb = _expected(p, q / q.sum(), cap, n)
        print('learned', cap, r['acceptance'], acc, r['xeb'], xeb, r['xeb_se'], flush=True)
        assert abs(r['acceptance'] - acc) < 5 * np.sqrt(acc * (1 - acc) / r['proposals']) + 2e-3
        assert abs(r['xeb'] - xeb) < 5 * r['xeb_se'] + 2e-3


@pytest.mark.parametrize('learned', [False, True])
def test_capped_decide_rounds_as_tensor_expressions(learned):
    """rgs_capped_decide against the same decision written as device tensor
    expressions (p from probabilities_from_scores, which the complex128 audit
    also uses): u is set exactly at the tensor-computed acceptance ratio for
    half the proposals (not kept) and one ulp below it for the other half
    (kept), so a ratio off by one rounding flips a decision.  The count is
    not a multiple of the wavefront or workgroup size."""
    torch = _torch()
    from rgs70_classical_challenge_amd.scan import probabilities_from_scores
    from rgs70_classical_challenge_amd._lib import call, ptr
    dev = torch.device('cuda', 0)
    n, cap, count = 34, 1.7 if learned else 8.0, 10007
    g = torch.Generator(device=dev)
    g.manual_seed(7 + learned)
    val = torch.randn((count, 2), generator=g, device=dev, dtype=torch.float64) * 0.6
    ex = torch.randint(-3, 4, (count,), generator=g, device=dev, dtype=torch.int32)
    p = probabilities_from_scores(val, ex, n)
    # the exact power of two, against an independent construction
    assert torch.equal(p, (val[:, 0] * val[:, 0] + val[:, 1] * val[:, 1])
                       * torch.tensor([2.0 ** k for k in (2 * ex.cpu().numpy() - n).tolist()], device=dev))
    if learned:
        logq = -n * np.log(2.0) + torch.randn(count, generator=g, device=dev, dtype=torch.float64)
        acc = torch.clamp(p / (cap * torch.exp(logq)), max=1.0)
        scale = cap
    else:
        logq = None
        acc = torch.clamp(p * (2.0 ** n / cap), max=1.0)
        scale = 2.0 ** n / cap
    below = torch.nextafter(acc, torch.zeros_like(acc))
    half = torch.arange(count, device=dev) % 2 == 1
    u = torch.where(half, below, acc).contiguous()
    keep = torch.empty(count, dtype=torch.uint8, device=dev)
    accepted = torch.zeros(1, dtype=torch.int64, device=dev)
    call('rgs_capped_decide', ptr(val), ptr(ex), ptr(u), None if logq is None else ptr(logq), count, n, scale,
         ptr(keep), ptr(accepted), torch.cuda.current_stream(dev).cuda_stream)
    torch.cuda.synchronize(dev)
    expect = (u < acc)
    assert torch.equal(keep.bool(), expect)
    assert torch.equal(expect, half & (acc > 0))
    assert int(accepted.item()) == int(expect.sum().item())


def test_capped_rejection_replays_reference_stream(golden_dir):
    """The reference's cap-1 run on its 70-qubit instance (seed 314159,
    rejection_sampler.py:18-24) replayed through CappedRejection with the
    reference's call schedule: complex64-table scorer and complex128-table
    auditor (rank-30 HBM scan).  Against the 64 recorded proposals
    (cap1_archive, cap1_full_audit_artifact): the same bitstrings and
    uniforms, p_fp64 to 1e-9, the complex64 p and the acceptance within 1e-5
    relative, and the same decision wherever the FP64 margin exceeds that
    tolerance (audit_all_proposals_fp64.py:43-50)."""
    import json
    _torch()
    from rgs70_classical_challenge_amd.rejection import CappedRejection, FLIP_TOLERANCE
    from rgs70_classical_challenge_amd.scan import ScanOracle
    d = np.load(f'{golden_dir}/rgs_graph.npz')
    g, u = d['graph'], d['measurement_unitaries']
    fixture = json.load(open(f'{golden_dir}/rgs70_cap1_proposals.json'))
    meta, rows = fixture['metadata'], fixture['proposals']
    order = np.load(f'{golden_dir}/scan_order_optimized.npy').tolist()
    scorer, auditor = ScanOracle(g, u, order), ScanOracle(g, u, order, fp64=True)
    cr = CappedRejection(scorer, auditor, meta['n'], meta['cap'])
    r = cr.run(np.random.default_rng(meta['seed']), 10 ** 9, batch=len(rows), max_batches=1, schedule='reference',
               record_all=True)
    a = r['all']
    assert len(a['u']) == len(rows)
    for k, row in enumerate(rows):
        assert ''.join(map(str, a['bits'][k])) == row['bitstring_q0_left'], k
        assert a['u'][k] == row['acceptance_uniform'], k
    p_fp64 = np.array([row['p_fp64'] for row in rows])
    p_rec = np.array([row['probability_raw'] for row in rows])
    acc_rec = np.array([row['acceptance_probability'] for row in rows])
    keep_rec = np.array([row['accepted'] for row in rows])
    margin = np.array([row['margin_fp64'] for row in rows])
    rel128 = np.abs(a['p_fp64'] - p_fp64) / p_fp64
    rel64 = np.abs(a['p'] - p_fp64) / p_fp64
    # acceptance against the FP64 acceptance of the recorded complex128 p (the
    # recorded acceptance itself came from the reference's complex64 p)
    acc_fp64 = np.minimum(2.0 ** meta['n'] * p_fp64 / meta['cap'], 1.0)
    relacc = np.abs(a['acceptance'] - acc_fp64) / acc_fp64
    print(f'complex128 p max rel {rel128.max():.2e}; complex64 p max rel {rel64.max():.2e} '
          f'(the reference\'s complex64: {np.max(np.abs(p_rec - p_fp64) / p_fp64):.2e}); '
          f'acceptance max rel {relacc.max():.2e}; accepted {int(a["accepted"].sum())} of {len(rows)}', flush=True)
    assert rel128.max() < 1e-9
    assert rel64.max() < 1e-5
    assert relacc.max() < 1e-5
    assert np.max(np.abs(acc_rec - acc_fp64) / acc_fp64) < 1e-4
    sure = margin > FLIP_TOLERANCE * acc_fp64
    assert np.array_equal(a['accepted'][sure], keep_rec[sure])
    assert np.array_equal(a['accepted_fp64'][sure], keep_rec[sure])
    assert r['decision_flips_beyond_tolerance'] == 0
