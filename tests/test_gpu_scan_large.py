"""HBM-table scan (one proposal at a time, ranks up to 32): against the
one-launch batch scan and the state vector at small rank, and against the
reference's recorded complex128 amplitudes of the 70-qubit RGS instance
(rank 30 orderings)."""
import json

import numpy as np
import pytest

from oracle import graph_state as gs

pytestmark = pytest.mark.gpu


def _torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip('no HIP device')
    return torch


def _bits(s):
    return np.array([int(c) for c in s], np.uint8)        # qubit 0 first


@pytest.mark.parametrize('n', [10, 14])
def test_large_path_matches_batch_and_statevector(n):
    _torch()
    from rgs70_classical_challenge_amd.instances import instance
    from rgs70_classical_challenge_amd.scan import ScanOracle
    g, u = instance(n, 900 + n)
    ref = gs.full_state(g, u)
    rng = np.random.default_rng(n)
    order = rng.permutation(n).tolist()
    bits = rng.integers(0, 2, (40, n), dtype=np.uint8)
    idx = (bits.astype(np.int64) << np.arange(n)).sum(1)
    want = ref[idx]
    for fp64, tol in ((False, 1e-5), (True, 1e-12)):
        large = ScanOracle(g, u, order, fp64=fp64, batch_rmax=0)
        batch = ScanOracle(g, u, order, fp64=fp64)
        print(f'n={n} fp64={fp64} rmax={large.rmax}', flush=True)
        a_large = large.amplitudes(bits)
        a_batch = batch.amplitudes(bits)
        assert np.max(np.abs(a_large - want)) < tol * np.max(np.abs(want)) * 10
        p_large, p_want = np.abs(a_large) ** 2, np.abs(want) ** 2
        big = p_want > 1e-3 / 2 ** n
        assert np.max(np.abs(p_large[big] - p_want[big]) / p_want[big]) < (1e-5 if not fp64 else 1e-10)
        if fp64:
            assert np.allclose(a_large, a_batch, rtol=1e-12, atol=1e-15 * np.max(np.abs(want)))


def test_rgs70_recorded_amplitudes(golden_dir):
    """The reference's complex128 rescoring of its own 70-qubit outputs
    (born/batch1/born_group0_fp64.jsonl rows kept in the fixture) and of its
    capped-rejection proposals (cap1_full_audit_artifact)."""
    _torch()
    from rgs70_classical_challenge_amd.scan import ScanOracle
    d = np.load(f'{golden_dir}/rgs_graph.npz')
    g, u = d['graph'], d['measurement_unitaries']
    order = np.load(f'{golden_dir}/scan_order_optimized.npy').tolist()
    born = json.load(open(f'{golden_dir}/rgs70_born_records.json'))['records']
    cap1 = json.load(open(f'{golden_dir}/rgs70_cap1_proposals.json'))['proposals'][:24]
    so = ScanOracle(g, u, order, fp64=True)
    print(f'rgs70 rmax={so.rmax} sum_states={so.plan.sum_states}', flush=True)
    assert so.rmax == 30
    bits = np.stack([_bits(r['bitstring_q0_left']) for r in born])
    amp = so.amplitudes(bits)
    want = np.array([complex(r['re'], r['im']) for r in born])
    rel = np.abs(amp - want) / np.abs(want)
    print('born records: max relative amplitude error', rel.max(), flush=True)
    assert rel.max() < 1e-9
    bits = np.stack([_bits(r['bitstring_q0_left']) for r in cap1])
    p = so.probabilities(bits)
    want_p = np.array([r['p_fp64'] for r in cap1])
    relp = np.abs(p - want_p) / want_p
    print('cap1 proposals: max relative p error', relp.max(), flush=True)
    assert relp.max() < 1e-9
    # complex64 tables with FP64 sums on the same strings: the reference's own
    # complex64 scan agreed with its complex128 rescoring to 3.5e-5 at most
    p64 = ScanOracle(g, u, order).probabilities(bits)
    rel64 = np.abs(p64 - want_p) / want_p
    print('complex64 tables: median / max relative p error', np.median(rel64), rel64.max(), flush=True)
    assert rel64.max() < 1e-5
