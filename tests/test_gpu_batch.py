"""Uneven batched instances (one captured multi-stream graph per memory wave,
or one graph for the whole batch over a shared arena) against one launch
sequence per instance and against the oracle."""
import numpy as np
import pytest

from oracle import draw as od
from oracle import graph_state as gs

pytestmark = pytest.mark.gpu


def _torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip('no HIP device')
    return torch


def test_graph_replays_match_eager_and_oracle():
    torch = _torch()
    from rgs70_classical_challenge_amd.batch import BatchSampler, uneven_instances
    from rgs70_classical_challenge_amd.instances import instance
    specs = uneven_instances(count=10, seed=77, nmin=10, nmax=17)
    assert len({n for n, _ in specs}) > 3                       # really uneven
    print('specs', specs, flush=True)
    eager = BatchSampler(specs, samples=300, streams=3, seed=5).run_eager()
    bs = BatchSampler(specs, samples=300, streams=3, seed=5)
    first = bs.run_graph(replays=1, keep=True)
    assert set(first) == set(eager) == set(range(len(specs)))
    for i in eager:
        assert np.array_equal(first[i], eager[i]), i
    # replay the captured graphs again with eager work and synchronisation in between
    for rep in range(3):
        for wave, (graph, items) in bs.graphs.items():
            for it in items:
                it['idx'].fill_(-1)                             # eager work on the outputs
            torch.cuda.synchronize()
            graph.replay()
            torch.cuda.synchronize()
            for it in items:
                got = it['idx'].cpu().numpy()
                assert np.array_equal(got, eager[it['i']]), (rep, it['i'])
    # every instance against the oracle: state, pyramid-driven draws on the same uniforms
    for wave, (graph, items) in bs.graphs.items():
        for it in items:
            n, seed = specs[it['i']]
            g, u = instance(n, seed)
            ref = gs.full_state(g, u)
            s = it['s']
            assert np.linalg.norm(s.host() - ref) < 1e-5, it['i']
            raw = s.device_amplitudes()
            levels = od.pyramid(od.probabilities(raw))
            exp_idx, _ = od.tree_draws(levels, it['uni'].cpu().numpy())
            assert np.array_equal(eager[it['i']], exp_idx), it['i']


def test_single_graph_over_shared_arena_matches_eager():
    """The whole batch as one graph: waves share one arena (later waves reuse
    the memory of earlier ones) and every instance draws what its eager
    launch sequence draws, on every replay."""
    torch = _torch()
    from rgs70_classical_challenge_amd.batch import BatchSampler, footprint, uneven_instances
    specs = uneven_instances(count=12, seed=78, nmin=10, nmax=16)
    budget = max(footprint(n, 300) for n, _ in specs) * 3      # several waves
    eager = BatchSampler(specs, samples=300, streams=3, seed=6, budget_bytes=budget).run_eager()
    bs = BatchSampler(specs, samples=300, streams=3, seed=6, budget_bytes=budget)
    assert len(bs.waves) >= 2
    print('waves', [len(w) for w in bs.waves], flush=True)
    graph, waves = bs.build_single()
    assert bs.arena.high <= max(sum(footprint(specs[i][0], 300) for i in w) for w in bs.waves) + (64 << 20)
    for rep in range(3):
        for items in waves:
            for it in items:
                it['idx'].fill_(-1)                             # eager work between replays
        torch.cuda.synchronize()
        got = bs.run_single()
        assert set(got) == set(eager)
        for i in eager:
            assert np.array_equal(got[i], eager[i]), (rep, i)


def test_waves_respect_the_memory_budget():
    _torch()
    from rgs70_classical_challenge_amd.batch import BatchSampler, footprint, uneven_instances
    specs = uneven_instances(count=12, seed=3, nmin=12, nmax=18)
    budget = 3 * footprint(18, 100)
    bs = BatchSampler(specs, samples=100, budget_bytes=budget)
    assert sorted(i for w in bs.waves for i in w) == list(range(len(specs)))
    for w in bs.waves:
        assert len(w) == 1 or sum(footprint(specs[i][0], 100) for i in w) <= budget
