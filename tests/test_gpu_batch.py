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


FUSED_SPECS = [(n, 4000 + 10 * n + k) for n in range(16, 23) for k in range(2)]


def test_fused_batch_n16_22_bit_identical_to_per_instance_and_oracle():
    """The native multi-state batch over n = 16..22 (two instances per n, so
    every pass launch serves two states): draws, probabilities, amplitudes and
    pyramids bit-identical to each instance's own launch sequence on the same
    buffers and uniforms, amplitudes within 1e-5 of the oracle's state, draws
    equal to the oracle's tree descent on the device amplitudes, and graph
    replays of the batch equal to its eager run."""
    torch = _torch()
    from rgs70_classical_challenge_amd.batch import BatchSampler
    from rgs70_classical_challenge_amd.instances import instance
    bs = BatchSampler(FUSED_SPECS, samples=500, streams=2, seed=11)
    assert len(bs.waves) == 1
    fused = bs.run_fused()
    launches, groups = bs.fused_launches()
    per_instance = sum(it['s'].passes + 2 for _, items in bs.fused for it in items)
    print(f'fused: {launches} launches for {len(FUSED_SPECS)} instances in {groups} groups '
          f'(per-instance sequences: at least {per_instance})', flush=True)
    assert groups <= 7 and launches < per_instance
    _, items = bs.fused[0]
    snap = {}
    for it in items:
        sl = it['s'].slab
        snap[it['i']] = (sl.state.cpu().numpy().copy(), sl.pyramid.cpu().numpy().copy(), it['p'].cpu().numpy().copy())
    # each instance's own launch sequence on the same buffers
    cur = torch.cuda.current_stream()
    for it in items:
        it['idx'].fill_(-1)
        it['p'].fill_(-1)
        bs._launch(it, cur)
    torch.cuda.synchronize()
    for it in items:
        i = it['i']
        sl = it['s'].slab
        st, pyr, p = snap[i]
        st2 = sl.state.cpu().numpy()
        bad = np.nonzero(st2 != st)[0]
        print(f'  instance {i} n={FUSED_SPECS[i][0]}: {len(bad)} state entries differ'
              + (f' (first {bad[:4]}, max |diff| {np.max(np.abs(st2[bad] - st[bad])):.3e}, '
                 f'max |amp| {np.max(np.abs(st)):.3e})' if len(bad) else '')
              + f', {np.sum(it["p"].cpu().numpy() != p)} p differ', flush=True)
    for it in items:
        i = it['i']
        sl = it['s'].slab
        st, pyr, p = snap[i]
        assert np.array_equal(it['idx'].cpu().numpy(), fused[i]), i
        assert np.array_equal(it['p'].cpu().numpy(), p), i
        assert np.array_equal(sl.state.cpu().numpy(), st), i
        assert np.array_equal(sl.pyramid.cpu().numpy(), pyr), i
    # the oracle: state, pyramid and draws
    for it in items:
        i = it['i']
        n, seed = FUSED_SPECS[i]
        print(f'  oracle n={n} seed={seed}', flush=True)
        g, u = instance(n, seed)
        ref = gs.full_state(g, u)
        s = it['s']
        assert np.max(np.abs(s.host() - ref)) <= 1e-5 * np.max(np.abs(ref)), i
        raw = s.device_amplitudes()
        levels = od.pyramid(od.probabilities(raw))
        exp_idx, _ = od.tree_draws(levels, it['uni'].cpu().numpy())
        assert np.array_equal(fused[i], exp_idx), i
        assert np.array_equal(snap[i][2], levels[0][exp_idx]), i
    # the batch captured as one graph: every replay equals the eager run
    graph = bs.capture_fused()
    for rep in range(3):
        for it in items:
            it['idx'].fill_(-1)
            it['s'].slab.state.zero_()
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        for it in items:
            assert np.array_equal(it['idx'].cpu().numpy(), fused[it['i']]), (rep, it['i'])
    del graph
    bs.close_fused()


def test_fused_batch_over_waves_matches_eager():
    """Several memory waves, each a native batch over one shared arena, in
    one run: every instance draws what its eager launch sequence draws."""
    _torch()
    from rgs70_classical_challenge_amd.batch import BatchSampler, footprint, uneven_instances
    specs = uneven_instances(count=20, seed=79, nmin=12, nmax=22)
    budget = max(footprint(n, 200) for n, _ in specs) * 3
    eager = BatchSampler(specs, samples=200, seed=12, budget_bytes=budget).run_eager()
    bs = BatchSampler(specs, samples=200, seed=12, budget_bytes=budget)
    assert len(bs.waves) >= 2
    print('waves', [len(w) for w in bs.waves], flush=True)
    for rep in range(2):
        got = bs.run_fused()
        assert set(got) == set(eager)
        for i in eager:
            assert np.array_equal(got[i], eager[i]), (rep, i)
    bs.close_fused()
