"""Deferred leaves: the last contiguous pass writes the masses only, and draws
and reads recompute the tiles they touch.  Everything a sampler returns must
be the same bits as with the pass stored: the pyramid, the draws and their p,
probability_at, the recomputed tiles, and the state after rgs_slab_finish.
Run on the plain sampler over the pipelines and tile shapes of the GPU slab
tests (with the rotations' outcome flips), and on the conditional sampler
with its deferred terminal gate."""
import gc

import numpy as np
import pytest

from oracle import draw as od
from rgs70_classical_challenge_amd.instances import instance

pytestmark = pytest.mark.gpu

torch = pytest.importorskip('torch')

TUNINGS = [
    None,
    dict(low_bits=13, tile_bits=13, lane_bits=6, round_bits=0, group_order=1),
    dict(low_bits=8, tile_bits=11, lane_bits=3, round_bits=1),
    dict(low_bits=12, groups=[(4, 6, 0), (6, 6, 1)]),
]


def _engine():
    from rgs70_classical_challenge_amd import engine
    return engine


@pytest.mark.parametrize('tuning', TUNINGS)
@pytest.mark.parametrize('pipeline', ['slab', 'literal', 'folded'])
def test_deferred_leaves_equal_stored_pass(pipeline, tuning):
    E = _engine()
    n = 22
    g, u = instance(n, 2222)
    ref = E.GraphStateSampler(g, u, pipeline=pipeline, tuning=tuning)
    dfr = E.GraphStateSampler(g, u, pipeline=pipeline, tuning=tuning, defer_leaves=True)
    print(f'case n={n} pipeline={pipeline} tuning={tuning} plan={dfr.plan_text} flip={ref.flip:#x}', flush=True)
    last = dfr.plan_text.split()[-1]
    assert last.startswith('low[') and '[gates][mass][leaf]:' in last, dfr.plan_text
    assert ref.flip != 0
    ref.compute()
    dfr.compute()
    assert np.array_equal(dfr.slab.pyramid.cpu().numpy(), ref.slab.pyramid.cpu().numpy())
    rng = np.random.default_rng(22)
    uni = rng.random(4096)
    i_ref, p_ref = ref.sample(uni)
    i_dfr, p_dfr = dfr.sample(uni)
    assert np.array_equal(i_dfr, i_ref) and np.array_equal(p_dfr, p_ref)
    probe = np.concatenate([i_ref[:64], rng.integers(0, 1 << n, 64)])
    assert [dfr.probability_at(int(x)) for x in probe] == [ref.probability_at(int(x)) for x in probe]
    # recomputed tiles, including repeats and an index past the last tile
    # (its slot is left as it was)
    sl = dfr.slab
    lb = sl.leaf_bits(dfr._desc)
    tiles = np.concatenate([rng.integers(0, 1 << (n - lb), 16), [0, (1 << (n - lb)) - 1, 3, 3]])
    got = sl.leaf_tiles(dfr._desc, tiles).cpu().numpy()
    want = ref.slab.state.cpu().numpy().reshape(-1, 1 << lb)[tiles]
    assert np.array_equal(got, want)
    t = torch.as_tensor(np.array([1 << (n - lb)], np.int64)).to(sl.device)
    out = torch.full((1, 1 << lb), 7.0, dtype=torch.complex64, device=sl.device)
    from rgs70_classical_challenge_amd import _lib
    from rgs70_classical_challenge_amd.engine import ctypes_ref
    _lib.call('rgs_slab_leaf_tiles', ctypes_ref(dfr._desc), _lib.ptr(t), 1, _lib.ptr(out), sl.stream.cuda_stream)
    assert bool((out == 7.0).all())
    # the finished state and the host view
    assert np.array_equal(dfr.device_amplitudes(), ref.device_amplitudes())
    assert not dfr._pending
    i2, p2 = dfr.sample(uni)                   # the stored-state draw path now
    assert np.array_equal(i2, i_ref) and np.array_equal(p2, p_ref)
    # a second compute defers again
    dfr.compute()
    assert dfr._pending
    i3, _ = dfr.sample(uni[:256])
    assert np.array_equal(i3, i_ref[:256])
    # draws of the deferred state against the oracle's tree draw on the
    # finished values (the same complex64 numbers)
    levels = od.pyramid(od.probabilities(ref.device_amplitudes()))
    exp_idx, _ = od.tree_draws(levels, uni[:512])
    assert np.array_equal(i_dfr[:512], exp_idx)
    del ref, dfr
    gc.collect()
    torch.cuda.empty_cache()


@pytest.mark.parametrize('tuning', [None, dict(low_bits=8, tile_bits=11, lane_bits=3)])
@pytest.mark.parametrize('name', ['induced_24q.npz', 'induced_40q.npz'])
def test_conditional_deferred_leaves(golden_dir, name, tuning):
    E = _engine()
    d = np.load(f'{golden_dir}/{name}')
    g, u, A, B = d['graph'], d['measurement_unitaries'], d['A'].tolist(), d['B'].tolist()
    ref = E.ConditionalSampler(g, u, A, B, [0], tuning=tuning, defer_leaves=False)
    dfr = E.ConditionalSampler(g, u, A, B, [0], tuning=tuning)
    # a one-pass schedule makes the state itself: nothing to defer
    assert dfr.defer_leaves and ('[leaf]' in dfr.plan_text) == (dfr.passes > 1), dfr.plan_text
    rng = np.random.default_rng(2424)
    for trial in range(2):
        ob = rng.integers(0, 2, len(B)).tolist()
        print(f'case {name} trial={trial} tuning={tuning} plan={dfr.plan_text}', flush=True)
        ref.compute(ob)
        dfr.compute(ob)
        seed = int(rng.integers(1 << 30))
        a = ref.draw(np.random.default_rng(seed))
        b = dfr.draw(np.random.default_rng(seed))
        assert a == b and ref.last_margins == dfr.last_margins
        uni, bits = rng.random(2048), rng.random(2048)
        ia, pa = ref.sample(uni, bits)
        ib, pb = dfr.sample(uni, bits)
        assert np.array_equal(ia, ib) and np.array_equal(pa, pb)
        for x in ia[:16]:
            assert dfr.probability_at(int(x)) == ref.probability_at(int(x))
        for loc in rng.integers(0, ref.n, 8):
            assert np.array_equal(dfr.transformed_pair(int(loc)), ref.transformed_pair(int(loc)))
        assert np.array_equal(dfr.host(), ref.host())
    del ref, dfr
    gc.collect()
    torch.cuda.empty_cache()
