"""Pins of the scan, draw, RNG-schedule and statistics oracles (CPU only)."""
import json

import numpy as np
import pytest

from oracle import draw as od
from oracle import graph_state as gs
from oracle import scan as osc
from oracle.stats import xeb_stats
from rgs70_classical_challenge_amd.instances import instance, haar_unitaries


def random_graph(n, rng):
    g = np.triu(rng.integers(0, 2, (n, n), dtype=np.uint8), 1)
    return g | g.T


# ---------------------------------------------------------------- scan

def test_plan_profile_equals_dense_cut_rank():
    rng = np.random.default_rng(8)
    for n in (6, 12, 20):
        g = random_graph(n, rng)
        order = rng.permutation(n).tolist()
        steps, prof, _ = osc.compile_plan(g, order)
        assert prof == osc.cut_rank_profile(g, order)
        for st in steps:
            assert len(st.kernel) == st.r_in + 1 - st.r_out


@pytest.mark.parametrize('order_name,key', [('scan_order.npy', 'scan_order'),
                                            ('scan_order_optimized.npy', 'scan_order_optimized')])
def test_rgs70_plan_profile_matches_reference(golden_dir, order_name, key):
    d = np.load(f'{golden_dir}/rgs_graph.npz')
    order = np.load(f'{golden_dir}/{order_name}').tolist()
    ref = json.load(open(f'{golden_dir}/rgs70_profiles.json'))[key]
    _, prof, _ = osc.compile_plan(d['graph'], order)
    assert prof == ref['profile']
    assert osc.plan_cost(prof) == ref['sum_states']


def test_scan_amplitude_vs_statevector():
    """scan_oracle.selftest design (scan_oracle.py:166-180) at complex128."""
    rng = np.random.default_rng(20261030)
    for n in (8, 10, 12):
        g = random_graph(n, rng)
        u = haar_unitaries(n, rng)
        full = gs.full_state(g, u)
        steps, _, _ = osc.compile_plan(g, rng.permutation(n).tolist())
        for _ in range(5):
            bits = rng.integers(0, 2, n).tolist()
            amp = osc.amplitude(steps, u, bits)
            ref = full[sum(b << j for j, b in enumerate(bits))]
            assert abs(amp - ref) < 1e-13


def test_scan_amplitude_on_induced_fixture(golden_dir):
    d = np.load(f'{golden_dir}/induced_24q.npz')
    g, u = d['graph'], d['measurement_unitaries']
    full = gs.full_state(g, u)
    rng = np.random.default_rng(5)
    steps, prof, _ = osc.compile_plan(g, list(range(len(g))))
    for _ in range(4):
        x = int(rng.integers(0, 1 << len(g)))
        bits = [(x >> j) & 1 for j in range(len(g))]
        assert abs(osc.amplitude(steps, u, bits) - full[x]) < 1e-15 + 1e-12 * abs(full[x])


# ---------------------------------------------------------------- draw

def test_tree_draw_is_inverse_cdf():
    rng = np.random.default_rng(1)
    p = rng.random(1 << 10) ** 3
    p[rng.integers(0, len(p), 50)] = 0.0
    levels = od.pyramid(p)
    assert levels[-1][0] == pytest.approx(p.sum(), rel=1e-14)
    u = rng.random(2000)
    idx, margin = od.tree_draws(levels, u)
    ref = np.searchsorted(np.cumsum(p), u * p.sum(), side='right')
    ok = margin > 1e-14
    assert ok.mean() > 0.99
    assert np.array_equal(idx[ok], ref[ok])
    assert np.all(p[idx] > 0)


def test_reference_draw_factorisation(golden_dir):
    """Deferred draw: P(index) = |final[index]|^2 exactly (auxiliary marginal
    times the post-gate conditional, deferred_global_sampler.py:117-123)."""
    d = np.load(f'{golden_dir}/induced_12q.npz')
    g, u, A, B = d['graph'], d['measurement_unitaries'], d['A'].tolist(), d['B'].tolist()
    ob = [1, 0, 1, 1, 0, 0]
    raw, final = gs.deferred(g, u, A, B, ob)
    half = len(raw) // 2
    before = np.abs(raw.reshape(2, -1)) ** 2
    pair = np.abs(final.reshape(2, -1)) ** 2
    joint = before.sum(0)[None, :] * pair / pair.sum(0)[None, :]
    assert np.allclose(joint.reshape(-1), np.abs(final) ** 2, atol=1e-15)
    # one concrete draw: uniforms consumed exactly three times
    class Rec:
        def __init__(self, r):
            self.r, self.n = r, 0

        def random(self):
            self.n += 1
            return self.r.random()
    rec = Rec(np.random.default_rng(0))
    idx, p, norm = od.reference_draw(raw, rec, tilebits=3, terminal=u[A[-1]])
    assert rec.n == 3 and 0 <= idx < 2 * half and abs(norm - 1) < 1e-12
    assert p == pytest.approx(abs(final[idx]) ** 2, rel=1e-12)


def test_reference_draw_rejects_bad_norm():
    with pytest.raises(ArithmeticError):
        od.reference_draw(np.full(16, 0.5, complex), np.random.default_rng(0), tilebits=2)


# ---------------------------------------------------------------- RNG schedule

def _state(d):
    return {'bit_generator': 'PCG64', 'state': {'state': int(d['state']['state']), 'inc': int(d['state']['inc'])},
            'has_uint32': int(d['has_uint32']), 'uinteger': int(d['uinteger'])}


def test_production_rng_schedule_replay(golden_dir):
    """35 B integers then 3 uniforms per sample (born_production.py:44-47,
    born/verify.py:65-75), replayed from the recorded PCG64 states."""
    data = json.load(open(f'{golden_dir}/rgs70_born_records.json'))
    rng = np.random.Generator(np.random.PCG64(int(data['group']['cli_seed_decimal'])))
    assert rng.bit_generator.state == _state(data['group']['initial_pcg64_state_decimal128'])
    for rec in data['records']:
        assert rng.bit_generator.state == _state(rec['rngstate_before'])
        assert rng.integers(0, 2, 35).tolist() == rec['B_bits']
        assert [float(rng.random()) for _ in range(3)] == rec['draw_uniforms']
        assert rng.bit_generator.state == _state(rec['rngstate_after'])
        bits = [0] * 70
        for j, q in enumerate(data['A_vertices']):
            bits[q] = (rec['A_index'] >> j) & 1
        for j, q in enumerate(data['B_vertices']):
            bits[q] = rec['B_bits'][j]
        assert ''.join(map(str, bits)) == rec['bitstring_q0_left']


def test_capped_rejection_replay(golden_dir):
    """Uniform capped rejection: bits = integers(0,2,n), then one uniform;
    accept iff u < min(2^n p / cap, 1) (rejection_sampler.py:19-24,
    analyze_challenge.py:155-158)."""
    data = json.load(open(f'{golden_dir}/rgs70_cap1_proposals.json'))
    meta = data['metadata']
    rng = np.random.default_rng(meta['seed'])
    for row in data['proposals']:
        bits = ''.join(map(str, rng.integers(0, 2, meta['n'])))
        uacc = float(rng.random())
        acc = min(2.0 ** meta['n'] * row['probability_raw'] / meta['cap'], 1.0)
        assert bits == row['bitstring_q0_left']
        assert uacc == row['acceptance_uniform']
        assert abs(acc - row['acceptance_probability']) < 1e-12
        assert row['accepted'] == (uacc < acc)
        acc64 = min(2.0 ** meta['n'] * row['p_fp64'] / meta['cap'], 1.0)
        assert abs(abs(uacc - acc64) - row['margin_fp64']) < 1e-12


def test_xeb_statistics_match_recorded(golden_dir):
    z = np.array(json.load(open(f'{golden_dir}/rgs70_batch1_group0_z.json'))['z'])
    summ = json.load(open(f'{golden_dir}/rgs70_batch1_summary.json'))
    rec = summ['groups'][0]['score']
    st = xeb_stats(z, resamples=2000)
    assert st['xeb'] == pytest.approx(rec['mean'], rel=1e-12)
    assert st['standard_error'] == pytest.approx(rec['standard_error'], rel=1e-12)


def test_instances_are_seeded_and_valid():
    for n in (16, 17, 34, 70):
        g1, u1 = instance(n, 42)
        g2, u2 = instance(n, 42)
        assert np.array_equal(g1, g2) and np.array_equal(u1, u2)
        assert np.array_equal(g1, g1.T) and not np.any(np.diag(g1))
        if n % 2 == 0:
            assert np.all(g1.sum(1) == 3)
        assert np.allclose(u1 @ np.conj(np.transpose(u1, (0, 2, 1))), np.eye(2), atol=1e-14)


def test_split_draws_equal_full_descent():
    """Descending the block level and then each block's own pairwise pyramid
    picks the same index as descending the full pyramid."""
    rng = np.random.default_rng(44)
    p = rng.exponential(size=1 << 12) * 2.0 ** -12
    p[rng.integers(0, 1 << 12, 50)] = 0.0
    levels = od.pyramid(p)
    uni = rng.random(400)
    full, _ = od.tree_draws(levels, uni)
    for bb in (1, 4, 8):
        got = od.split_draws(levels[bb:], lambda blocks: p.reshape(-1, 1 << bb)[blocks], bb, uni)
        assert np.array_equal(got, full)


def test_batched_sweep_equals_single_sweep():
    from rgs70_classical_challenge_amd.instances import instance
    g, u = instance(14, 3)
    order = np.random.default_rng(1).permutation(14).tolist()
    steps, _, _ = osc.compile_plan(g, order)
    bits = np.random.default_rng(2).integers(0, 2, (40, 14))
    got = osc.amplitudes(steps, u, bits)
    one = np.array([osc.amplitude(steps, u, r.tolist()) for r in bits])
    full = gs.full_state(g, u)
    idx = (bits << np.arange(14)).sum(1)
    assert np.max(np.abs(got - one)) < 1e-15 * np.max(np.abs(one))
    assert np.max(np.abs(got - full[idx])) < 1e-14 * np.max(np.abs(full))
