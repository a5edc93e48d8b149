"""Explicit strided groups: each group with its own width, row width (lane
bits) and LDS rounds, against the complex128 oracle.

The plans mix the tile shapes the large-state plans use - 32-, 64- and
128-amplitude rows, 13- and 14-bit strided tiles, with and without an LDS
round - in one schedule, at sizes the oracle holds in full.  Bars as in the
other slab tests: L2 distance below 1e-5 of the normalised state, p(x) within
1e-5 relative wherever p >= 0.1 / 2^n, the FP64 pyramid and the draws bit for
bit against the oracle run on the device's own amplitudes, and the sign
stages (init + Walsh + graph phase) bit-exact.
"""
import numpy as np
import pytest

from oracle import draw as od
from oracle import graph_state as gs
from rgs70_classical_challenge_amd.instances import instance

from test_gpu_exact import _exact, _run
from test_gpu_slab import check_pyramid_and_draws, check_state

pytestmark = pytest.mark.gpu

torch = pytest.importorskip('torch')

# n = 22: 8 contiguous bits, then two 7-bit groups
PLANS_22 = [
    dict(low_bits=8, groups=[(7, 6, 0), (7, 7, 1)]),
    dict(low_bits=8, groups=[(7, 7, 1), (7, 6, 0)], group_order=1),
    dict(low_bits=8, groups=[(7, 5, 0), (7, 7, 0)]),
    dict(low_bits=8, groups=[(7, 7, 2), (7, 6, 1)], dphase_global=1),
    dict(low_bits=7, groups=[(8, 6, 1), (7, 7, 1)]),
]


@pytest.mark.parametrize('pipeline', ['slab', 'literal', 'folded'])
@pytest.mark.parametrize('plan', PLANS_22)
def test_explicit_groups_graph_state(pipeline, plan):
    from rgs70_classical_challenge_amd.engine import GraphStateSampler
    n = 22
    g, u = instance(n, 2200)
    ref = gs.full_state(g, u)
    s = GraphStateSampler(g, u, pipeline=pipeline, tuning=plan)
    print(f'case pipeline={pipeline} plan={plan} -> {s.plan_text}', flush=True)
    s.compute()
    check_state(s.host(), ref)
    assert abs(s.norm() - 1) < 1e-5
    check_pyramid_and_draws(s, ref, np.random.default_rng(5).random(500))


@pytest.mark.parametrize('plan', PLANS_22[:2])
def test_explicit_groups_sign_stages_bit_exact(plan):
    from rgs70_classical_challenge_amd import _lib
    from rgs70_classical_challenge_amd.engine import _Slab, form_rows, graph_phase_lut
    n = 22
    g, _ = instance(n, 2201)
    sl = _Slab(n, 0, tuning=plan)
    lut = sl.to_device(graph_phase_lut(g))
    rows = form_rows(g)
    sign = np.ones(1 << n)
    gs.quadratic_phase(sign, g)
    print(f'case n={n} plan={plan}', flush=True)
    want = sign * 2.0 ** -n
    gs.walsh(want)
    _exact(_run(sl, sl.desc(_lib.INIT_GRAPH, 2.0 ** -n, init_lut=lut, init_rows=rows, walsh=True, mass=False)), want)
    _exact(_run(sl, sl.desc(_lib.INIT_DELTA, 1.0, walsh=True, phase_lut=lut, phase_rows=rows, mass=False)), sign)


# m = 20 (the induced 40-qubit instance): 7 or 8 contiguous bits, then two groups
PLANS_C20 = [
    dict(low_bits=7, groups=[(6, 7, 1), (7, 6, 0)]),
    dict(low_bits=8, groups=[(5, 6, 1), (7, 7, 1)], dphase_global=1),
    dict(low_bits=8, groups=[(5, 5, 0), (7, 7, 1)], group_order=1),
]


@pytest.mark.parametrize('plan', PLANS_C20)
def test_explicit_groups_conditional(golden_dir, plan):
    from rgs70_classical_challenge_amd.engine import ConditionalSampler
    d = np.load(f'{golden_dir}/induced_40q.npz')
    g, u, A, B = d['graph'], d['measurement_unitaries'], d['A'].tolist(), d['B'].tolist()
    rng = np.random.default_rng(4040)
    eng = ConditionalSampler(g, u, A, B, [0], tuning=plan)
    ob = rng.integers(0, 2, len(B)).tolist()
    print(f'case plan={plan} -> {eng.plan_text}', flush=True)
    eng.compute(ob)
    raw, final = gs.deferred(g, u, A, B, ob)
    assert np.linalg.norm(eng.host() - final) < 1e-5
    mass, norm, tb, nmass = eng.masses()
    assert abs(norm - 1) < 1e-5
    seed = int(rng.integers(1 << 30))
    i1, p1, _ = eng.draw(np.random.default_rng(seed))
    i2, p2, _ = od.reference_draw(raw, np.random.default_rng(seed), terminal=u[A[-1]])
    assert i1 == i2
    floor = 0.1 * 2.0 ** -len(A)
    assert abs(p1 - p2) <= 1e-5 * max(p2, floor), (p1, p2)
