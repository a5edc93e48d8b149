"""The n = 34 plan's kernels against the oracle, at a size the oracle holds in
full and at the full n = 34 size.

The bench's plan runs two tile shapes: 13-bit contiguous tiles (k_tile<13, 0, 0>
for the init pass, k_tile_mass<13, 0, 25> for the gates + masses pass) and
13-bit strided tiles with 64-amplitude rows and no LDS rounds (k_tile<13, 6,
0>), with the strided groups walked from the high bits down.
At n = 27 the same tuning gives the same pass kinds (two 7-bit strided groups:
Walsh-only, Walsh + gates, gates-only), and the whole state fits the oracle.
At n = 34 no oracle holds 2^34 amplitudes, so the checks are the FP64 norm,
p(x) at drawn and uniform indices against the complex128 rank-scan oracle,
the pyramid's pairwise sums, every draw replayed on the host from the device
pyramid, and a size-independent statistic of the draws.

Bars: p(x) within 1e-5 relative of the complex128 oracle wherever
p >= 0.1 / 2^n (complex64 rounding is absolute in the amplitude, so below
that the bar is absolute: 1e-5 * 0.1 / 2^n); pyramid and draws bit-exact
against the oracle run on the device's own complex64 amplitudes.
"""
import gc

import numpy as np
import pytest

from oracle import draw as od
from oracle import scan as osc
from rgs70_classical_challenge_amd.instances import instance

pytestmark = pytest.mark.gpu

torch = pytest.importorskip('torch')

# the seven-pass schedule's n = 34 tuning (its 14-bit strided tiles narrow to 13 bits)
HEADLINE = dict(low_bits=13, tile_bits=14, lane_bits=6, round_bits=0, group_order=1)
# the autotuned n = 34 entry the bench runs: the fused schedule with strided
# groups of 8, 8 and 5 bits (32-amplitude rows), walked from the top:
# k_tile<10, 5, 0> makes the state, k_tile_fixed<13, 5, 0, 143> runs the 8-bit
# groups and k_tile_mass<13, 0, 159> the contiguous pass
BENCH_PLAN = dict(low_bits=13, tile_bits=13, lane_bits=6, round_bits=0, group_order=1, fuse_walsh=1,
                  groups=[[8, 5, 0], [8, 5, 0], [5, 5, 0]])
# the same kernels at n = 26: one 8-bit and one 5-bit group
BENCH_PLAN_26 = dict(BENCH_PLAN, groups=[[8, 5, 0], [5, 5, 0]])
BENCH_KERNELS = {'tb10/c5/rb0', 'tb13/c5/rb0', 'tb13/c0/rb0'}
BENCH_SEED = 20261203


def _kernels(plan_text):
    return {tok.split(':')[1].split('/', 1)[1] for tok in plan_text.split()}


def _d1_at(gates, idx):
    """prod_{q < gate_limit} D1_q[x_q] at indices idx."""
    f = np.ones(len(idx), np.complex128)
    for q in range(gates.gate_limit):
        f *= gates.d1[q][(idx >> q) & 1]
    return f


def _oracle_amplitudes(g, u, idx, n):
    from rgs70_classical_challenge_amd.scan import order_search
    order, _ = order_search(g, iterations=200000, seed=3)
    steps, _, _ = osc.compile_plan(g, order)
    bits = ((idx[:, None] >> np.arange(n)) & 1).astype(np.int64)
    out = np.empty(len(idx), np.complex128)
    for k in range(0, len(idx), 1024):
        out[k:k + 1024] = osc.amplitudes(steps, u, bits[k:k + 1024])
    return out


def _check_p(p_dev, p_ref, n):
    floor = 0.1 * 2.0 ** -n
    err = np.abs(p_dev - p_ref)
    bar = 1e-5 * np.maximum(p_ref, floor)
    big = p_ref >= floor
    rel = err[big] / p_ref[big]
    print(f'  p rel error (p >= 0.1/2^n, {big.sum()} of {len(p_ref)}): median {np.median(rel):.2e} '
          f'max {rel.max():.2e}', flush=True)
    assert np.all(err <= bar), (err / bar).max()


def _sampler(n, seed, pipeline, tuning):
    from rgs70_classical_challenge_amd.engine import GraphStateSampler
    g, u = instance(n, seed)
    s = GraphStateSampler(g, u, pipeline=pipeline, tuning=tuning)
    return g, u, s


SEVEN = {'tb13/c0/rb0', 'tb13/c6/rb0'}


@pytest.mark.parametrize('pipeline,tuning,n,passes,kernels', [
    ('slab', HEADLINE, 27, 5, SEVEN), ('slab', dict(HEADLINE, fuse_walsh=1), 27, 3, SEVEN),
    ('literal', HEADLINE, 27, 5, SEVEN), ('folded', HEADLINE, 27, 3, SEVEN),
    ('slab', BENCH_PLAN_26, 26, 3, BENCH_KERNELS)], ids=['slab', 'slab-fused', 'literal', 'folded', 'slab-bench-plan'])
def test_headline_kernels_at_n27(pipeline, tuning, n, passes, kernels):
    g, u, s = _sampler(n, 2700 + n, pipeline, tuning)
    print(f'case n={n} pipeline={pipeline} tuning={tuning} plan={s.plan_text}', flush=True)
    assert s.passes == passes and _kernels(s.plan_text) == kernels, s.plan_text
    s.compute()
    root = s.slab.root()
    assert abs(root - 1) < 1e-6, root
    raw = s.device_amplitudes()
    # pyramid bit-exact against the oracle's pairwise sums of the same values
    levels = od.pyramid(od.probabilities(raw))
    sl = s.slab
    dev = sl.pyramid.cpu().numpy()
    off = 0
    for lv in levels[sl.mass_bits:]:
        assert np.array_equal(dev[off:off + len(lv)], lv)
        off += len(lv)
    assert off == len(dev)
    rng = np.random.default_rng(27)
    uni = rng.random(1000)
    idx, p = s.sample(uni)
    exp_idx, _ = od.tree_draws(levels, uni)
    assert np.array_equal(idx, exp_idx)
    assert np.array_equal(p, levels[0][idx])
    # amplitudes and p at drawn and uniform indices against the complex128 scan
    probe = np.concatenate([idx, rng.integers(0, 1 << n, 3000)]).astype(np.int64)
    ref = _oracle_amplitudes(g, u, probe, n)
    got = raw[probe] * _d1_at(s._gates, probe ^ s.flip)     # D1 acts on the state index
    amp_err = np.abs(got - ref)
    assert np.all(amp_err <= 1e-5 * np.maximum(np.abs(ref), np.sqrt(0.1 * 2.0 ** -n))), amp_err.max()
    _check_p(od.probabilities(raw[probe]), np.abs(ref) ** 2, n)
    del s, raw, levels
    gc.collect()
    torch.cuda.empty_cache()


@pytest.mark.parametrize('fuse', [0, 1], ids=['seven-pass', 'bench-plan'])
def test_headline_n34_full_size(fuse):
    """The bench's instance and plans at full size: 2^34 complex64 amplitudes
    (the seven-pass plan, and the four-pass plan of fuse_walsh)."""
    n = 34
    gc.collect()
    torch.cuda.empty_cache()
    free, _ = torch.cuda.mem_get_info()
    need = (8 << n) + (16 << (n - 8)) + (4 << 30)
    if free < need:
        pytest.skip(f'the 2^34 state needs {need} bytes of free device memory; {free} free')
    g, u, s = _sampler(n, BENCH_SEED, 'slab', BENCH_PLAN if fuse else HEADLINE)
    print(f'case n=34 bench instance fuse_walsh={fuse} plan={s.plan_text}', flush=True)
    assert (s.passes, _kernels(s.plan_text)) == ((4, BENCH_KERNELS) if fuse else (7, SEVEN)), s.plan_text
    s.compute()
    sl = s.slab
    root = sl.root()
    print(f'  FP64 root {root!r}', flush=True)
    assert abs(root - 1) < 1e-6
    # pyramid: every level above the block masses is the pairwise sum of the one below
    pyr = sl.pyramid.cpu().numpy()
    B = sl.base_bits
    levels, off = [], 0
    for k in range(B + 1):
        levels.append(pyr[off:off + (1 << (B - k))])
        off += 1 << (B - k)
    for k in range(1, B + 1):
        assert np.array_equal(levels[k], levels[k - 1][0::2] + levels[k - 1][1::2]), k
    mb = sl.mass_bits
    state = sl.state
    flip = s.flip                     # outcome x is held at state index x ^ flip

    def fetch(blocks):
        b = torch.as_tensor(np.asarray(blocks, np.int64)).to(sl.device)
        ix = ((b[:, None] << mb) + torch.arange(1 << mb, device=sl.device)[None, :]) ^ flip
        vals = state[ix.reshape(-1)].cpu().numpy()
        return od.probabilities(vals).reshape(len(blocks), 1 << mb)

    rng = np.random.default_rng(3434)
    # block masses: the pairwise tree of their 256 leaves, bit for bit
    blocks = rng.integers(0, 1 << B, 64)
    leaves = fetch(blocks)
    for b, lv in zip(blocks, leaves):
        assert od.pyramid(lv)[-1][0] == levels[0][b]
    # 10,000 device draws replayed on the host from the device pyramid
    uni = rng.random(10000)
    idx, p = s.sample(uni)
    assert np.all((idx >= 0) & (idx < (1 << n)))
    host = od.split_draws(levels, fetch, mb, uni)
    assert np.array_equal(idx, host)
    assert np.array_equal(p, s.slab.gather_prob(idx ^ flip))
    # p(x) at drawn and uniform indices against the complex128 rank scan
    probe = np.concatenate([idx[:512], rng.integers(0, 1 << n, 512)]).astype(np.int64)
    ref = _oracle_amplitudes(g, u, probe, n)
    _check_p(s.slab.gather_prob(probe ^ flip), np.abs(ref) ** 2, n)
    # size-independent draw statistic: E_draw[2^n p(x)] = 2^n sum p^2 =
    # E_uniform[(2^n p)^2], estimated from the draws and from 2^20 uniform
    # indices independently; they must agree within 5 standard errors
    z_draw = (2.0 ** n) * p
    z_unif = (2.0 ** n) * s.slab.gather_prob(rng.integers(0, 1 << n, 1 << 20))
    a, b = z_draw.mean(), (z_unif ** 2).mean()
    se = np.sqrt(z_draw.var() / len(z_draw) + (z_unif ** 2).var() / len(z_unif))
    print(f'  E_draw[2^n p] = {a:.4f}, E_unif[(2^n p)^2] = {b:.4f}, s.e. {se:.4f}', flush=True)
    assert abs(a - b) < 5 * se
    # the same step with the last pass deferred (masses only; draws recompute
    # the tiles they read): the same pyramid, draws and p bit for bit, and the
    # finished state equal to the stored one
    p_probe = s.slab.gather_prob(probe ^ flip)
    s.defer_leaves = True
    s._desc = s._make_desc()
    assert s.plan_text.split()[:-1] == s.slab.describe(s._desc)[1].split()[:-1]
    s.compute()
    assert np.array_equal(sl.pyramid.cpu().numpy(), pyr)
    idx2, p2 = s.sample(uni)
    assert np.array_equal(idx2, idx) and np.array_equal(p2, p)
    assert np.array_equal([s.probability_at(int(x)) for x in probe[:64]], p_probe[:64])
    s.finish()
    assert np.array_equal(fetch(blocks), leaves)
    del s, state
    gc.collect()
    torch.cuda.empty_cache()
