"""Phase and lookup kernels on the device, bit for bit.

The sign stages of every pipeline are exact in binary floating point:
(-1)^{q(x)} times a power of two, and every partial Walsh sum of a quadratic
sign (-1)^{Q(z)} is 0 or +-2^j.  So the device's complex64 output of these
stages must equal the oracle's float64 values exactly (array_equal), with no
tolerance:

* graph-phase init alone, and graph-phase init + Walsh (the slab pipeline's
  first stage);
* delta init + Walsh + graph phase (the literal pipeline without gates);
* the covariant init of the conditional pipeline (chunked map and sign
  tables) with unit weights, alone and followed by Walsh and the A-side
  phase, on the reference's induced 24- and 40-qubit instances.

Each runs under several tile plans, the n = 34 plan's shape among them.
"""
import numpy as np
import pytest

from oracle import graph_state as gs
from rgs70_classical_challenge_amd.instances import instance

pytestmark = pytest.mark.gpu

torch = pytest.importorskip('torch')

PLANS = [None, dict(low_bits=8, tile_bits=8, lane_bits=3), dict(low_bits=10, tile_bits=12, lane_bits=4, round_bits=2),
         dict(low_bits=12, tile_bits=14, lane_bits=5, round_bits=2),
         dict(low_bits=8, tile_bits=12, lane_bits=4, top_bits=4),
         dict(low_bits=13, tile_bits=14, lane_bits=6, round_bits=0, group_order=1)]


def _run(sl, d):
    sl.run(d, mass=False)
    sl.sync()
    return sl.state.cpu().numpy()


def _exact(got, want):
    got = np.asarray(got)
    assert np.array_equal(got.real.astype(np.float64), np.real(want)), \
        np.max(np.abs(got.real.astype(np.float64) - np.real(want)))
    assert np.array_equal(got.imag.astype(np.float64), np.imag(want) if np.iscomplexobj(want) else 0 * want)


@pytest.mark.parametrize('plan', PLANS)
@pytest.mark.parametrize('n', [16, 20])
def test_graph_phase_stages_bit_exact(n, plan):
    from rgs70_classical_challenge_amd import _lib
    from rgs70_classical_challenge_amd.engine import _Slab, form_rows, graph_phase_lut
    g, _ = instance(n, 500 + n)
    sl = _Slab(n, 0, tuning=plan)
    lut = sl.to_device(graph_phase_lut(g))
    rows = form_rows(g)
    sign = np.ones(1 << n)
    gs.quadratic_phase(sign, g)
    print(f'case n={n} plan={plan}', flush=True)
    # graph-phase init alone
    _exact(_run(sl, sl.desc(_lib.INIT_GRAPH, 1.0, init_lut=lut, init_rows=rows, mass=False)), sign)
    # init + unnormalised Walsh over every bit
    want = sign * 2.0 ** -n
    gs.walsh(want)
    _exact(_run(sl, sl.desc(_lib.INIT_GRAPH, 2.0 ** -n, init_lut=lut, init_rows=rows, walsh=True, mass=False)), want)
    # delta -> Walsh -> graph phase
    _exact(_run(sl, sl.desc(_lib.INIT_DELTA, 1.0, walsh=True, phase_lut=lut, phase_rows=rows, mass=False)), sign)


@pytest.mark.parametrize('wbits', [8, 9, 'popcount'])
@pytest.mark.parametrize('plan', [PLANS[0], PLANS[2], PLANS[5]])
@pytest.mark.parametrize('name', ['induced_24q.npz', 'induced_40q.npz'])
def test_covariant_init_bit_exact(golden_dir, name, plan, wbits):
    from rgs70_classical_challenge_amd import _lib
    from rgs70_classical_challenge_amd._lib import call, ptr
    from rgs70_classical_challenge_amd.engine import _Slab, cov_form, form_rows
    d = np.load(f'{golden_dir}/{name}')
    g, A, B = d['graph'], d['A'].tolist(), d['B'].tolist()
    m = len(A)
    assert m % 2 == 0                       # 2^{-m/2} is then a power of two
    ones = np.ones((len(g), 2, 2), np.complex128)
    pop = wbits == 'popcount'
    rng = np.random.default_rng(m + (0 if pop else wbits))
    outB = rng.integers(0, 2, m).astype(np.uint8)
    sl = _Slab(m, 0, tuning=plan, kind='c')
    map_lut = np.zeros(((m + 7) // 8) * 256 * 2, np.uint64)
    aphase = np.zeros(((m + 7) // 8) * 256, np.uint64)
    Aa, Ba = np.asarray(A, np.int32), np.asarray(B, np.int32)
    gg = np.ascontiguousarray(g, np.uint8)
    call('rgs_cov_luts', ptr(gg), len(g), ptr(Aa), ptr(Ba), m, ptr(map_lut), ptr(aphase))
    rows, cols = cov_form(map_lut, m)
    if pop:
        # the popcount table: with unit weights every valid entry is 2^{-m/2}
        wl = np.zeros((m + 1) ** 2 * 2, np.float32)
        call('rgs_cov_weight_table', ptr(np.ascontiguousarray(ones[0])), ptr(outB), m, 2.0 ** (-m / 2), ptr(wl))
        assert np.all(wl[1::2] == 0)
    else:
        wl = np.zeros((((m + wbits - 1) // wbits) << wbits) * 2, np.float32)
        call('rgs_cov_weight_lut_bits', ptr(ones), len(g), ptr(Ba), ptr(outB), m, 2.0 ** (-m / 2), wbits, ptr(wl))
        assert np.all(wl[0::2] != 0) and np.all(wl[1::2] == 0)
    dmap, dwl, dph = sl.to_device(map_lut), sl.to_device(wl), sl.to_device(aphase)
    print(f'case {name} m={m} plan={plan} wbits={wbits}', flush=True)
    init = gs.covariant_init(g, ones, A, B, outB.tolist())

    def desc(**kw):
        d = sl.desc(_lib.INIT_COV, 0.0, init_lut=dmap, init_wlut=dwl, init_rows=rows, init_cols=cols,
                    init_wbits=8 if pop else wbits, mass=False, **kw)
        if pop:
            d.init_wpop = 1
            d.init_omask = int(sum(1 << j for j in range(m) if outB[j]))
        return d

    got = _run(sl, desc())
    _exact(got, init)
    want = init.copy()
    gs.walsh(want)
    gs.quadratic_phase(want, g[np.ix_(A, A)])
    got = _run(sl, desc(walsh=True, phase_lut=dph, phase_rows=form_rows(g[np.ix_(A, A)])))
    _exact(got, want)
