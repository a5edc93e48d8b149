"""The C ABI library on the CPU: it loads, exports every declared symbol, and
its host planning functions (GF(2) algebra, lookup tables, pass schedules)
agree with the oracle.  No device calls."""
import ctypes
import os
import re

import numpy as np
import pytest

from oracle import graph_state as gs
from rgs70_classical_challenge_amd import _lib
from rgs70_classical_challenge_amd.engine import cov_form
from rgs70_classical_challenge_amd.instances import instance

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    inc = os.path.join(ROOT, 'include')
    for f in os.listdir(inc):
        if f.endswith('.h'):
            text = open(os.path.join(inc, f)).read()
            text = re.sub(r'/\*.*?\*/', '', text, flags=re.S)
            names |= set(re.findall(r'\b(rgs_[a-z0-9_]+)\s*\(', text))
    return names


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    names = declared_symbols()
    assert len(names) >= 10
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing
    assert lib.rgs_abi_version() == 4


def test_binding_covers_declarations():
    # the scan and MADE modules register their entry points on import
    from rgs70_classical_challenge_amd import made, scan  # noqa: F401
    missing = declared_symbols() - set(_lib.SIGNATURES)
    assert not missing, missing


def test_gf2_inverse_and_rank(golden_dir):
    from rgs70_classical_challenge_amd.engine import gf2_inverse, gf2_rank
    d = np.load(f'{golden_dir}/rgs_graph.npz')
    g, A, B = d['graph'], d['A'].tolist(), d['B'].tolist()
    M = g[np.ix_(B, A)]
    assert np.array_equal(gf2_inverse(M), gs.gf2_inverse(M))
    assert gf2_rank(M) == gs.gf2_rank(M) == 35
    rng = np.random.default_rng(0)
    for _ in range(5):
        m = rng.integers(0, 2, (30, 70), dtype=np.uint8)
        assert gf2_rank(m) == gs.gf2_rank(m)
    with pytest.raises(ValueError):
        gf2_inverse(np.ones((4, 4), np.uint8))


def test_graph_phase_lut_matches_oracle():
    from rgs70_classical_challenge_amd.engine import graph_phase_lut
    for n in (6, 20, 34, 63):
        g, _ = instance(n if n % 2 == 0 else n, 5)
        lut = graph_phase_lut(g)
        ref = gs.graph_phase_tables(g)[:, :, 1].reshape(-1)
        assert np.array_equal(lut, ref)


def test_popcount_weight_table(golden_dir):
    """The identical-unitary weight table: the oracle's table equals the
    generic product prod_j u0[o_j, b_j] for every b at m = 10 (brute force)
    and is zero at impossible counts; the library's complex64 table is the
    oracle's rounded once, on the 70-qubit instance's unitary at m = 35."""
    rng = np.random.default_rng(7)
    m = 10
    u0 = np.linalg.qr(rng.normal(size=(2, 2)) + 1j * rng.normal(size=(2, 2)))[0]
    ob = rng.integers(0, 2, m)
    o = int(sum(1 << j for j in range(m) if ob[j]))
    tab = gs.popcount_weight_table(u0, ob, 0.5)
    for b in range(1 << m):
        want = 0.5 * np.prod([u0[ob[j], (b >> j) & 1] for j in range(m)])
        got = tab[bin(b).count('1'), bin(b & o).count('1')]
        assert abs(got - want) <= 1e-14, b
    r = int(ob.sum())
    for k in range(m + 1):
        for h in range(m + 1):
            if not max(0, r + k - m) <= h <= min(r, k):
                assert tab[k, h] == 0
    d = np.load(f'{golden_dir}/rgs_graph.npz')
    u = d['measurement_unitaries']
    assert all(np.array_equal(u[q], u[0]) for q in range(len(u)))    # the instance's unitaries are identical
    m = 35
    for seed in range(3):
        ob = np.random.default_rng(seed).integers(0, 2, m).astype(np.uint8)
        out = np.zeros(2 * (m + 1) ** 2, np.float32)
        u0 = np.ascontiguousarray(u[0])
        _lib.call('rgs_cov_weight_table', _lib.ptr(u0), _lib.ptr(ob), m, 2.0 ** (-m / 2), _lib.ptr(out))
        got = out.view(np.complex64).reshape(m + 1, m + 1)
        want = gs.popcount_weight_table(u0, ob, 2.0 ** (-m / 2))
        assert np.array_equal(got == 0, want == 0)
        nz = want != 0
        assert np.all(np.abs(got[nz] - want[nz]) <= 6e-8 * np.abs(want[nz]))
    with pytest.raises(ValueError):
        _lib.call('rgs_cov_weight_table', _lib.ptr(u0), _lib.ptr(np.full(m, 2, np.uint8)), m, 1.0, _lib.ptr(out))


def test_cov_luts_match_oracle(golden_dir):
    for name in ('induced_12q.npz', 'induced_40q.npz', 'rgs_graph.npz'):
        d = np.load(f'{golden_dir}/{name}')
        g, u, A, B = d['graph'], d['measurement_unitaries'], d['A'].tolist(), d['B'].tolist()
        m = len(A)
        chunks = (m + 7) // 8
        map_lut = np.zeros(chunks * 256 * 2, np.uint64)
        aph = np.zeros(chunks * 256, np.uint64)
        gc, Aa, Ba = np.ascontiguousarray(g), np.asarray(A, np.int32), np.asarray(B, np.int32)
        _lib.call('rgs_cov_luts', _lib.ptr(gc), len(g), _lib.ptr(Aa), _lib.ptr(Ba), m, _lib.ptr(map_lut),
                  _lib.ptr(aph))
        N = gs.gf2_inverse(g[np.ix_(B, A)]).T
        Q, ell = gs.transformed_quadratic(g[np.ix_(B, B)], N)
        ref = gs.phase_map_tables(m, N, Q, ell).reshape(-1)
        assert np.array_equal(map_lut, ref), name
        assert np.array_equal(aph, gs.graph_phase_tables(g[np.ix_(A, A)])[:, :, 1].reshape(-1))
        # rows of the sign form and columns of the map, read off the table by
        # the product's cov_form, against the oracle's Q, ell and N
        rows, cols = cov_form(map_lut, m)
        want = [r | (int(ell[a]) << a) for a, r in enumerate(gs.row_masks(Q | Q.T))]
        assert rows.tolist() == want, name
        assert cols.tolist() == gs.row_masks(np.asarray(N).T), name
        ob = np.random.default_rng(1).integers(0, 2, m).astype(np.uint8)
        w = np.zeros(chunks * 256 * 2, np.float32)
        uc = np.ascontiguousarray(u)
        _lib.call('rgs_cov_weight_lut', _lib.ptr(uc), len(g), _lib.ptr(Ba), _lib.ptr(ob), m, 2.0 ** (-m / 2),
                  _lib.ptr(w))
        wc = w.view(np.complex64).reshape(chunks, 256)
        wref = gs.weight_tables(u, B, ob.tolist(), 2.0 ** (-m / 2))
        for c in range(chunks):
            valid = 1 << min(8, m - 8 * c)
            assert np.allclose(wc[c, :valid], wref[c, :valid], rtol=2e-7, atol=0)
        # nine-bit chunks: the same weight for every B string as the pinned 8-bit tables
        c9 = (m + 8) // 9
        w9 = np.zeros(c9 * 512 * 2, np.float32)
        _lib.call('rgs_cov_weight_lut_bits', _lib.ptr(uc), len(g), _lib.ptr(Ba), _lib.ptr(ob), m, 2.0 ** (-m / 2),
                  9, _lib.ptr(w9))
        w9 = w9.view(np.complex64).reshape(c9, 512)
        w8 = np.zeros(chunks * 256 * 2, np.float32)
        _lib.call('rgs_cov_weight_lut_bits', _lib.ptr(uc), len(g), _lib.ptr(Ba), _lib.ptr(ob), m, 2.0 ** (-m / 2),
                  8, _lib.ptr(w8))
        assert np.array_equal(w8, w)
        for b in np.random.default_rng(3).integers(0, 1 << m, 300, dtype=np.int64):
            b = int(b)
            p8 = np.prod([complex(wc[c, (b >> (8 * c)) & 255]) for c in range(chunks)])
            p9 = np.prod([complex(w9[c, (b >> (9 * c)) & 511]) for c in range(c9)])
            assert abs(p9 - p8) <= 1e-6 * abs(p8), (name, b)


def test_cov_luts_reject_singular_cut():
    g = np.zeros((8, 8), np.uint8)
    A, B, out = np.arange(4, dtype=np.int32), np.arange(4, 8, dtype=np.int32), np.zeros(512, np.uint64)
    with pytest.raises(ValueError, match='singular'):
        _lib.call('rgs_cov_luts', _lib.ptr(g), 8, _lib.ptr(A), _lib.ptr(B), 4, _lib.ptr(out), None)


def describe(nbits, walsh, gates=True, phase=False, mass=True, init=_lib.INIT_GRAPH, low=14, tile=14, lane=4,
             rounds=0, phase_rounds=-1, top=0, order=0, groups=(), defer=False, desc_out=None, dphase=False,
             fuse=0, dphase_host=True):
    d = _lib.SlabDesc()
    d.state = 1
    d.nbits = nbits
    d.init_kind = init
    d.init_lut = 1
    d.do_walsh = int(walsh)
    d.phase_lut = 1 if phase else None
    d.rot = 1 if gates else None
    d.gate_limit = nbits
    d.mass = 1 if mass else None
    d.mass_bits = 8
    d.low_bits, d.tile_bits, d.lane_bits, d.round_bits = low, tile, lane, rounds
    d.phase_round_bits = phase_rounds
    d.top_bits = top
    d.group_order = order
    for i, (bits, glane, grounds) in enumerate(groups):
        d.group_bits[i], d.group_lanes[i], d.group_rounds[i] = bits, glane, grounds
    d.defer_leaves = int(defer)
    d.dphase_lut = 1 if dphase else None
    d.dphase_host = 1 if dphase and dphase_host else None     # planned only, never read here
    d.fuse_walsh = fuse
    if desc_out is not None:
        desc_out.append(d)
    n = ctypes.c_int()
    buf = ctypes.create_string_buffer(4096)
    _lib.call('rgs_slab_describe', ctypes.byref(d), ctypes.byref(n), buf, 4096)
    return n.value, buf.value.decode()


def test_pass_schedules():
    # n=34: 14 contiguous + two strided groups of 10 bits
    n, text = describe(34, walsh=True)
    assert n == 5, text
    assert text.split()[0].startswith('low[init][wht]')
    assert '[wht][gates]:24+10' in text and text.split()[-1].startswith('low[gates][mass]')
    n, text = describe(34, walsh=False)
    assert n == 3, text
    n, text = describe(12, walsh=True)
    assert n == 1 and '[wht]' in text and '[mass]' in text
    n, text = describe(20, walsh=True, low=8, tile=8, lane=3)
    assert n == 1 + 2 + 1 + 2 + 1, text
    with pytest.raises(ValueError):
        describe(34, walsh=True, tile=15)
    with pytest.raises(ValueError):
        describe(34, walsh=True, lane=8)           # no kernel for 8 lane bits
    with pytest.raises(ValueError):
        describe(34, walsh=True, rounds=3)
    # LDS rounds: at most the lane bits, two, and tile bits - 10
    n, text = describe(34, walsh=True, rounds=2)
    assert all(tok.endswith('rb2') for tok in text.split() if tok.startswith('strided')), text
    n, text = describe(20, walsh=True, low=10, tile=12, lane=4, rounds=2)
    assert all(tok.endswith('rb0') for tok in text.split() if tok.startswith('strided')), text  # 9-bit tiles
    # the phase pass may take its own rounds; -1 keeps round_bits
    n, text = describe(34, walsh=True, phase=True, low=13, tile=13, lane=6, rounds=0, phase_rounds=2)
    toks = text.split()
    assert [t for t in toks if '[phase]' in t][0].endswith('rb2'), text
    assert all(t.endswith('rb0') for t in toks if t.startswith('strided') and '[phase]' not in t), text
    n, text = describe(34, walsh=True, phase=True, low=13, tile=13, lane=6, rounds=1, phase_rounds=-1)
    assert all(t.endswith('rb1') for t in text.split() if t.startswith('strided')), text
    with pytest.raises(ValueError):
        describe(34, walsh=True, phase=True, phase_rounds=3)
    # a narrow phase group: last with the Walsh transform, first without
    n, text = describe(34, walsh=True, phase=True, low=14, tile=14, lane=6, top=4)
    assert text.split()[1:4] == ['strided[wht]:14+8/tb14/c6/rb0', 'strided[wht]:22+8/tb14/c6/rb0',
                                 'strided[wht][phase][gates]:30+4/tb10/c6/rb0'], text
    n, text = describe(34, walsh=False, phase=True, low=14, tile=14, lane=6, top=4)
    assert text.split()[0].startswith('strided[init][phase][gates]:14+4/tb10'), text
    for bad in (3, 9, -1):
        with pytest.raises(ValueError):
            describe(34, walsh=True, phase=True, low=14, tile=14, lane=6, top=bad)
    # groups walked from the high bits down: the phase lands in the lowest group
    n, text = describe(34, walsh=True, phase=True, low=13, tile=13, lane=6, order=1)
    assert [t.split('/')[0] for t in text.split()] == [
        'low[init][wht]:0+13', 'strided[wht]:27+7', 'strided[wht]:20+7', 'strided[wht][phase][gates]:13+7',
        'strided[gates]:20+7', 'strided[gates]:27+7', 'low[gates][mass]:0+13'], text
    n, text = describe(34, walsh=True, phase=True, low=14, tile=14, lane=6, top=4, order=1)
    assert '[phase]' in text.split()[3] and text.split()[3].startswith('strided[wht][phase][gates]:14+4/tb10'), text


def test_explicit_group_schedules():
    # explicit groups from the low bits up, each with its own lane bits and
    # rounds; tile bits = lane bits + width; rounds capped at tile bits - 10
    n, text = describe(35, walsh=True, phase=True, low=12, groups=[(8, 5, 0), (8, 6, 1), (7, 7, 1)])
    assert text.split() == [
        'low[init][wht]:0+12/tb12/c0/rb0', 'strided[wht]:12+8/tb13/c5/rb0', 'strided[wht]:20+8/tb14/c6/rb1',
        'strided[wht][phase][gates]:28+7/tb14/c7/rb1', 'strided[gates]:20+8/tb14/c6/rb1',
        'strided[gates]:12+8/tb13/c5/rb0', 'low[gates][mass]:0+12/tb12/c0/rb0'], text
    n, text = describe(34, walsh=True, phase=True, low=13, order=1, groups=[(7, 6, 0), (7, 6, 0), (7, 7, 2)])
    assert text.split()[1] == 'strided[wht]:27+7/tb14/c7/rb2' and '[phase]' in text.split()[3], text
    n, text = describe(22, walsh=True, phase=True, low=8, groups=[(7, 7, 2), (7, 6, 1)])
    assert text.split()[1:3] == ['strided[wht]:8+7/tb14/c7/rb2', 'strided[wht][phase][gates]:15+7/tb13/c6/rb1'], text
    n, text = describe(20, walsh=False, phase=True, low=8, groups=[(5, 6, 1), (7, 7, 1)])
    assert text.split()[0] == 'strided[init][phase][gates]:8+5/tb11/c6/rb1', text
    for bad in ([(7, 6, 0), (6, 7, 1)],          # bits 8..20 only
                [(7, 8, 0), (7, 6, 0)],          # lane bits above 7
                [(7, 1, 0), (7, 6, 0)],          # lane bits below 2
                [(8, 7, 0), (6, 6, 0)],          # 15 tile bits
                [(3, 6, 0), (11, 3, 0)],         # narrower than a register window
                [(7, 6, 3), (7, 6, 0)]):         # rounds above 2
        with pytest.raises(ValueError):
            describe(22, walsh=True, low=8, groups=bad)
    with pytest.raises(ValueError):
        describe(20, walsh=True, low=6, groups=[(7, 7, 0), (7, 6, 0)])   # 7 lane bits, 6 bits below the group


def test_fused_walsh_schedules():
    """fuse_walsh: every pass takes the Walsh butterflies, the diagonal
    phase and the rotations of its own bits - the first strided group makes
    the state and the contiguous pass ends with the masses; refused with a
    sign table after the Walsh transform or without the transform."""
    n, text = describe(34, walsh=True, dphase=True, low=13, tile=13, lane=6, fuse=1)
    assert [t.split('/')[0] for t in text.split()] == [
        'strided[init][wht][phase][gates]:13+7', 'strided[wht][phase][gates]:20+7',
        'strided[wht][phase][gates]:27+7', 'low[wht][phase][gates][mass]:0+13'], text
    n, text = describe(34, walsh=True, dphase=True, low=13, tile=13, lane=6, order=1, fuse=1)
    assert [t.split('/')[0] for t in text.split()][:3] == [
        'strided[init][wht][phase][gates]:27+7', 'strided[wht][phase][gates]:20+7',
        'strided[wht][phase][gates]:13+7'], text
    n, text = describe(34, walsh=True, dphase=True, low=13, tile=13, lane=6, fuse=1, defer=True)
    assert n == 4 and text.split()[-1] == 'low[wht][phase][gates][mass][leaf]:0+13/tb13/c0/rb0', text
    n, text = describe(12, walsh=True, dphase=True, fuse=1)
    assert n == 1 and text.startswith('low[init][wht][phase][gates][mass]'), text
    with pytest.raises(ValueError, match='fuse_walsh'):
        describe(34, walsh=True, phase=True, fuse=1)
    with pytest.raises(ValueError, match='fuse_walsh'):
        describe(34, walsh=False, dphase=True, fuse=1)
    with pytest.raises(ValueError, match='fuse_walsh'):
        describe(34, walsh=True, dphase=True, fuse=2)
    with pytest.raises(ValueError, match='dphase_host'):
        describe(34, walsh=True, dphase=True, fuse=1, dphase_host=False)


def test_deferred_leaf_schedules():
    """defer_leaves marks the last contiguous pass [leaf] (masses only) and
    is refused where that pass would make the state or has no mass level;
    the leaf entry points check their arguments before any launch."""
    keep = []
    n, text = describe(34, walsh=True, phase=True, low=13, tile=13, lane=6, order=1, defer=True, desc_out=keep)
    full = describe(34, walsh=True, phase=True, low=13, tile=13, lane=6, order=1)[1]
    assert text.split()[:-1] == full.split()[:-1], text
    assert text.split()[-1] == 'low[gates][mass][leaf]:0+13/tb13/c0/rb0', text
    d = keep[0]
    bits = ctypes.c_int()
    _lib.call('rgs_slab_leaf_bits', ctypes.byref(d), ctypes.byref(bits))
    assert bits.value == 13
    nb = ctypes.c_size_t()
    _lib.call('rgs_draw_leaves_workspace', ctypes.byref(d), 10000, ctypes.byref(nb))
    assert nb.value == ((10000 * 24 + 255) // 256) * 256 + (10000 << 13) * 8
    with pytest.raises(ValueError, match='2\\^24'):
        _lib.call('rgs_slab_leaf_tiles', ctypes.byref(d), 1, (1 << 24) + 1, 1, None)
    with pytest.raises(ValueError, match='null'):
        _lib.call('rgs_slab_leaf_tiles', ctypes.byref(d), None, 4, 1, None)
    with pytest.raises(ValueError, match='2\\^24'):
        _lib.call('rgs_draw_leaves', ctypes.byref(d), 1, 1, (1 << 24) + 1, 1, 1 << 40, 1, 1, None)
    with pytest.raises(ValueError, match='aligned'):
        _lib.call('rgs_draw_leaves', ctypes.byref(d), 1, 1, 16, 8, 1 << 40, 1, 1, None)
    with pytest.raises(ValueError, match='smaller'):
        _lib.call('rgs_draw_leaves', ctypes.byref(d), 1, 1, 16, 256, 1024, 1, 1, None)
    n, text = describe(35, walsh=True, phase=True, low=12, groups=[(8, 5, 0), (8, 6, 1), (7, 7, 1)], defer=True)
    assert text.split()[-1] == 'low[gates][mass][leaf]:0+12/tb12/c0/rb0', text
    # one pass makes the state: nothing to defer; no mass level: nothing to draw
    with pytest.raises(ValueError, match='defer_leaves'):
        describe(12, walsh=True, defer=True)
    with pytest.raises(ValueError, match='defer_leaves'):
        describe(34, walsh=True, mass=False, defer=True)
    # a descriptor that does not defer has no leaf tiles
    keep = []
    describe(34, walsh=True, desc_out=keep)
    with pytest.raises(ValueError, match='does not defer'):
        _lib.call('rgs_slab_leaf_bits', ctypes.byref(keep[0]), ctypes.byref(bits))
    with pytest.raises(ValueError, match='does not defer'):
        _lib.call('rgs_slab_finish', ctypes.byref(keep[0]), None)


def test_gate_shear_factorisation(golden_dir):
    """u = X^f D1 m M(t) D2 with M(t) = [[1, -t], [t, 1]], |t| <= 1 and
    m = (1 + t^2)^(-1/2): reconstructed from the library's outputs to float32
    accuracy for the 70-qubit instance's unitary, Haar unitaries, and the
    diagonal and anti-diagonal corner cases; the flips are exactly the qubits
    below gate_limit whose rotation is past 45 degrees; the common scale is
    the product of the m's; phase-table entries have the smallest norm error
    of their one-ulp neighbourhood."""
    import ctypes
    d = np.load(f'{golden_dir}/rgs_graph.npz')
    diag = np.diag([np.exp(0.3j), np.exp(-1.1j)])
    anti = np.array([[0, -np.exp(0.7j)], [np.exp(0.2j), 0]])
    mats = np.concatenate([d['measurement_unitaries'][:4], instance(12, 5)[1], [diag, anti]])
    n = len(mats)
    u = np.ascontiguousarray(mats.astype(np.complex128))
    X = np.array([[0, 1], [1, 0]])
    for limit in (n, n - 3):
        tan = np.zeros(n, np.float32)
        dph = np.zeros(((n + 7) // 8) * 256 * 2, np.float32)
        d1 = np.zeros(4 * n, np.float64)
        flip = ctypes.c_uint64()
        _lib.call('rgs_gate_shear', _lib.ptr(u), n, limit, _lib.ptr(tan), _lib.ptr(dph), _lib.ptr(d1),
                  ctypes.byref(flip))
        d1c = d1.view(np.complex128).reshape(n, 2)
        dphc = dph.reshape(-1, 2).astype(np.float64) @ [1, 1j]
        assert np.all(np.abs(tan[:limit]) <= 1)
        for q in range(n):
            c, s_ = abs(u[q, 0, 0]), abs(u[q, 1, 0])
            f = (flip.value >> q) & 1
            assert f == (q < limit and s_ > c), q
            if q >= limit:
                continue
            t = float(tan[q])
            e_b = dphc[(q // 8) * 256 + (1 << (q % 8))]            # single-qubit entry of the D2 table
            rec = np.diag(d1c[q]) @ (np.array([[1, -t], [t, 1]]) / np.sqrt(1 + t * t)) @ np.diag([1.0, e_b])
            if f:
                rec = X @ rec
            assert np.max(np.abs(rec - u[q])) < 3e-7, q
        assert flip.value >> limit == 0
        scale = _lib.lib().rgs_gate_scale(_lib.ptr(tan), limit)
        assert scale == pytest.approx(np.prod(1 / np.sqrt(1 + tan[:limit].astype(np.float64) ** 2)), rel=1e-14)
    deps = np.abs(dphc) ** 2 - 1
    assert np.max(np.abs(deps)) < 1.2e-7 and abs(np.mean(deps)) < 1e-8


def test_tuning_table_key_and_entries():
    """The committed autotuner table is keyed by architecture and CU count and
    holds a valid plan for every size it lists."""
    import json
    from types import SimpleNamespace
    from rgs70_classical_challenge_amd import autotune
    props = SimpleNamespace(gcnArchName='gfx950:sramecc+:xnack-', multi_processor_count=256)
    assert autotune.device_key_of(props) == 'gfx950_256cu'
    table = json.load(open(autotune.TABLE))
    assert 'gfx950_256cu' in table
    for key, e in table['gfx950_256cu'].items():
        if key.startswith('scan:'):          # scan orders: checked in test_scan_host
            continue
        cond = key.startswith('c')           # conditional pipeline entries: 'c<m>'
        n = int(key[1:] if cond else key)
        init = _lib.INIT_COV if cond else _lib.INIT_GRAPH
        _, text = describe(n, walsh=True, phase=cond, init=init, low=e['low_bits'], tile=e['tile_bits'],
                           lane=e['lane_bits'], rounds=e.get('round_bits', 0),
                           phase_rounds=e.get('phase_round_bits', -1), top=e.get('top_bits', 0),
                           order=e.get('group_order', 0))
        assert text


def test_plain_tuner_refuses_sizes_above_its_limit():
    from rgs70_classical_challenge_amd import autotune
    with pytest.raises(ValueError, match='n <= 34'):
        autotune.tune(35)


def test_bench_plan_is_the_plan_the_gpu_tests_check():
    """The autotuned n = 34 entry the bench runs equals the tuning the GPU
    tests run at n = 26 and 34 (so its three tile kernels are the ones
    checked against the oracle)."""
    import json
    import os
    from test_gpu_headline import BENCH_PLAN, BENCH_PLAN_26, BENCH_KERNELS
    table = json.load(open(os.path.join(os.path.dirname(_lib.LIB_PATH), 'tuned.json')))
    for dev, entries in table.items():
        e = entries['34']
        keys = ('low_bits', 'tile_bits', 'lane_bits', 'round_bits', 'group_order', 'fuse_walsh', 'groups')
        assert {k: e.get(k, 0) for k in keys} == {k: BENCH_PLAN.get(k, 0) for k in keys}, dev
        assert e.get('top_bits', 0) == 0 and e.get('phase_round_bits', -1) in (-1, 0) and not e.get('dphase_global')
    for n, plan in ((34, BENCH_PLAN), (26, BENCH_PLAN_26)):
        _, text = describe(n, walsh=True, dphase=True, low=13, tile=13, lane=6, rounds=0, order=1, fuse=1,
                           groups=plan['groups'])
        assert {tok.split('/', 1)[1] for tok in text.split()} == BENCH_KERNELS, text


def test_slab_batch_rejects_bad_items_before_touching_the_device():
    """rgs_slab_batch_create validates every item on the host (no allocation,
    no launch) and names the offending item."""
    def item(desc, samples=10):
        it = _lib.SlabBatchItem()
        it.desc = ctypes.pointer(desc)
        it.uniforms, it.samples, it.out_index, it.out_p = 1, samples, 1, 1
        return it

    def desc(mass_bits=8, state=1):
        d = _lib.SlabDesc()
        d.state, d.nbits, d.init_kind, d.init_lut, d.init_rows, d.do_walsh = state, 16, _lib.INIT_GRAPH, 1, 1, 1
        d.mass, d.mass_bits = 1, mass_bits
        d.low_bits, d.tile_bits, d.lane_bits = 12, 12, 4
        d.phase_round_bits = -1
        return d

    h = ctypes.c_void_p()
    with pytest.raises(ValueError, match='need items'):
        _lib.call('rgs_slab_batch_create', None, 0, ctypes.byref(h))
    good, no_state, wide = desc(), desc(state=None), desc(mass_bits=9)
    for bad, msg in ((no_state, 'state and a mass'), (wide, 'mass_bits')):
        arr = (_lib.SlabBatchItem * 2)(item(good), item(bad))
        with pytest.raises(ValueError, match=msg):
            _lib.call('rgs_slab_batch_create', arr, 2, ctypes.byref(h))
        assert not h.value
    deferred = desc()
    deferred.defer_leaves = 1
    arr = (_lib.SlabBatchItem * 1)(item(deferred))
    with pytest.raises(ValueError, match='defer_leaves'):
        _lib.call('rgs_slab_batch_create', arr, 1, ctypes.byref(h))
    assert not h.value
    arr = (_lib.SlabBatchItem * 1)(item(good, samples=-1))
    with pytest.raises(ValueError, match='draw buffers'):
        _lib.call('rgs_slab_batch_create', arr, 1, ctypes.byref(h))
    # a null handle is an error for run and info, and a no-op for destroy
    with pytest.raises(ValueError, match='null handle'):
        _lib.call('rgs_slab_batch_run', None, None)
    _lib.call('rgs_slab_batch_destroy', None)
