"""Pins of the complex128 state-vector oracle against brute force, library
routines, the reference's invariants and its fixtures (CPU only)."""
import numpy as np
import pytest
import scipy.linalg

from oracle import graph_state as gs
from rgs70_classical_challenge_amd.instances import instance, haar_unitaries


def kron_all(mats):
    out = np.array([[1.0 + 0j]])
    for m in mats:
        out = np.kron(m, out)      # qubit 0 is the least significant index bit
    return out


def brute_graph_state(g):
    """prod CZ_ab H^n |0> with explicit dense matrices."""
    n = len(g)
    v = np.zeros(1 << n, complex)
    v[0] = 1
    v = kron_all([gs.HADAMARD] * n) @ v
    idx = np.arange(1 << n)
    for a in range(n):
        for b in range(a + 1, n):
            if g[a, b]:
                v[((idx >> a) & 1) & ((idx >> b) & 1) == 1] *= -1
    return v


def random_graph(n, rng):
    g = np.triu(rng.integers(0, 2, (n, n), dtype=np.uint8), 1)
    return g | g.T


def load(golden_dir, name):
    d = np.load(f'{golden_dir}/{name}', allow_pickle=False)
    return d['graph'], d['measurement_unitaries'], d['A'].tolist(), d['B'].tolist()


def test_graph_state_vs_dense_circuit():
    rng = np.random.default_rng(1)
    for n in (3, 5, 7):
        g = random_graph(n, rng)
        assert np.allclose(gs.graph_state(g), brute_graph_state(g), atol=1e-14)


def test_walsh_is_hadamard_matrix():
    rng = np.random.default_rng(2)
    for n in (1, 4, 8):
        v = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
        w = v.copy()
        gs.walsh(w)
        assert np.allclose(w, scipy.linalg.hadamard(1 << n) @ v, atol=1e-10)
        gs.walsh(w)                                   # involution up to 2^n
        assert np.allclose(w, v * (1 << n), atol=1e-9)


def test_apply_one_is_kron():
    rng = np.random.default_rng(3)
    n = 5
    u = haar_unitaries(n, rng)
    v = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    for q in range(n):
        w = v.copy()
        gs.apply_one(w, u[q], q)
        mats = [np.eye(2)] * n
        mats[q] = u[q]
        assert np.allclose(w, kron_all(mats) @ v, atol=1e-13)


def test_full_state_vs_kron_and_dual_pipeline():
    rng = np.random.default_rng(4)
    for n in (4, 6, 9):
        g = random_graph(n, rng)
        u = haar_unitaries(n, rng)
        ref = kron_all(list(u)) @ brute_graph_state(g)
        assert np.allclose(gs.full_state(g, u), ref, atol=1e-13)
        assert np.allclose(gs.dual_pipeline(g, u), ref, atol=1e-13)
        assert abs(np.vdot(ref, ref).real - 1) < 1e-12


def test_dual_walsh_stage_is_exact_powers_of_two():
    """The Walsh sums of (-1)^q are 0 or +-2^j: the init+WHT stage is exact."""
    rng = np.random.default_rng(5)
    g, _ = instance(12, 7)
    v = np.full(1 << 12, 1.0)
    gs.quadratic_phase(v, g)
    gs.walsh(v)
    nz = v[v != 0]
    assert np.all(np.abs(nz) == 2.0 ** np.round(np.log2(np.abs(nz))))
    v32 = np.full(1 << 12, 1.0, np.float32)
    gs.quadratic_phase(v32, g)
    gs.walsh(v32)
    assert np.array_equal(v32.astype(np.float64), v)
    del rng


def test_quadratic_parity_matches_edge_sum():
    rng = np.random.default_rng(6)
    g = random_graph(11, rng)
    ids = np.arange(1 << 11, dtype=np.uint64)
    bits = ((ids[:, None] >> np.arange(11, dtype=np.uint64)) & np.uint64(1)).astype(np.int64)
    q = np.einsum('xa,ab,xb->x', bits, np.triu(g.astype(np.int64), 1), bits) & 1
    assert np.array_equal(gs.quadratic_parity(ids, g).astype(np.int64), q)


def test_conditional_slices_full_state():
    """graph_sampler.selfcheck (graph_sampler.py:147-158): the conditional A
    vector equals the slice of the full state, and a full-rank cut gives an
    exactly uniform B marginal."""
    rng = np.random.default_rng(20261030)
    for n in (4, 6, 8):
        for _ in range(3):
            g = random_graph(n, rng)
            u = haar_unitaries(n, rng)
            full = gs.full_state(g, u)
            A, B = list(range(n // 2)), list(range(n // 2, n))
            full_rank = gs.gf2_rank(g[np.ix_(A, B)]) == len(B)
            for z in range(1 << len(B)):
                ob = [(z >> i) & 1 for i in range(len(B))]
                ca, pb = gs.conditional(g, u, A, B, ob)
                assert np.allclose(ca, full[z << len(A):(z + 1) << len(A)], atol=2e-14)
                if full_rank:
                    assert abs(pb - 2.0 ** (-len(B))) < 1e-14


@pytest.mark.parametrize('name', ['induced_12q.npz', 'induced_24q.npz'])
def test_covariant_and_deferred_match_conditional(golden_dir, name):
    g, u, A, B = load(golden_dir, name)
    rng = np.random.default_rng(2026103007)
    assert gs.gf2_rank(g[np.ix_(B, A)]) == len(A)
    for _ in range(2):
        ob = rng.integers(0, 2, len(B)).tolist()
        ref, pb = gs.conditional(g, u, A, B, ob)
        ref /= np.sqrt(pb)
        v = gs.covariant(g, u, A, B, ob)
        assert np.max(np.abs(v - ref)) < 1e-12
        raw, final = gs.deferred(g, u, A, B, ob)
        assert np.max(np.abs(final - ref)) < 1e-12
        before = np.abs(raw.reshape(2, -1)) ** 2
        after = np.abs(final.reshape(2, -1)) ** 2
        assert np.max(np.abs(before.sum(0) - after.sum(0))) < 1e-12


def test_fourier_covariance_identity(golden_dir):
    """(M x).(M^{-T} z) = x.z mod 2 on RGS70's 35x35 cut, exhaustively on a
    10-dimensional sublattice (fourier_covariant_sampler.py:160-175)."""
    g, _, A, B = load(golden_dir, 'rgs_graph.npz')
    M = g[np.ix_(B, A)]
    N = gs.gf2_inverse(M).T
    m = len(A)
    rng = np.random.default_rng(11)
    bz = rng.integers(0, 1 << m, 10, dtype=np.uint64)
    bx = rng.integers(0, 1 << m, 10, dtype=np.uint64)
    sub = np.arange(1024, dtype=np.uint64)
    iz = np.zeros(1024, np.uint64)
    ix = np.zeros(1024, np.uint64)
    for j in range(10):
        sel = (sub >> np.uint64(j)) & np.uint64(1)
        iz ^= sel * bz[j]
        ix ^= sel * bx[j]
    lhs = gs.parity64(gs.map_indices(ix, M)[:, None] & gs.map_indices(iz, N)[None, :])
    rhs = gs.parity64(ix[:, None] & iz[None, :])
    assert np.array_equal(lhs, rhs)
    ids = rng.integers(0, 1 << m, 4096, dtype=np.uint64)
    assert np.array_equal(gs.map_indices(gs.map_indices(ids, N), M.T), ids)


@pytest.mark.parametrize('name', ['induced_12q.npz', 'induced_32q.npz'])
def test_chunk_tables_vs_per_site(golden_dir, name):
    """Chunk LUTs against per-site evaluation on every id for m <= 16
    (init_specialization.py:268-280 test design)."""
    g, u, A, B = load(golden_dir, name)
    m = len(A)
    N = gs.gf2_inverse(g[np.ix_(B, A)]).T
    gB = g[np.ix_(B, B)]
    Q, ell = gs.transformed_quadratic(gB, N)
    for w in (3, 8):
        lut = gs.phase_map_tables(m, N, Q, ell, w)
        ids = np.arange(1 << m, dtype=np.uint64)
        b, par = gs.lookup_map_phase(lut, ids, w)
        assert np.array_equal(b, gs.map_indices(ids, N))
        assert np.array_equal(par, gs.quadratic_parity(b, gB))
    rng = np.random.default_rng(3)
    ob = rng.integers(0, 2, m).tolist()
    tab = gs.weight_tables(u, B, ob, 2.0 ** (-m / 2))
    bb = gs.map_indices(np.arange(1 << m, dtype=np.uint64), N)
    wt = np.ones(len(bb), complex)
    for c in range(tab.shape[0]):
        wt *= tab[c][((bb >> np.uint64(8 * c)) & np.uint64(255)).astype(np.int64)]
    sign = 1.0 - 2.0 * gs.quadratic_parity(bb, gB).astype(float)
    assert np.allclose(wt * sign, gs.covariant_init(g, u, A, B, ob), rtol=1e-13, atol=0)


def test_graph_phase_tables():
    g, _ = instance(20, 99)
    lut = gs.graph_phase_tables(g)
    ids = np.random.default_rng(0).integers(0, 1 << 20, 20000, dtype=np.uint64)
    b, par = gs.lookup_map_phase(lut, ids)
    assert np.array_equal(b, ids)
    assert np.array_equal(par, gs.quadratic_parity(ids, g))


def test_threaded_state_equals_full_state():
    """The CPU baseline's threaded state build is the same arithmetic split
    over slices: bit-identical to full_state."""
    for n, threads in ((10, 3), (14, 8)):
        g, u = instance(n, 40 + n)
        assert np.array_equal(gs.full_state_threads(g, u, threads), gs.full_state(g, u))
