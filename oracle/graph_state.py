"""Complex128 restatement of the reference's state-vector algebra.

Conventions (work/rgs_attack/graph_sampler.py:1): qubit k is bit k of the
array index (qubit 0 = least significant bit); a graph is a symmetric
zero-diagonal uint8 adjacency matrix; ``u[k]`` is the 2x2 unitary measured on
qubit k, so an output amplitude is

    <x| (x)_k u[k] |G> = 2^{-n/2} sum_y prod_k u[k][x_k, y_k] (-1)^{q(y)},
    q(y) = sum_{a<b} G_ab y_a y_b                    (scan_plan.py:6-9).

Everything here is vectorised NumPy, exponential in the qubit count and meant
for n <= ~24 (tests) or as the CPU baseline at larger n.
"""
import numpy as np

HADAMARD = np.array([[1, 1], [1, -1]], np.complex128) / np.sqrt(2.0)


# --------------------------------------------------------------------------
# bit helpers

def parity64(x):
    """Parity of each uint64 (0/1 as uint64)."""
    return (np.bitwise_count(np.asarray(x, np.uint64)) & np.uint8(1)).astype(np.uint64)


def mask_of(bits):
    """Python int with bit j set for every j in ``bits``."""
    out = 0
    for j in bits:
        out |= 1 << int(j)
    return out


def upper_masks(g):
    """Row masks of the strict upper triangle: bit b of masks[a] is G_ab, b > a.

    Restates qmasks (distributed_sampler.py:57) and the mask built inside
    quadratic_phase (graph_sampler.py:79).  Python ints, so n may exceed 64.
    """
    g = np.asarray(g)
    return [mask_of(np.flatnonzero(g[a, a + 1:]) + a + 1) for a in range(len(g))]


def row_masks(matrix):
    """Python-int mask of each row of a 0/1 matrix (fourier_covariant_sampler.py:39-40)."""
    return [mask_of(np.flatnonzero(row)) for row in np.asarray(matrix)]


# --------------------------------------------------------------------------
# phase, gates, transforms

def quadratic_parity(ids, g):
    """q(x) mod 2 for uint64 indices x (graph_sampler.py:76-85).

    q(x) = sum_a x_a * popcount(x & uppermask_a)  (mod 2).
    """
    ids = np.asarray(ids, np.uint64)
    par = np.zeros_like(ids)
    for a, mask in enumerate(upper_masks(g)):
        if mask:
            par ^= ((ids >> np.uint64(a)) & np.uint64(1)) & parity64(ids & np.uint64(mask))
    return par


def quadratic_phase(v, g):
    """In place v *= (-1)^{q(index)} (graph_sampler.py:76-85)."""
    ids = np.arange(len(v), dtype=np.uint64)
    v *= 1.0 - 2.0 * quadratic_parity(ids, g).astype(np.float64)


def apply_one(v, u, q):
    """In place 2x2 gate u on qubit q (graph_sampler.py:71-74)."""
    blk = v.reshape(-1, 2, 1 << q)
    a = blk[:, 0, :].copy()
    b = blk[:, 1, :].copy()
    blk[:, 0, :] = u[0, 0] * a + u[0, 1] * b
    blk[:, 1, :] = u[1, 0] * a + u[1, 1] * b


def walsh(v):
    """In place unnormalised Walsh-Hadamard transform over every bit (graph_sampler.py:90-93)."""
    nbits = int(np.log2(len(v)))
    for q in range(nbits):
        blk = v.reshape(-1, 2, 1 << q)
        a = blk[:, 0, :].copy()
        b = blk[:, 1, :].copy()
        blk[:, 0, :] = a + b
        blk[:, 1, :] = a - b


def graph_state(g):
    """|G> = prod CZ H^n |0> in the computational basis (graph_sampler.py:87-88)."""
    n = len(g)
    v = np.full(1 << n, 2.0 ** (-n / 2), np.complex128)
    quadratic_phase(v, g)
    return v


def full_state(g, u):
    """(x)_k u[k] |G> (graph_sampler.py:150-151)."""
    v = graph_state(g)
    for q in range(len(g)):
        apply_one(v, u[q], q)
    return v


def full_state_threads(g, u, threads):
    """``full_state`` with its elementwise passes split over ``threads``
    threads (NumPy releases the interpreter lock inside ufuncs): each pass
    works on disjoint slices, in the same arithmetic, so the result equals
    ``full_state`` bit for bit.  A CPU baseline with every host core busy."""
    from concurrent.futures import ThreadPoolExecutor
    n = len(g)
    v = np.empty(1 << n, np.complex128)
    chunk = max(1, (1 << n) // threads)
    spans = [(a, min(a + chunk, 1 << n)) for a in range(0, 1 << n, chunk)]
    masks = upper_masks(g)
    with ThreadPoolExecutor(threads) as pool:
        def init(span):
            a, b = span
            ids = np.arange(a, b, dtype=np.uint64)
            par = np.zeros_like(ids)
            for k, mask in enumerate(masks):
                if mask:
                    par ^= ((ids >> np.uint64(k)) & np.uint64(1)) & parity64(ids & np.uint64(mask))
            v[a:b] = 2.0 ** (-n / 2) * (1.0 - 2.0 * par.astype(np.float64))
        list(pool.map(init, spans))
        for q in range(n):
            blk = v.reshape(-1, 2, 1 << q)
            outer, inner = blk.shape[0], blk.shape[2]
            uq = u[q]

            def gate(k):
                if outer >= threads:
                    sl = blk[k * outer // threads:(k + 1) * outer // threads]
                else:
                    sl = blk[:, :, k * inner // threads:(k + 1) * inner // threads]
                a = sl[:, 0, :].copy()
                b = sl[:, 1, :].copy()
                sl[:, 0, :] = uq[0, 0] * a + uq[0, 1] * b
                sl[:, 1, :] = uq[1, 0] * a + uq[1, 1] * b
            list(pool.map(gate, range(threads)))
    return v


def dual_pipeline(g, u):
    """The same state through init -> Walsh -> product, the slab pipeline.

    (x)u |G> = (x)(u_k H) . H^n |G>  and  H^n |G> = 2^{-n/2} WHT(|G>), so the
    device evaluates g(y) = 2^{-n} (-1)^{q(y)}, an unnormalised WHT over all
    bits, then the gates V_k = u_k H.  Every partial Walsh sum of (-1)^{q} is
    0 or +-2^j, so in binary floating point the init and the transform are
    exact; only the gates round.
    """
    n = len(g)
    v = np.full(1 << n, 2.0 ** (-n), np.complex128)
    quadratic_phase(v, g)
    walsh(v)
    for q in range(n):
        apply_one(v, u[q] @ HADAMARD, q)
    return v


# --------------------------------------------------------------------------
# GF(2) linear algebra

def gf2_rank(a):
    """Rank over GF(2) by row elimination (graph_sampler.py:59-69)."""
    a = np.array(a, np.uint8) & 1
    rank = 0
    rows, cols = a.shape
    for c in range(cols):
        piv = np.flatnonzero(a[rank:, c])
        if not len(piv):
            continue
        p = rank + int(piv[0])
        a[[rank, p]] = a[[p, rank]]
        below = np.flatnonzero(a[rank + 1:, c]) + rank + 1
        a[below] ^= a[rank]
        rank += 1
        if rank == rows:
            break
    return rank


def gf2_inverse(m):
    """Inverse over GF(2) by Gauss-Jordan (fourier_covariant_sampler.py:23-36).

    Raises ValueError when singular.
    """
    m = np.array(m, np.uint8) & 1
    n = len(m)
    if m.shape != (n, n):
        raise ValueError('square matrix required')
    aug = np.concatenate([m, np.eye(n, dtype=np.uint8)], axis=1)
    for c in range(n):
        piv = np.flatnonzero(aug[c:, c])
        if not len(piv):
            raise ValueError('singular GF(2) matrix')
        p = c + int(piv[0])
        aug[[c, p]] = aug[[p, c]]
        others = np.flatnonzero(aug[:, c])
        others = others[others != c]
        aug[others] ^= aug[c]
    inv = aug[:, n:].copy()
    if not np.array_equal((m.astype(np.int64) @ inv) & 1, np.eye(n, dtype=np.int64)):
        raise AssertionError('GF(2) inverse failed')
    return inv


def map_indices(ids, matrix):
    """out bit j = parity(ids & row_j): the GF(2) product matrix . id
    (fourier_covariant_sampler.py:49-53)."""
    ids = np.asarray(ids, np.uint64)
    out = np.zeros_like(ids)
    for j, mask in enumerate(row_masks(matrix)):
        out |= parity64(ids & np.uint64(mask)) << np.uint64(j)
    return out


# --------------------------------------------------------------------------
# conditional states (balanced cut A|B, M = G[B, A] of full rank)

def conditional(g, u, A, B, outB):
    """Original-method conditional A state and B marginal (graph_sampler.py:95-109).

    Returns (unnormalised psi_A complex128[2^|A|], its squared norm = P(x_B)).
    """
    g = np.asarray(g)
    ma, mb = len(A), len(B)
    ids = np.arange(1 << mb, dtype=np.uint64)
    f = np.ones(1 << mb, np.complex128)
    for i, b in enumerate(B):
        f *= u[b, outB[i], ((ids >> np.uint64(i)) & np.uint64(1)).astype(np.int64)]
    quadratic_phase(f, g[np.ix_(B, B)])
    walsh(f)
    ida = np.arange(1 << ma, dtype=np.uint64)
    mapped = np.zeros_like(ida)
    for i, a in enumerate(A):
        col = mask_of(j for j, b in enumerate(B) if g[a, b])
        mapped ^= ((ida >> np.uint64(i)) & np.uint64(1)) * np.uint64(col)
    va = f[mapped.astype(np.int64)] / (2.0 ** (len(g) / 2))
    quadratic_phase(va, g[np.ix_(A, A)])
    for i, a in enumerate(A):
        apply_one(va, u[a], i)
    return va, float(np.vdot(va, va).real)


def covariant_init(g, u, A, B, outB, ids=None):
    """Fourier-covariant init g(z) = 2^{-m/2} f(N z), N = M^{-T}
    (fourier_covariant_sampler.py:64-70)."""
    g = np.asarray(g)
    m = len(A)
    N = gf2_inverse(g[np.ix_(B, A)]).T
    z = np.arange(1 << m, dtype=np.uint64) if ids is None else np.asarray(ids, np.uint64)
    b = map_indices(z, N)
    v = np.full(len(z), 2.0 ** (-m / 2), np.complex128)
    for j, site in enumerate(B):
        v *= u[site, outB[j], ((b >> np.uint64(j)) & np.uint64(1)).astype(np.int64)]
    v *= 1.0 - 2.0 * quadratic_parity(b, g[np.ix_(B, B)]).astype(np.float64)
    return v


def covariant(g, u, A, B, outB):
    """Normalised conditional state by Fourier covariance (fourier_covariant_sampler.py:64-74)."""
    g = np.asarray(g)
    v = covariant_init(g, u, A, B, outB)
    walsh(v)
    quadratic_phase(v, g[np.ix_(A, A)])
    for j, site in enumerate(A):
        apply_one(v, u[site], j)
    return v


def deferred(g, u, A, B, outB):
    """Deferred-top-bit conditional state (deferred_global_sampler.py:93-107).

    Returns (raw, final): raw lacks the terminal gate u[A[-1]] on bit m-1,
    final has it.  The top Walsh bit is folded into the init.
    """
    g = np.asarray(g)
    m = len(A)
    half = np.arange(1 << (m - 1), dtype=np.uint64)
    lo = covariant_init(g, u, A, B, outB, half)
    hi = covariant_init(g, u, A, B, outB, half | np.uint64(1 << (m - 1)))
    raw = np.concatenate([lo + hi, lo - hi])
    for part in (raw[:len(half)], raw[len(half):]):
        walsh(part)
    quadratic_phase(raw, g[np.ix_(A, A)])
    for j, site in enumerate(A[:-1]):
        apply_one(raw, u[site], j)
    final = raw.copy()
    apply_one(final, u[A[-1]], m - 1)
    return raw, final


# --------------------------------------------------------------------------
# chunked lookup tables

def transformed_quadratic(gB, N):
    """q_B(N z) = sum_{i<j} Q_ij z_i z_j + ell.z over GF(2) (init_specialization.py:53-59).

    Returns (Q strict-upper uint8, ell uint8 vector).
    """
    up = np.triu(np.asarray(gB, np.int64), 1)
    Nn = np.asarray(N, np.int64)
    P = (Nn.T @ up @ Nn) & 1
    ell = np.diag(P).astype(np.uint8)
    Q = np.triu((P ^ P.T) & 1, 1).astype(np.uint8)
    return Q, ell


def phase_map_tables(m, N, Q, ell, chunk_bits=8):
    """Chunk tables that map z -> b = N z and give the parity of q_B(N z).

    Restates build_specialization's LUT (init_specialization.py:160-189): for
    chunk c and code k (bits of z in [c*w, c*w+w)), entry[0] = XOR of the
    columns of N selected by the code, entry[1] = (cross mask onto higher z
    bits) | (parity of the in-chunk quadratic + linear part) << 63.  A caller
    XORs entry[0] over chunks into b and ANDs entry[1] with (z | 1<<63),
    XORing into one word whose parity is q_B(N z).  Requires m <= 63.
    """
    if not 1 <= m <= 63:
        raise ValueError('1..63 bits')
    chunks = -(-m // chunk_bits)
    cols = row_masks(np.asarray(N).T)          # column j of N as a mask over b bits
    rows = row_masks(Q)                        # strict-upper rows of Q
    lin = mask_of(np.flatnonzero(ell))
    lut = np.zeros((chunks, 1 << chunk_bits, 2), np.uint64)
    full = (1 << m) - 1
    for c in range(chunks):
        lo, hi = c * chunk_bits, min(m, (c + 1) * chunk_bits)
        higher = full ^ ((1 << hi) - 1)
        for code in range(1 << (hi - lo)):
            z = code << lo
            mapped, cross = 0, 0
            inner = bin(z & lin).count('1') & 1
            for j in range(lo, hi):
                if (z >> j) & 1:
                    mapped ^= cols[j]
                    cross ^= rows[j] & higher
                    inner ^= bin(z & rows[j]).count('1') & 1
            lut[c, code, 0] = mapped
            lut[c, code, 1] = cross | (inner << 63)
    return lut


def lookup_map_phase(lut, ids, chunk_bits=8):
    """Evaluate phase_map_tables: returns (b uint64, parity uint64)."""
    ids = np.asarray(ids, np.uint64)
    b = np.zeros_like(ids)
    acc = np.zeros_like(ids)
    marked = ids | np.uint64(1 << 63)
    mask = np.uint64((1 << chunk_bits) - 1)
    for c in range(lut.shape[0]):
        code = ((ids >> np.uint64(c * chunk_bits)) & mask).astype(np.int64)
        b ^= lut[c, code, 0]
        acc ^= lut[c, code, 1] & marked
    return b, parity64(acc)


def weight_tables(u, B, outB, scale, chunk_bits=8):
    """Per-chunk products for arbitrary per-site unitaries (no reference
    counterpart: generalises the identical-U popcount table of
    init_specialization.py:98-115 to the Haar-random gates of the synthetic
    workloads).

    table[c][k] = prod_{j in chunk c} u[B[j]][outB[j], bit (j - c*w) of k],
    with ``scale`` folded into chunk 0; prod_c table[c][chunk_c(b)] is the
    covariant init weight of b.  Built in complex128.
    """
    m = len(B)
    chunks = -(-m // chunk_bits)
    tab = np.ones((chunks, 1 << chunk_bits), np.complex128)
    codes = np.arange(1 << chunk_bits)
    for c in range(chunks):
        for t in range(chunk_bits):
            j = c * chunk_bits + t
            if j >= m:
                break
            tab[c] *= u[B[j]][outB[j], (codes >> t) & 1]
    tab[0] *= scale
    return tab


def popcount_weight_table(u0, outB, scale):
    """The covariant init weight for identical B unitaries u0, as a table in
    two popcounts (init_specialization.py:98-115): with o = outB and
    r = popcount(o), table[k, h] = scale u00^n00 u01^n01 u10^n10 u11^n11 with
    n11 = h, n01 = k - h, n10 = r - h, n00 = m - r - k + h where those counts
    are valid, else 0; prod_j u0[o_j, b_j] = table[popc(b), popc(b & o)].
    Built in complex128 (the reference rounds it once to complex64).
    """
    m = len(outB)
    r = int(np.sum(outB))
    p = np.stack([np.power(complex(a), np.arange(m + 1)) for a in np.asarray(u0, np.complex128).flat])
    tab = np.zeros((m + 1, m + 1), np.complex128)
    for k in range(m + 1):
        for h in range(max(0, r + k - m), min(r, k) + 1):
            tab[k, h] = scale * p[0, m - r - k + h] * p[1, k - h] * p[2, r - h] * p[3, h]
    return tab


def graph_phase_tables(g, chunk_bits=8):
    """Chunk tables for the plain graph phase (-1)^{q(x)}: the covariant
    tables with N = identity, Q = upper triangle of G, ell = 0."""
    n = len(g)
    return phase_map_tables(n, np.eye(n, dtype=np.uint8), np.triu(np.asarray(g, np.uint8), 1),
                            np.zeros(n, np.uint8), chunk_bits)
