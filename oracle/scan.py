"""Linear rank-width scan: plan compiler and NumPy executor (complex128).

Restates the algorithm of the vendored dcs-rankwidth scan
(work/rgs_attack/upstream_scan/scan_plan.py:10-23, make_plan :58-105,
future_masks :45-55, byte_tables :133-145, kernel_combos :148-152,
sweep_reference :161-195) and the scorer's amplitude
(work/rgs_attack/scan_oracle.py:25-29).  Third-party provenance: Kim Renaud's
dcs-rankwidth at commit 2ddc1c94c31b922f0eb3118d0a8356cc92e6f1a7 (reference
NOTICE), vendored in the reference, read as text only.

Vertices are summed in a fixed order.  After position j the environment E is
indexed by the coordinates c of the syndrome s = o + sum_i c_i basis_i (the
parities of processed neighbours of every later position), so |E| = 2^{r_j}
with r_j the GF(2) rank of the cut G[order[:j+1], order[j+1:]].  Summing the
vertex at position j with value y multiplies by f(y) (-1)^{y s_j} and moves
to the next syndrome space; in table form

    E'[c'] = sum_{p in L c' + span K} E[p_c] f(p_y) (-1)^{p_y (ov + p_c.beta)}

with p = (c, y) packed as c | y << r_in, L a right inverse and K a kernel
basis of the linear map p -> c'.  Only free vertices are supported (no
structurally fixed vertex: none exist for QR-drawn unitaries or for RGS70,
order_optimization/prepare_and_verify.py:46).

The bases chosen here are this build's own (pivot = highest set bit); any
valid basis gives the same amplitude and the same rank profile.
"""
import math
import numpy as np


class Step:
    __slots__ = ('v', 'r_in', 'r_out', 'beta', 'lcols', 'kernel')


def future_masks(n, g, order):
    """fut[j] = bitmask over positions of the later neighbours of order[j]."""
    pos = np.empty(n, np.int64)
    pos[np.asarray(order)] = np.arange(n)
    fut = [0] * n
    gi = np.asarray(g)
    for j, v in enumerate(order):
        for w in np.flatnonzero(gi[v]):
            if pos[w] > j:
                fut[j] |= 1 << int(pos[w])
    return fut


def _echelon_insert(vec, basis, pivots):
    """Reduce ``vec`` by the basis; return (residual, coordinate mask)."""
    coord = 0
    for k, (b, p) in enumerate(zip(basis, pivots)):
        if (vec >> p) & 1:
            vec ^= b
            coord |= 1 << k
    return vec, coord


def compile_plan(g, order):
    """Returns (steps, profile, fut) for an ordering of all vertices."""
    n = len(g)
    fut = future_masks(n, g, order)
    basis = []
    steps, profile = [], []
    for j in range(n):
        st = Step()
        st.v = int(order[j])
        st.r_in = len(basis)
        st.beta = 0
        for i, b in enumerate(basis):
            st.beta |= ((b >> j) & 1) << i
        gens = [b & ~(1 << j) for b in basis] + [fut[j]]
        nb, piv, coords, creator = [], [], [], {}
        for i, w in enumerate(gens):
            res, coord = _echelon_insert(w, nb, piv)
            if res:
                piv.append(res.bit_length() - 1)
                nb.append(res)
                coord |= 1 << (len(nb) - 1)
                creator[len(nb) - 1] = i
            coords.append(coord)
        r_out = len(nb)
        lcols = [0] * r_out
        for k in range(r_out):
            i = creator[k]
            p = 1 << i
            rest = coords[i] & ~(1 << k)
            for l in range(k):
                if (rest >> l) & 1:
                    p ^= lcols[l]
            lcols[k] = p
        pivot_gens = set(creator.values())
        kern = []
        for i, cmask in enumerate(coords):
            if i in pivot_gens:
                continue
            kv = 1 << i
            for k in range(r_out):
                if (cmask >> k) & 1:
                    kv ^= lcols[k]
            kern.append(kv)
        st.r_out = r_out
        st.lcols = lcols
        st.kernel = kern
        basis = nb
        steps.append(st)
        profile.append(r_out)
    return steps, profile, fut


def cut_rank_profile(g, order):
    """Closed form: rank over GF(2) of G[order[:k], order[k:]] for k = 1..n
    (order_optimization/prepare_and_verify.py:31-32)."""
    from .graph_state import gf2_rank
    g = np.asarray(g)
    order = list(order)
    out = []
    for k in range(1, len(g) + 1):
        left, right = order[:k], order[k:]
        out.append(gf2_rank(g[np.ix_(left, right)]) if right else 0)
    return out


def byte_tables(cols, nbits):
    """T[b][val] = XOR of cols[8b + t] over set bits t of val (scan_plan.py:133-145)."""
    nb = max(1, (nbits + 7) // 8)
    T = np.zeros((nb, 256), np.uint64)
    for b in range(nb):
        for val in range(256):
            acc = 0
            for t in range(8):
                i = 8 * b + t
                if i < nbits and (val >> t) & 1:
                    acc ^= cols[i]
            T[b, val] = acc
    return T


def kernel_combos(kernel):
    """All XOR combinations of the kernel basis (scan_plan.py:148-152)."""
    combos = [0]
    for kv in kernel:
        combos = combos + [c ^ kv for c in combos]
    return combos


def sweep(steps, fvec):
    """Execute a plan in complex128 with max-modulus renormalisation
    (sweep_reference, scan_plan.py:161-195).  fvec[v] = (f0, f1).
    Returns (value, logscale): sum = value * exp(logscale)."""
    E = np.ones(1, np.complex128)
    logscale = 0.0
    for st in steps:
        f0, f1 = complex(fvec[st.v][0]), complex(fvec[st.v][1])
        T = byte_tables(st.lcols, st.r_out)
        cp = np.arange(1 << st.r_out, dtype=np.uint64)
        p0 = np.zeros_like(cp)
        for b in range(T.shape[0]):
            p0 ^= T[b][((cp >> np.uint64(8 * b)) & np.uint64(255)).astype(np.int64)]
        mask = np.uint64((1 << st.r_in) - 1)
        beta = np.uint64(st.beta)
        acc = np.zeros(1 << st.r_out, np.complex128)
        for kc in kernel_combos(st.kernel):
            p = p0 ^ np.uint64(kc)
            c = p & mask
            y = (p >> np.uint64(st.r_in)) & np.uint64(1)
            par = (np.bitwise_count(c & beta) & np.uint8(1)).astype(np.uint64)
            sgn = 1.0 - 2.0 * (y & par).astype(np.float64)
            w = np.where(y == 1, f1, f0)
            acc += E[c.astype(np.int64)] * w * sgn
        E = acc
        mx = float(np.max(np.abs(E))) if E.size else 0.0
        if mx > 0:
            E = E / mx
            logscale += math.log(mx)
    return complex(E[0]), logscale


def amplitude(steps, u, bits):
    """<x| (x)u |G> for one bitstring (qubit 0 first), ScanOracle.amplitude."""
    n = len(bits)
    fvec = {i: u[i, int(bits[i]), :] for i in range(n)}
    val, ls = sweep(steps, fvec)
    return val * math.exp(ls) * 2.0 ** (-n / 2)


def sweep_batch(steps, f0, f1):
    """``sweep`` for many proposals at once: f0[b][v], f1[b][v] are the two
    weights of vertex v for proposal b (complex arrays [B][n]).  The same
    table walk and the same per-proposal max-modulus renormalisation as
    ``sweep``, vectorised over proposals.  Returns (values [B], logscale [B])."""
    f0 = np.asarray(f0, np.complex128)
    f1 = np.asarray(f1, np.complex128)
    nb = f0.shape[0]
    E = np.ones((nb, 1), np.complex128)
    logscale = np.zeros(nb)
    for st in steps:
        T = byte_tables(st.lcols, st.r_out)
        cp = np.arange(1 << st.r_out, dtype=np.uint64)
        p0 = np.zeros_like(cp)
        for b in range(T.shape[0]):
            p0 ^= T[b][((cp >> np.uint64(8 * b)) & np.uint64(255)).astype(np.int64)]
        mask = np.uint64((1 << st.r_in) - 1)
        beta = np.uint64(st.beta)
        acc = np.zeros((nb, 1 << st.r_out), np.complex128)
        w0 = f0[:, st.v][:, None]
        w1 = f1[:, st.v][:, None]
        for kc in kernel_combos(st.kernel):
            p = p0 ^ np.uint64(kc)
            c = (p & mask).astype(np.int64)
            y = ((p >> np.uint64(st.r_in)) & np.uint64(1)).astype(bool)
            par = (np.bitwise_count(p & mask & beta) & np.uint8(1)).astype(bool)
            sgn = np.where(y & par, -1.0, 1.0)
            acc += E[:, c] * np.where(y[None, :], w1, w0) * sgn[None, :]
        E = acc
        mx = np.max(np.abs(E), axis=1)
        pos = mx > 0
        E[pos] /= mx[pos, None]
        logscale[pos] += np.log(mx[pos])
    return E[:, 0].copy(), logscale


def amplitudes(steps, u, bits):
    """<x| (x)u |G> for each row of ``bits`` ([B][n], qubit 0 first)."""
    bits = np.asarray(bits, np.int64)
    n = bits.shape[1]
    q = np.arange(n)
    f0 = np.asarray(u)[q[None, :], bits, 0]
    f1 = np.asarray(u)[q[None, :], bits, 1]
    val, ls = sweep_batch(steps, f0, f1)
    return val * np.exp(ls) * 2.0 ** (-n / 2)


def plan_cost(profile):
    return sum(1 << r for r in profile)
