"""Seeded synthetic instances: the inputs of every workload and test.

* Random d-regular graphs by the configuration (pairing) model with whole
  restarts until the multigraph is simple.  Every draw comes from one
  ``numpy.random.Generator(PCG64(seed))`` stream, so a seed fixes the graph.
* Erdos-Renyi G(n, p) for odd n, where no 3-regular graph exists.
* Per-qubit 2x2 unitaries: Q of the QR factorisation of a complex Gaussian
  matrix, drawn with the same NumPy calls as the reference
  (work/rgs_attack/graph_sampler.py:149, scan_oracle.py:173).  No phase fix is
  applied, as in the reference, so the law is not exactly Haar; the unitaries
  are exactly unitary to rounding.
"""
import numpy as np


def random_regular_graph(n, d, rng, max_tries=10000):
    if (n * d) % 2 or d >= n:
        raise ValueError('no %d-regular graph on %d vertices' % (d, n))
    stubs = np.repeat(np.arange(n), d)
    for _ in range(max_tries):
        perm = rng.permutation(stubs)
        a, b = perm[0::2], perm[1::2]
        if np.any(a == b):
            continue
        g = np.zeros((n, n), np.uint8)
        lo, hi = np.minimum(a, b), np.maximum(a, b)
        key = lo * n + hi
        if len(np.unique(key)) != len(key):
            continue
        g[lo, hi] = 1
        g[hi, lo] = 1
        return g
    raise RuntimeError('pairing model failed')


def erdos_renyi(n, p, rng):
    upper = np.triu((rng.random((n, n)) < p).astype(np.uint8), 1)
    return upper | upper.T


def haar_unitaries(n, rng):
    out = np.empty((n, 2, 2), np.complex128)
    for q in range(n):
        z = rng.normal(size=(2, 2)) + 1j * rng.normal(size=(2, 2))
        out[q] = np.linalg.qr(z)[0]
    return out


def instance(n, seed, kind=None):
    """(graph uint8[n,n], unitaries complex128[n,2,2]) for a seed.

    kind None: 3-regular for even n, G(n, 3/n) for odd n (the uneven-batch rule).
    """
    rng = np.random.Generator(np.random.PCG64(seed))
    if kind is None:
        kind = 'regular3' if n % 2 == 0 else 'er'
    if kind == 'regular3':
        g = random_regular_graph(n, 3, rng)
    elif kind == 'er':
        g = erdos_renyi(n, 3.0 / n, rng)
    else:
        raise ValueError(kind)
    return g, haar_unitaries(n, rng)
