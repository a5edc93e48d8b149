"""Sampling semantics (float64 on the host).

Two draws are restated:

* ``reference_draw`` - the production draw of the reference: FP64 tile masses,
  ``searchsorted(cumsum(mass), u*norm, side='right')`` for the tile, the same
  inside the tile, then (deferred variant) the terminal 2x2 gate applied to the
  two amplitudes at the chosen local index in complex128 and a third uniform
  for the top bit (distributed_sampler.py:145-157,
  deferred_global_sampler.py:65-78).  Exactly 3 uniforms per sample in the
  deferred variant (born_production.py:47).

* ``pyramid`` / ``tree_draw`` - the device draw of this build: a perfect
  binary tree of FP64 pairwise sums over p(x) = re^2 + im^2 (level 0 = leaves),
  descended from the root with one uniform per sample, qubit n-1 first.  With
  the same leaf values and the same pairwise association, host and device
  masses are bit-identical, so the draws are identical; against another
  state (the complex128 oracle) they agree wherever the uniform is farther
  from a node boundary than the L1 distance of the two distributions allows
  (``tree_draw`` reports the margin).
"""
import numpy as np


def probabilities(v):
    """p = re^2 + im^2 in float64 (re^2, im^2 are exact for complex64 inputs)."""
    v = np.asarray(v)
    re = v.real.astype(np.float64)
    im = v.imag.astype(np.float64)
    return re * re + im * im


def tile_masses(p, tilebits):
    """FP64 mass per 2^tilebits tile (the blockmass kernel, distributed_sampler.py:50-54)."""
    t = 1 << tilebits
    n = len(p)
    ntile = (n + t - 1) // t
    out = np.zeros(ntile, np.float64)
    for k in range(ntile):
        out[k] = p[k * t:(k + 1) * t].sum()
    return out


def reference_draw(raw, rng, tilebits=20, terminal=None):
    """Restates DistributedSampler.draw (+ DeferredGlobalMixin.draw when
    ``terminal`` is a 2x2 matrix) on one array holding both halves.

    ``raw``: complex state of 2^m amplitudes (before the terminal gate when
    deferred).  The halves on bit m-1 play the reference's two shards.
    Returns (index, p_at_index, norm).
    """
    raw = np.asarray(raw)
    size = len(raw)
    half = size // 2
    shards = [raw[:half], raw[half:]]
    tb = min(tilebits, int(np.log2(half)))
    nmass = (half + (1 << tb) - 1) >> tb
    mass = np.concatenate([tile_masses(probabilities(s), tb) for s in shards])
    norm = float(mass.sum())
    if not np.isfinite(norm) or abs(norm - 1) > 1e-3:
        raise ArithmeticError('Conditional norm check failed: %r' % norm)
    block = int(np.searchsorted(np.cumsum(mass), rng.random() * norm, side='right'))
    gpu, chunk = block // nmass, block % nmass
    tile = shards[gpu][chunk << tb:min((chunk + 1) << tb, half)]
    p = probabilities(tile)
    within = int(np.searchsorted(np.cumsum(p), rng.random() * p.sum(), side='right'))
    index = gpu * half + (chunk << tb) + within
    if terminal is None:
        return index, float(p[within]), norm
    local = index % half
    pair = np.asarray(terminal, np.complex128) @ np.array(
        [complex(shards[0][local]), complex(shards[1][local])], np.complex128)
    pp = np.abs(pair) ** 2
    total = float(pp.sum())
    if not np.isfinite(total) or total <= 0:
        raise ArithmeticError('Invalid transformed local mass')
    bit = int(rng.random() * total >= pp[0])
    return bit * half + local, float(pp[bit]), norm


def pyramid(p):
    """List of levels, levels[0] = p (float64), levels[k][i] = levels[k-1][2i] + levels[k-1][2i+1]."""
    levels = [np.asarray(p, np.float64)]
    while len(levels[-1]) > 1:
        cur = levels[-1]
        levels.append(cur[0::2] + cur[1::2])
    return levels


def descend(levels, t, top, idx=0):
    """Descend from node ``idx`` of level ``top`` to level 0 with target t.

    At a node with children (L, R): go right iff t >= L and R > 0, then t -= L.
    Returns (index at level 0, remaining t, smallest |t - L| met on the path).
    """
    margin = np.inf
    for lev in range(top - 1, -1, -1):
        left = levels[lev][2 * idx]
        right = levels[lev][2 * idx + 1]
        margin = min(margin, abs(t - left))
        if t >= left and right > 0:
            t -= left
            idx = 2 * idx + 1
        else:
            idx = 2 * idx
    return idx, t, margin


def tree_draw(levels, u):
    """Descend the pyramid with target t = u * root.

    Returns (index, margin) where the margin is the smallest |t - L| met on
    the path, in units of total mass (absolute).  If two mass pyramids differ
    in their leaves by delta in L1 norm, every decision agrees whenever the
    margin exceeds 3 * delta / root (t moves by at most 2 delta, L by delta).
    """
    top = len(levels) - 1
    root = levels[top][0]
    idx, _, margin = descend(levels, u * root, top)
    return idx, margin / root


def split_draws(upper, fetch_block, block_bits, uniforms):
    """The same descent when only the pyramid's block level and above
    (``upper``, upper[0] = masses of the aligned 2^block_bits blocks) are at
    hand: descend them to a block, then descend the pairwise pyramid of that
    block's leaf values, ``fetch_block(blocks)`` -> float64 [len(blocks)][2^block_bits].
    Equal to ``tree_draws`` on the full pyramid whenever upper[0] holds the
    pairwise sums of the leaves."""
    top = len(upper) - 1
    root = upper[top][0]
    firsts = [descend(upper, float(u) * root, top) for u in uniforms]
    blocks = np.array([b for b, _, _ in firsts], np.int64)
    leaves = fetch_block(blocks)
    out = np.empty(len(uniforms), np.int64)
    for k, (b, t, _) in enumerate(firsts):
        i, _, _ = descend(pyramid(leaves[k]), t, block_bits)
        out[k] = (b << block_bits) | i
    return out


def tree_draws(levels, uniforms):
    out = np.empty(len(uniforms), np.int64)
    margins = np.empty(len(uniforms), np.float64)
    for k, u in enumerate(uniforms):
        out[k], margins[k] = tree_draw(levels, float(u))
    return out, margins
