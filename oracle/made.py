"""Float64 MADE: masks, forward, loss, gradients and one Adam step.

No reference counterpart (the learned proposal is new in this build); the
network is Germain et al.'s MADE (arXiv:1502.03509) with two ReLU hidden
layers and sigmoid conditionals:

    h1 = relu(x W1^T + b1),  h2 = relu(h1 W2^T + b2),  l = h2 W3^T + b3,
    -log q(x) = sum_o softplus(l_o) - x_o l_o,

with W_k multiplied elementwise by the autoregressive masks.  Gradients are
the analytic backward pass; tests pin them by central finite differences.
"""
import numpy as np


def made_masks(n, hidden, seed=0):
    """Degrees: inputs 1..n in order, hidden units take degrees in 1..n-1 in
    sorted order (hidden_degrees; ``seed`` is kept for the signature, the
    assignment is deterministic); a hidden unit of degree d sees inputs of
    degree <= d and output o (degree o+1) sees hidden units of degree < o+1."""
    deg_in = np.arange(1, n + 1)
    d1 = hidden_degrees(n, hidden)
    d2 = hidden_degrees(n, hidden)
    m1 = (d1[:, None] >= deg_in[None, :]).astype(np.float64)     # [H, n]
    m2 = (d2[:, None] >= d1[None, :]).astype(np.float64)         # [H, H]
    m3 = (deg_in[:, None] > d2[None, :]).astype(np.float64)      # [n, H]
    return m1, m2, m3


def bf16_round(a):
    """Round to the nearest bfloat16 (ties to even) through float32, as a
    float64 array."""
    u = np.asarray(a, np.float32).view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint32) << 16
    return u.view(np.float32).astype(np.float64)


def _exact(a):
    return a


def hidden_degrees(n, hidden):
    """Sorted hidden-unit degrees in 1 .. n-1.  Widths that are whole 16-unit
    blocks give each degree whole blocks: m blocks per degree when the block
    count is m (n-1), one block per degree 1 .. k when there are k < n-1
    blocks.  Other widths spread the degrees evenly:
    floor(i (n-1) / hidden) + 1."""
    d = max(1, n - 1)
    units = np.arange(hidden)
    if hidden % 16 == 0 and (hidden // 16) % d == 0:
        per = 16 * ((hidden // 16) // d)
        return np.array([u // per + 1 for u in units])
    if hidden % 16 == 0 and hidden // 16 < d:
        return np.array([u // 16 + 1 for u in units])
    return (units * d) // hidden + 1


def forward(params, masks, x, store=_exact):
    """``store`` is applied where an implementation keeps an activation in a
    narrower type (the hidden layers); the identity gives plain float64."""
    W1, b1, W2, b2, W3, b3 = params
    m1, m2, m3 = masks
    a1 = x @ (W1 * m1).T + b1
    h1 = store(np.maximum(a1, 0))
    a2 = h1 @ (W2 * m2).T + b2
    h2 = store(np.maximum(a2, 0))
    logits = h2 @ (W3 * m3).T + b3
    return a1, h1, a2, h2, logits


def nll(logits, x):
    sp = np.logaddexp(0.0, logits)
    return (sp - x * logits).sum(1)


def loss_and_grads(params, masks, x, store=_exact):
    """Mean NLL over the batch and gradients (same order as params).

    With ``store`` (e.g. ``bf16_round``) the hidden activations and the
    backpropagated signals that feed matrix products are rounded as an
    implementation that keeps them in that type would; bias gradients sum the
    unrounded signals."""
    W1, b1, W2, b2, W3, b3 = params
    m1, m2, m3 = masks
    B = len(x)
    a1, h1, a2, h2, logits = forward(params, masks, x, store)
    loss = nll(logits, x).mean()
    dl = (1.0 / (1.0 + np.exp(-logits)) - x) / B
    dlq = store(dl)
    gW3 = (dlq.T @ h2) * m3
    gb3 = dl.sum(0)
    dh2 = dlq @ (W3 * m3)
    da2 = dh2 * (h2 > 0)
    da2q = store(da2)
    gW2 = (da2q.T @ h1) * m2
    gb2 = da2.sum(0)
    dh1 = da2q @ (W2 * m2)
    da1 = dh1 * (h1 > 0)
    gW1 = (store(da1).T @ x) * m1
    gb1 = da1.sum(0)
    return loss, [gW1, gb1, gW2, gb2, gW3, gb3]


def adam_step(w, g, m, v, t, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** t)
    vh = v / (1 - b2 ** t)
    return w - lr * mh / (np.sqrt(vh) + eps), m, v
