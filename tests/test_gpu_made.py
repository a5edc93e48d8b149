"""MADE kernels on the device against the float64 oracle.

GEMM: products of BF16 inputs are exact in FP32, so results agree with a
float64 matmul of the same BF16 values to FP32 accumulation error (1e-5
relative).  Gradients: the device rounds activations and backpropagated
signals to BF16, so each gradient tensor must match, in relative L2, the
float64 oracle that rounds at the same points within 2e-3, and the plain
float64 oracle (evaluated on the device's BF16 weights) within 0.1.
"""
import numpy as np
import pytest

from oracle import made as om

pytestmark = pytest.mark.gpu
torch = pytest.importorskip('torch')


def bf16_round(a):
    a = np.asarray(a, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint32) << 16
    return u.view(np.float32).astype(np.float64)


def to_bf16_dev(a):
    a = np.asarray(a, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    h = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)
    return torch.as_tensor(h.view(np.int16)).cuda()


@pytest.mark.parametrize('ta,tb', [(0, 0), (0, 1), (1, 0), (1, 1)])
def test_gemm_layouts_and_epilogues(ta, tb):
    from rgs70_classical_challenge_amd import made  # noqa: F401  (registers the entry points)
    from rgs70_classical_challenge_amd._lib import call, ptr
    rng = np.random.default_rng(ta * 2 + tb)
    M, N, K = 37, 70, 45
    A = rng.normal(size=(K, M) if ta else (M, K))
    B = rng.normal(size=(N, K) if tb else (K, N))
    bias = rng.normal(size=N).astype(np.float32)
    dm = rng.normal(size=(M, N))
    Ad, Bd, dmd = to_bf16_dev(A), to_bf16_dev(B), to_bf16_dev(dm)
    C = torch.zeros((M, N), dtype=torch.float32, device='cuda')
    Cb = torch.zeros((M, N), dtype=torch.int16, device='cuda')
    bd = torch.as_tensor(bias).cuda()
    s = torch.cuda.current_stream().cuda_stream
    print(f'case ta={ta} tb={tb}', flush=True)
    call('rgs_gemm_bf16', ta, tb, M, N, K, ptr(Ad), A.shape[1], ptr(Bd), B.shape[1], 0.5, ptr(C), N, ptr(Cb), N,
         ptr(bd), 1, ptr(dmd), N, 0, s)
    torch.cuda.synchronize()
    Ar, Br = bf16_round(A), bf16_round(B)
    ref = 0.5 * ((Ar.T if ta else Ar) @ (Br.T if tb else Br)) + bias
    ref = np.maximum(ref, 0) * (bf16_round(dm) > 0)
    got = C.cpu().numpy()
    assert np.max(np.abs(got - ref)) < 1e-5 * np.max(np.abs(ref)) + 1e-6
    gotb = (Cb.cpu().numpy().view(np.uint16).astype(np.uint32) << 16).view(np.float32)
    assert np.array_equal(gotb, bf16_round(got).astype(np.float32))
    # accumulate mode adds into C
    call('rgs_gemm_bf16', ta, tb, M, N, K, ptr(Ad), A.shape[1], ptr(Bd), B.shape[1], 0.5, ptr(C), N, None, 0,
         ptr(bd), 1, ptr(dmd), N, 1, s)
    torch.cuda.synchronize()
    assert np.max(np.abs(C.cpu().numpy() - 2 * ref)) < 2e-5 * np.max(np.abs(ref)) + 1e-6


def test_made_gradients_match_oracle():
    from rgs70_classical_challenge_amd.made import MADE
    n, H, B = 12, 64, 48
    net = MADE(n, hidden=H, batch=B, seed=7)
    rng = np.random.default_rng(1)
    x = rng.integers(0, 2, (B, n)).astype(np.uint8)
    from rgs70_classical_challenge_amd.made import bits_to_bf16
    net.x.copy_(torch.as_tensor(bits_to_bf16(x)))
    net.forward()
    net.loss()
    net.backward()
    torch.cuda.synchronize()
    # the GEMMs read BF16 weights; biases are added from the FP32 master copy
    wq, wm = net.bf16_params_numpy(), net.params_numpy()
    params = [wq[0], wm[1], wq[2], wm[3], wq[4], wm[5]]
    masks = om.made_masks(n, H, 7)
    xf = x.astype(np.float64)
    # (1) the oracle rounding activations and signals to BF16 where the device
    # stores them: agreement to FP32 accumulation order (plus rare one-ulp
    # BF16 rounding differences at tie boundaries)
    loss_q, grads_q = om.loss_and_grads(params, masks, xf, store=om.bf16_round)
    # (2) the plain float64 oracle: BF16 storage is the only difference, and
    # the batch sums cancel, so only a loose bound holds
    loss, grads = om.loss_and_grads(params, masks, xf)
    assert float(net.nll.mean().item()) == pytest.approx(loss_q, rel=1e-4)
    assert float(net.nll.mean().item()) == pytest.approx(loss, rel=1e-2)
    mask_all = net.mask.cpu().numpy()
    for k, g in enumerate(net.grads_numpy()):
        # the device leaves masked-out weight gradients unmasked; Adam applies the mask
        mk = mask_all[net.offsets[k]:net.offsets[k + 1]].reshape(net.shapes[k])
        err_q = np.linalg.norm(g * mk - grads_q[k]) / max(np.linalg.norm(grads_q[k]), 1e-12)
        err = np.linalg.norm(g * mk - grads[k]) / max(np.linalg.norm(grads[k]), 1e-12)
        assert err_q < 2e-3, (k, err_q)
        assert err < 0.1, (k, err)
    # one Adam step against the oracle on the FP32 master weights
    w0 = net.params_numpy()
    g0 = net.grads_numpy()
    net.adam()
    torch.cuda.synchronize()
    w1 = net.params_numpy()
    for k in range(6):
        mk = net.mask.cpu().numpy()[net.offsets[k]:net.offsets[k + 1]].reshape(net.shapes[k])
        exp, _, _ = om.adam_step(w0[k], g0[k] * mk, 0.0, 0.0, 1)
        assert np.allclose(w1[k], exp * mk, rtol=1e-5, atol=1e-7)


def test_training_learns_and_sampling_matches_q():
    """n=8: train on a skewed distribution, then draw from q and compare the
    empirical frequencies with the oracle's exact q (enumerated)."""
    from rgs70_classical_challenge_amd.made import MADE
    n = 8
    rng = np.random.default_rng(3)
    logits = rng.normal(size=1 << n) * 1.5
    p = np.exp(logits) / np.exp(logits).sum()
    data_idx = rng.choice(1 << n, size=20000, p=p)
    data = ((data_idx[:, None] >> np.arange(n)) & 1).astype(np.uint8)
    net = MADE(n, hidden=64, batch=256, seed=2, lr=3e-3)
    losses = net.train(data, 600, log_every=100)
    assert losses[-1] < losses[0] - 0.2, losses
    all_bits = ((np.arange(1 << n)[:, None] >> np.arange(n)) & 1).astype(np.uint8)
    logq = net.log_q(all_bits)
    q = np.exp(logq)
    assert abs(q.sum() - 1) < 2e-3
    kl = float(np.sum(p * (np.log(p) - logq)))
    assert kl < 0.5 * float(np.sum(p * np.log(p * (1 << n))))   # better than uniform by half
    bits, lq = net.sample(rng.random((20000, n)))
    idx = (bits.astype(np.int64) << np.arange(n)).sum(1)
    assert np.allclose(lq, logq[idx], atol=1e-4)
    counts = np.bincount(idx, minlength=1 << n)
    expect = q * len(idx)
    big = expect > 20
    chi2 = ((counts[big] - expect[big]) ** 2 / expect[big]).sum()
    assert chi2 < big.sum() + 6 * np.sqrt(2 * big.sum())


def test_incremental_sampler_equals_full_forward():
    """The incremental sampler (new units of degree o only, logit o only)
    reproduces the column-by-column full-forward sampler bit for bit, and its
    log q equals the full forward's log q of the drawn strings."""
    from rgs70_classical_challenge_amd.made import MADE
    n, H = 20, 128
    net = MADE(n, hidden=H, batch=256, seed=3)
    rng = np.random.default_rng(11)
    data = (rng.random((4096, n)) < np.linspace(0.2, 0.8, n)).astype(np.uint8)
    data[:, 5] = data[:, 2] ^ data[:, 3]                       # some structure to learn
    net.train(data, 200, seed=1)
    u = rng.random((3000, n))
    bits, logq = net.sample(u, batch=1024)                     # three partial batches
    bits_full, logq_full = net.sample_full(u)
    assert np.array_equal(bits, bits_full)
    ref = net.log_q(bits)
    assert np.max(np.abs(logq - ref)) < 1e-4 * n, np.max(np.abs(logq - ref))
    assert np.max(np.abs(logq - logq_full)) < 1e-4 * n
    # a decided bit follows its own logit: log q never exceeds 0
    assert np.all(logq <= 1e-12)


def test_fused_sampler_equals_incremental_sequence():
    """One-launch sampler against the launch-per-step incremental sampler:
    identical bits and identical log q (same arithmetic, same order).  Widths
    of whole 16-unit blocks per degree (H = 16 m (n - 1), m = 1 and 2, or
    H = 16 k with k < n - 1) take the one-wave kernel; a width without such
    blocks (H = 160 at n = 9: 20 units per degree) is refused."""
    from rgs70_classical_challenge_amd.made import MADE
    from rgs70_classical_challenge_amd.rejection import pack_words_device
    for n, H, count in ((20, 128, 3000), (34, 512, 1000), (9, 64, 77),
                        (34, 528, 1000), (9, 128, 77), (12, 352, 500), (20, 304, 3000)):
        net = MADE(n, hidden=H, batch=256, seed=5)
        rng = np.random.default_rng(n)
        data = (rng.random((4096, n)) < np.linspace(0.15, 0.85, n)).astype(np.uint8)
        net.train(data, 150, seed=3)
        sb = net._sampling_buffers(count)
        sb['u'][:count].copy_(torch.as_tensor(rng.random((count, n))))
        x_inc, lq_inc = net.sample_device(count, graph=False)
        words_inc = pack_words_device(x_inc, n).clone()
        lq_inc = lq_inc.clone()
        x_inc = x_inc.clone()
        words, lq = net.sample_fused(count, bits=True)
        torch.cuda.synchronize()
        print(f'case n={n} H={H} count={count}', flush=True)
        assert torch.equal(words, words_inc)
        assert torch.equal(sb['x'][:count], x_inc)
        assert torch.equal(lq, lq_inc)
        # the same draws from the row-major weights read in place
        words_raw, lq_raw = net.sample_fused_raw(count)
        assert torch.equal(words_raw, words_inc) and torch.equal(lq_raw, lq_inc)
    net = MADE(9, hidden=160, batch=64, seed=5)
    net._sampling_buffers(64)
    with pytest.raises(ValueError, match='aligned'):
        net.sample_fused(64)
    with pytest.raises(ValueError, match='aligned 16-unit blocks'):
        net.sample_fused_raw(64)
