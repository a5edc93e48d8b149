"""Host checks of the device rejection loop's proposal words (run with torch
on the CPU): layout of the packed uniform proposals and their bit balance."""
import numpy as np
import torch

from rgs70_classical_challenge_amd.rejection import uniform_words_device
from rgs70_classical_challenge_amd.scan import bits_to_words


def test_uniform_words_layout_and_balance():
    g = torch.Generator()
    g.manual_seed(3)
    count = 100_000
    for n in (9, 34, 64, 70, 127):
        w = uniform_words_device(count, n, g, 'cpu').numpy().view(np.uint64)
        bits = np.stack([(w[:, q // 64] >> np.uint64(q % 64)) & np.uint64(1) for q in range(n)], 1).astype(np.uint8)
        # the words are exactly bits_to_words of their own bits: nothing above bit n - 1
        assert np.array_equal(bits_to_words(bits), w)
        # each bit a fair coin (5 sigma), pairs of neighbouring bits uncorrelated
        mean = bits.mean(0)
        assert np.all(np.abs(mean - 0.5) < 5 * 0.5 / np.sqrt(count))
        pair = (bits[:, :-1] & bits[:, 1:]).mean(0)
        assert np.all(np.abs(pair - 0.25) < 5 * np.sqrt(0.25 * 0.75 / count))


def test_probabilities_from_scores_exact_power_of_two():
    """p = (re re + im im) 2^(2 ex - n): the power of two assembled from its
    exponent bits equals Python's exact 2.0 ** k, and each square and the sum
    are rounded once (numpy's elementwise operations)."""
    import torch
    from rgs70_classical_challenge_amd.scan import probabilities_from_scores
    g = torch.Generator().manual_seed(3)
    val = torch.randn((4096, 2), generator=g, dtype=torch.float64)
    ex = torch.randint(-40, 40, (4096,), generator=g, dtype=torch.int32)
    for n in (1, 34, 70, 128):
        p = probabilities_from_scores(val, ex, n).numpy()
        v = val.numpy()
        ref = (v[:, 0] * v[:, 0] + v[:, 1] * v[:, 1]) * np.array([2.0 ** k for k in (2 * ex.numpy() - n).tolist()])
        assert np.array_equal(p, ref)
