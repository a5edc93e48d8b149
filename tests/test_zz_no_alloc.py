"""The allocation-free large-rank scan (rgs_scan_large_static) against
rgs_scan_large with caller buffers: the same kernels on the same plan, so the
outputs are bit-identical.  Two calls in a row, the second on a smaller plan
(the fixed tables keep the previous call's contents), at both table
precisions; a plan above the fixed tables' rank is refused before any launch."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip('torch')


def _plan(n, seed, dense):
    from rgs70_classical_challenge_amd.instances import haar_unitaries, instance
    from rgs70_classical_challenge_amd.scan import ScanPlan
    rng = np.random.default_rng(seed)
    if dense:
        g = np.triu(rng.integers(0, 2, (n, n), dtype=np.uint8), 1)
        g = g | g.T
        u = haar_unitaries(n, rng)
    else:
        g, u = instance(n, seed)
    return ScanPlan(g, rng.permutation(n).tolist()), np.ascontiguousarray(u), rng


def _both(plan, u, bits, prec, dev):
    from rgs70_classical_challenge_amd._lib import call, ptr
    from rgs70_classical_challenge_amd.scan import bits_to_words
    blob = plan.packed_tables()
    dplan = torch.as_tensor(blob.view(np.int64)).to(dev)
    offs = np.ascontiguousarray(blob[2:2 + plan.n])
    words = bits_to_words(bits)
    dt = torch.complex64 if prec == 64 else torch.complex128
    b0 = torch.empty(1 << plan.rmax, dtype=dt, device=dev)
    b1 = torch.empty(1 << plan.rmax, dtype=dt, device=dev)
    mx = torch.zeros(plan.n, dtype=torch.int64, device=dev)
    out = {}
    for name in ('rgs_scan_large', 'rgs_scan_large_static'):
        val = torch.zeros((len(bits), 2), dtype=torch.float64, device=dev)
        ex = torch.zeros(len(bits), dtype=torch.int32, device=dev)
        for k in range(len(bits)):
            w = np.ascontiguousarray(words[k])
            if name == 'rgs_scan_large':
                call(name, ctypes.addressof(plan.steps), plan.n, ptr(dplan), ptr(offs), ptr(u), ptr(w), prec, ptr(b0),
                     ptr(b1), b0.numel(), ptr(mx), ptr(val) + 16 * k, ptr(ex) + 4 * k, None)
            else:
                call(name, ctypes.addressof(plan.steps), plan.n, ptr(dplan), ptr(offs), ptr(u), ptr(w), prec,
                     ptr(val) + 16 * k, ptr(ex) + 4 * k, None)
        torch.cuda.synchronize()
        out[name] = (val.cpu().numpy(), ex.cpu().numpy())
    return out


@pytest.mark.parametrize('prec', [64, 128])
def test_static_scan_matches_buffered_scan(prec):
    if not torch.cuda.is_available():
        pytest.skip('no HIP device')
    dev = torch.device('cuda', 0)
    big, ubig, rng = _plan(30, 11, dense=True)
    small, usmall, _ = _plan(16, 12, dense=False)
    print(f'case prec={prec} ranks {big.rmax} then {small.rmax}', flush=True)
    assert big.rmax > small.rmax and big.rmax <= 20
    for plan, u in ((big, ubig), (small, usmall)):
        bits = rng.integers(0, 2, (6, plan.n), dtype=np.uint8)
        out = _both(plan, u, bits, prec, dev)
        v0, e0 = out['rgs_scan_large']
        v1, e1 = out['rgs_scan_large_static']
        assert np.all(np.isfinite(v0))
        assert np.array_equal(v0, v1) and np.array_equal(e0, e1)


def test_static_scan_refuses_ranks_above_its_tables():
    from rgs70_classical_challenge_amd._lib import call, ptr
    if not torch.cuda.is_available():
        pytest.skip('no HIP device')
    plan, u, rng = _plan(46, 13, dense=True)
    assert plan.rmax > 21
    w = np.zeros(2, np.uint64)
    val = np.zeros(2)
    ex = np.zeros(1, np.int32)
    blob = np.zeros(4, np.uint64)
    offs = np.zeros(plan.n, np.uint64)
    for prec in (64, 128):
        with pytest.raises(ValueError, match='above the fixed tables'):
            # refused on the host before anything is read or launched
            call('rgs_scan_large_static', ctypes.addressof(plan.steps), plan.n, ptr(blob), ptr(offs), ptr(u), ptr(w),
                 prec, ptr(val), ptr(ex), None)
