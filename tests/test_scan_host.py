"""Host half of the rank scan on the CPU: the C++ plan compiler, profile and
annealer against the oracle and against the reference's own order optimizer
(built from its source into oracle/_ref/ when the reference is present)."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

from oracle import scan as osc
from rgs70_classical_challenge_amd import scan
from rgs70_classical_challenge_amd.instances import instance

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_SEARCH = os.path.join(ROOT, 'oracle', '_ref', 'search')


def test_compile_matches_oracle_plan():
    rng = np.random.default_rng(4)
    for n in (12, 34, 70):
        g, _ = instance(n, 77)
        order = rng.permutation(n).tolist()
        plan = scan.ScanPlan(g, order)
        steps, prof, _ = osc.compile_plan(g, order)
        assert plan.ranks == prof
        for st, ref in zip(plan.steps, steps):
            assert (st.v, st.r_in, st.r_out, st.beta) == (ref.v, ref.r_in, ref.r_out, ref.beta)
            assert list(st.lcols[:st.r_out]) == ref.lcols
            assert list(st.kern[:st.nker]) == ref.kernel


def test_profile_matches_dense_rank(golden_dir):
    d = np.load(f'{golden_dir}/rgs_graph.npz')
    for name in ('scan_order.npy', 'scan_order_optimized.npy'):
        order = np.load(f'{golden_dir}/{name}').tolist()
        assert scan.profile(d['graph'], order) == osc.cut_rank_profile(d['graph'], order)
        assert scan.ScanPlan(d['graph'], order).ranks == scan.profile(d['graph'], order)


def test_pack_layout():
    g, _ = instance(34, 3)
    order, _ = scan.order_search(g, seconds=0.5)
    plan = scan.ScanPlan(g, order)
    blob = plan.packed()
    assert blob[0] == 34 and blob[1] == plan.rmax
    for j, st in enumerate(plan.steps):
        off = int(blob[2 + j])
        hdr = int(blob[off])
        assert hdr & 255 == st.v and (hdr >> 8) & 255 == st.r_in and (hdr >> 16) & 255 == st.r_out
        for c in range(min(64, 1 << st.r_out)):
            want = 0
            for i in range(st.r_out):
                if i < 6 and (c >> i) & 1:
                    want ^= st.lcols[i]
            assert int(blob[off + 2 + c]) == want


def test_pack_tables_layout(golden_dir):
    """Byte-table plan: every output index's input vector L c' rebuilt from the
    byte tables equals the XOR of the lcols, and the combinations span the
    kernel (RGS70 ordering, ranks up to 30)."""
    d = np.load(f'{golden_dir}/rgs_graph.npz')
    plan = scan.ScanPlan(d['graph'], np.load(f'{golden_dir}/scan_order_optimized.npy').tolist())
    blob = plan.packed_tables()
    assert blob[0] == 70 and blob[1] == 30
    rng = np.random.default_rng(0)
    for j, st in enumerate(plan.steps):
        off = int(blob[2 + j])
        hdr = int(blob[off])
        nbytes = (st.r_out + 7) // 8
        assert (hdr & 255, (hdr >> 8) & 255, (hdr >> 16) & 255, (hdr >> 24) & 255, hdr >> 32) == \
            (st.v, st.r_in, st.r_out, st.nker, nbytes)
        assert int(blob[off + 1]) == st.beta
        for cp in rng.integers(0, 1 << st.r_out, 8).tolist() if st.r_out else [0]:
            want = 0
            for i in range(st.r_out):
                if (cp >> i) & 1:
                    want ^= st.lcols[i]
            got = 0
            for b in range(nbytes):
                got ^= int(blob[off + 2 + 256 * b + ((cp >> (8 * b)) & 255)])
            assert got == want
        combos = [int(x) for x in blob[off + 2 + 256 * nbytes: off + 2 + 256 * nbytes + (1 << st.nker)]]
        span = {0}
        for k in range(st.nker):
            span |= {x ^ st.kern[k] for x in span}
        assert sorted(combos) == sorted(span) and len(set(combos)) == 1 << st.nker
    words = blob.size
    assert words == 2 + 70 + sum(2 + 256 * ((st.r_out + 7) // 8) + (1 << st.nker) for st in plan.steps)


def test_order_search_improves_and_respects_cap():
    g, _ = instance(70, 1234)
    start = list(range(70))
    base = sum(1 << r for r in scan.profile(g, start))
    order, cost = scan.order_search(g, start, seconds=2.0, rmax_cap=20)
    prof = scan.profile(g, order)
    assert sorted(order) == start
    assert cost == sum(1 << r for r in prof) and max(prof) <= 20
    assert cost < base


@pytest.mark.skipif(not os.path.exists(REF_SEARCH), reason='reference optimizer not built (oracle/Makefile)')
def test_cost_agrees_with_reference_optimizer(golden_dir):
    """The reference's search.cpp (budget 0) prints the baseline cost of an order."""
    d = np.load(f'{golden_dir}/rgs_graph.npz')
    cases = [(d['graph'], np.load(f'{golden_dir}/scan_order.npy').tolist())]
    rng = np.random.default_rng(9)
    g, _ = instance(70, 5)
    cases.append((g, rng.permutation(70).tolist()))
    for g, order in cases:
        with tempfile.TemporaryDirectory() as tmp:
            inp = os.path.join(tmp, 'input.txt')
            np.savetxt(inp, np.concatenate([np.asarray(g, int).reshape(-1), order]), fmt='%d')
            out = subprocess.run([REF_SEARCH, inp, '0', tmp], capture_output=True, text=True, check=True).stdout
            base = [l for l in out.splitlines() if l.startswith('baseline')][0].split(',')
        prof = scan.profile(g, order)
        assert int(base[3]) == sum(1 << r for r in prof)
        assert int(base[4]) == max(prof)
