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
    order, _ = scan.order_search(g, iterations=100000)
    plan = scan.ScanPlan(g, order)
    blob = plan.packed()
    assert blob[0] == 34 and blob[1] == plan.rmax
    for j, st in enumerate(plan.steps):
        off = int(blob[2 + j])
        hdr = int(blob[off])
        assert hdr & 255 == st.v and (hdr >> 8) & 255 == st.r_in and (hdr >> 16) & 255 == st.r_out
        # split steps store the y = 0 member of p0's coset: p0 moved by the
        # kernel vector that carries y (the combination at the middle index)
        nhi = 1 << max(0, st.r_out - 6)
        kpiv = int(blob[off + 66 + nhi + (1 << st.nker) // 2]) if (hdr >> 32) & 1 else 0
        for c in range(min(64, 1 << st.r_out)):
            want = 0
            for i in range(st.r_out):
                if i < 6 and (c >> i) & 1:
                    want ^= st.lcols[i]
            if (want >> st.r_in) & 1:
                want ^= kpiv
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
    order, cost = scan.order_search(g, start, iterations=400000, rmax_cap=20)
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


def _packed_amplitude(blob, u, bits):
    """<x| (x)u |G> by walking the packed device plan in complex128 as the
    batch kernel does: per output c', p0 = lo[c' & 63] ^ hi[c' >> 6]; the
    first half of the combinations carries y = y(p0), the second (split
    steps) 1 - y(p0); each half is a signed sum of table entries times f_y.
    On split steps y(p0) is 0 for every output."""
    n = int(blob[0])
    E = np.ones(1, np.complex128)
    for j in range(n):
        off = int(blob[2 + j])
        hdr = int(blob[off])
        v, r_in, r_out, nker = hdr & 255, (hdr >> 8) & 255, (hdr >> 16) & 255, (hdr >> 24) & 255
        split = (hdr >> 32) & 1
        beta = int(blob[off + 1])
        nhi = 1 << max(0, r_out - 6)
        lo = [int(x) for x in blob[off + 2: off + 66]]
        hi = [int(x) for x in blob[off + 66: off + 66 + nhi]]
        combos = [int(x) for x in blob[off + 66 + nhi: off + 66 + nhi + (1 << nker)]]
        half = len(combos) // 2 if split else len(combos)
        f = u[v, int(bits[v])]
        out = np.zeros(1 << r_out, np.complex128)
        for cp in range(1 << r_out):
            p0 = lo[cp & 63] ^ hi[cp >> 6]
            ya = (p0 >> r_in) & 1
            assert not (split and ya)
            for part, ts in ((ya, combos[:half]), (1 - ya, combos[half:])):
                acc = 0j
                for t in ts:
                    c = (p0 ^ t) & ((1 << r_in) - 1)
                    acc += -E[c] if part and bin(c & beta).count('1') & 1 else E[c]
                out[cp] += acc * f[part]
        E = out
    return complex(E[0]) * 2.0 ** (-n / 2)


def test_packed_plan_halves_and_sum():
    """The packer reduces each kernel basis on the y bit: the combinations
    still span the kernel, their y bit is 0 in the first half and 1 in the
    second on split steps (0 throughout otherwise), and the halved signed
    sums over the packed plan equal the oracle sweep."""
    rng = np.random.default_rng(11)
    for n, p_edge in ((12, 0.5), (14, 0.35), (34, None)):
        if p_edge is None:
            g, u = instance(n, 5)
        else:
            a = np.triu(rng.random((n, n)) < p_edge, 1).astype(np.uint8)
            g = a | a.T
            z = rng.normal(size=(n, 2, 2)) + 1j * rng.normal(size=(n, 2, 2))
            u = np.stack([np.linalg.qr(m)[0] for m in z])
        order = list(range(n)) if p_edge else scan.order_search(g, iterations=60000)[0]
        plan = scan.ScanPlan(g, order)
        blob = plan.packed()
        steps, _, _ = osc.compile_plan(g, order)
        splits = 0
        moved = False          # a step after a coordinate change has rewritten p-vectors
        for j, ref in enumerate(steps):
            off = int(blob[2 + j])
            hdr = int(blob[off])
            nker, r_in, r_out = (hdr >> 24) & 255, (hdr >> 8) & 255, (hdr >> 16) & 255
            nhi = 1 << max(0, r_out - 6)
            combos = [int(x) for x in blob[off + 66 + nhi: off + 66 + nhi + (1 << nker)]]
            if not moved:
                assert sorted(combos) == sorted(osc.kernel_combos(ref.kernel))
            moved = moved or bool((hdr >> 33) & 1)
            ybits = [(t >> r_in) & 1 for t in combos]
            split = (hdr >> 32) & 1
            splits += split
            want = [0] * (len(combos) // 2) + [1] * (len(combos) // 2) if split else [0] * len(combos)
            assert ybits == want
        if p_edge:
            assert splits > 0
        for _ in range(3):
            bits = rng.integers(0, 2, n)
            if n <= 14:
                assert _packed_amplitude(blob, u, bits) == pytest.approx(osc.amplitude(steps, u, bits), abs=1e-13)


def test_kernel_dimension_at_most_two(golden_dir):
    """Each step clears one coordinate of the r_in basis vectors (rank drops
    by at most one) and adds one vector, so r_out >= r_in - 1 and the kernel
    of p = (c, y) -> c' has dimension r_in + 1 - r_out <= 2; the packer
    relies on it.  Checked on dense and sparse random graphs in random
    orders, and on the RGS70 graph with its optimised order."""
    d = np.load(f'{golden_dir}/rgs_graph.npz')
    plan = scan.ScanPlan(d['graph'], np.load(f'{golden_dir}/scan_order_optimized.npy').tolist())
    assert max(st.nker for st in plan.steps) <= 2
    rng = np.random.default_rng(21)
    seen = set()
    for _ in range(120):
        n = int(rng.integers(8, 19))
        a = np.triu(rng.random((n, n)) < rng.uniform(0.1, 0.95), 1).astype(np.uint8)
        steps, prof, _ = osc.compile_plan(a | a.T, rng.permutation(n).tolist())
        for st in steps:
            assert st.r_out >= st.r_in - 1 and len(st.kernel) == st.r_in + 1 - st.r_out <= 2
            seen.add(len(st.kernel))
    assert seen == {0, 1, 2}


def _twin(st):
    """d with e_y = L d ^ K z (the lcols and kernel vectors span p-space), or 0."""
    cols = list(st.lcols) + list(st.kernel)
    basis = []                       # (vector, combination mask), distinct top bits
    for i, v in enumerate(cols):
        c = 1 << i
        for bv, bc in basis:
            if v ^ bv < v:           # bv's top bit is set in v
                v, c = v ^ bv, c ^ bc
        if v:
            basis.append((v, c))
            basis.sort(key=lambda e: -e[0])
    t, tc = 1 << st.r_in, 0
    for bv, bc in basis:
        if t ^ bv < t:
            t, tc = t ^ bv, tc ^ bc
    return 0 if t else tc & ((1 << st.r_out) - 1)


def test_twin_outputs_share_entries():
    """Output c' ^ d reads the same table entries as output c' with y flipped,
    so a sweep that produces each output pair from one read of its entries
    equals the oracle sweep."""
    rng = np.random.default_rng(8)
    twins = 0
    for n in (10, 12, 14):
        for _ in range(4):
            a = np.triu(rng.random((n, n)) < rng.uniform(0.2, 0.8), 1).astype(np.uint8)
            g = a | a.T
            z = rng.normal(size=(n, 2, 2)) + 1j * rng.normal(size=(n, 2, 2))
            u = np.stack([np.linalg.qr(m)[0] for m in z])
            order = rng.permutation(n).tolist()
            steps, _, _ = osc.compile_plan(g, order)
            bits = rng.integers(0, 2, n)
            E = np.ones(1, np.complex128)
            for st in steps:
                f = u[st.v, int(bits[st.v])]
                d = _twin(st)
                combos = osc.kernel_combos(st.kernel)
                out = np.zeros(1 << st.r_out, np.complex128)
                done = np.zeros(1 << st.r_out, bool)
                for cp in range(1 << st.r_out):
                    if done[cp]:
                        continue
                    p0 = 0
                    for i in range(st.r_out):
                        if (cp >> i) & 1:
                            p0 ^= st.lcols[i]
                    acc = [0j, 0j]
                    for t in combos:
                        p = p0 ^ t
                        c, y = p & ((1 << st.r_in) - 1), (p >> st.r_in) & 1
                        par = bin(c & st.beta).count('1') & 1
                        acc[0] += E[c] * f[y] * (-1 if y and par else 1)
                        acc[1] += E[c] * f[1 - y] * (-1 if (1 - y) and par else 1)
                    out[cp], done[cp] = acc[0], True
                    if d:
                        out[cp ^ d], done[cp ^ d] = acc[1], True
                        twins += 1
                E = out
            want = osc.amplitude(steps, u, bits)
            assert complex(E[0]) * 2.0 ** (-n / 2) == pytest.approx(want, abs=1e-12)
    assert twins > 0


def test_packed_coordinates_make_weights_and_signs_uniform():
    """On steps with at least 256 outputs the packer re-indexes the output
    table so that y(p0) and the sign form vanish on the six lane bits (the
    header's uniform flag): every lane word has y = 0 and even parity against
    beta.  The re-indexed plan still gives the oracle's amplitudes."""
    rng = np.random.default_rng(12)
    n = 20
    a = np.triu(rng.random((n, n)) < 0.5, 1).astype(np.uint8)
    g = a | a.T
    z = rng.normal(size=(n, 2, 2)) + 1j * rng.normal(size=(n, 2, 2))
    u = np.stack([np.linalg.qr(m)[0] for m in z])
    order = list(range(n))
    plan = scan.ScanPlan(g, order)
    assert plan.rmax >= 9
    blob = plan.packed()
    uniform = 0
    for j in range(n):
        off = int(blob[2 + j])
        hdr = int(blob[off])
        r_in, r_out = (hdr >> 8) & 255, (hdr >> 16) & 255
        if r_out >= 8:
            assert (hdr >> 33) & 1, j
        if (hdr >> 33) & 1:
            uniform += 1
            beta = int(blob[off + 1])
            mask = (1 << r_in) - 1
            for c in range(64):
                lo = int(blob[off + 2 + c])
                assert (lo >> r_in) & 1 == 0 and bin(lo & mask & beta).count('1') % 2 == 0, (j, c)
    assert uniform >= 3
    steps, _, _ = osc.compile_plan(g, order)
    for _ in range(3):
        bits = rng.integers(0, 2, n)
        assert _packed_amplitude(blob, u, bits) == pytest.approx(osc.amplitude(steps, u, bits), abs=1e-13)


def test_tuned_scan_orders_are_valid(golden_dir):
    """Every scan order in the autotuner's table is a permutation whose
    recorded rank and sum of 2^r are its profile's; the 70-qubit instance's
    order costs no more than the reference's optimised order (7,919,632,381,
    order_optimization/search_summary.json)."""
    import json
    from rgs70_classical_challenge_amd import autotune
    table = json.load(open(autotune.TABLE))
    rgs = np.load(f'{golden_dir}/rgs_graph.npz')['graph']
    seen = 0
    for dev, entries in table.items():
        for key, e in entries.items():
            if not key.startswith('scan:'):
                continue
            seen += 1
            n = len(e['order'])
            assert sorted(e['order']) == list(range(n))
            if key == autotune.graph_key(rgs):
                prof = scan.profile(rgs, e['order'])
                assert max(prof) == e['rmax'] == 30
                assert sum(1 << r for r in prof) == e['sum_states'] <= 7919632381
    assert seen >= 3


@pytest.mark.slow
def test_order_search_beats_reference_optimizer_on_rgs70(golden_dir):
    """The annealer, from the reference's starting order with a fixed budget
    (8 chains x 8,000,000 moves, seed 20261030, rank cap 30), reaches a lower
    sum of 2^r than the reference's own 155-second search (7,919,632,381),
    and is reproducible."""
    d = np.load(f'{golden_dir}/rgs_graph.npz')
    start = np.load(f'{golden_dir}/scan_order.npy').tolist()
    order, cost = scan.order_search(d['graph'], start, iterations=8_000_000, rmax_cap=30, threads=8)
    prof = scan.profile(d['graph'], order)
    assert cost == sum(1 << r for r in prof) and max(prof) <= 30
    assert cost <= 7919632381, cost
    assert cost == 7854882813
