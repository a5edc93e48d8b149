#!/usr/bin/env python3
"""Headline benchmark: exact samples/s at n=34 on one MI355X.

Workload (one "step"): a seeded random 3-regular graph on n=34 vertices with
QR-drawn single-qubit unitaries; the state (x)_k u[k] |G> (2^34 complex64
amplitudes, 128 GiB) is built by the slab pipeline - GF(2) phase init ->
Walsh-Hadamard over all 34 bits -> per-qubit 2x2 gates -> FP64 masses - in
a few fused HBM passes (tile plan from the autotuner table, else the default
plan; the table's n=34 entry is the fused Walsh schedule: four passes, each
applying the Walsh butterflies, D2 phases and rotations of its own bits, the
first making the state and the last writing the masses), the FP64 mass
pyramid is built, and 10,000 exact samples are drawn
qubit by qubit on the device.  The graph, tables and the 10,000 uniforms of
every step are resident in HBM before the timed region.

Reported beside the headline value:
  roofline      the dominant pass kernel (the k_tile specialisation with the
                most time per step): algorithmic bytes per launch (16 B per
                amplitude for a read+write pass, 8 B for the write-only init
                pass) / its mean HIP-event duration on the launch stream,
                against the 8 TB/s HBM peak; `traffic` = measured HBM bytes per
                launch from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE
                summary, when present.
  kernels       the same figures for every specialisation in the plan.
  passes        per-pass ms and GB/s; `walsh_gbps` over the Walsh passes.
  deferred      the same step with the last pass deferred: it writes the
                FP64 masses but not the state, and each draw recomputes the
                tile holding its leaf block; samples/s, the last pass's ms and
                read GB/s, and whether the draws equal the stored-pass draws.
  folded        the same state with H^n|0> folded into the init (no Walsh
                passes), samples/s.
  stream_reference  torch copy_ (read + write) and fill_ (write only) of
                32 GiB on the same device: the library's own streaming rates,
                beside which the pass rates can be read.
  scan          batched rank-scan p(x) of 65,536 device-resident uniform 70-bit
                proposals of a 70-vertex 3-regular instance per launch (FP64
                sums), proposals/s with complex64 tables (the scorer) and
                complex128 tables (the audit), the complex64-vs-complex128
                error, and the reported p(x): complex64 tables with the
                proposals of 2^n p < 0.1 rescored in complex128, its rate and
                its error against complex128.
  scan_rgs70    the rank-30 scan of the reference's own 70-qubit instance
                (HBM tables): seconds per amplitude, complex64 and complex128.
  production    the reference's production conditional sampler on its own
                70-qubit instance (m=35: one 2^35 complex64 slab), replaying
                6 recorded samples: seconds per sample, index matches and p
                error against the recorded complex128 values.
  learned       learned (MADE, H = 16 (n - 1)) against uniform proposals for
                capped rejection at n=34: the MADE is trained on the GPU from
                exact samples, both proposal streams are judged by the exact
                scan and audited in complex128; accepted samples/s of each at
                FP64-audited linear XEB equal within one standard error.
  batch         config 4: 256 uneven instances (n in 16..30), 1,000 samples
                each, as one graph launch against one launch sequence per
                instance, and the launch-bound n <= 22 part alone.
  cpu_baseline  the oracle (NumPy restatement of the reference's CPU
                full-state path) on the host at n=26 - state, mass pyramid,
                10,000 draws - on every core this process may use (state
                passes split over threads) and on one core, extrapolated to
                n=34 (state n 2^n, pyramid 2^n, draws n).

`--calibrate-copy` instead runs one plain device copy of 2^32 bytes (torch
copy_), the reference kernel for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE.

Prints one JSON line on stdout; progress goes to stderr.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBPS = 8000.0
METRIC = 'exact samples/s at n=34 on 1 MI355X; WHT HBM GB/s vs peak; learned vs plain'


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def pass_bytes(plan_text, n):
    """Algorithmic HBM bytes of each pass: 8 B/amp write-only, 16 B/amp read+write."""
    out = []
    for tok in plan_text.split():
        out.append((8 if '[init]' in tok else 16) * (1 << n))
    return out


def kernel_of(tok, fused=False):
    """Kernel of a plan token such as 'strided[wht]:14+8/tb14/c6/rb2': the
    contiguous mass passes have their flags fixed at compile time
    (k_tile_mass<TB, RBW, flags>: gates + masses 25 storing the state, 89
    masses only; the fused schedule's Walsh + phase + gates + masses, layout
    by layout, 159 and 223), and so do the strided passes that load the state
    and apply Walsh, phase and gates (k_tile_fixed<TB, C, RBW, flags>: 143 in
    the fused schedule's single walk, 15 as the seven-pass schedule's phase
    pass); every other pass is the k_tile specialisation of its shape."""
    kind, tb, c, rb = tok.split('/')
    head = kind.split(':')[0]
    mass = {'low[gates][mass]': 25, 'low[gates][mass][leaf]': 89, 'low[wht][phase][gates][mass]': 159,
            'low[wht][phase][gates][mass][leaf]': 223}
    if head in mass:
        return f'k_tile_mass<{int(tb[2:])}, {int(rb[2:])}, {mass[head]}>'
    if head == 'strided[wht][phase][gates]':
        return f'k_tile_fixed<{int(tb[2:])}, {int(c[1:])}, {int(rb[2:])}, {143 if fused else 15}>'
    return f'k_tile<{int(tb[2:])}, {int(c[1:])}, {int(rb[2:])}>'


def run_headline(args, torch):
    from rgs70_classical_challenge_amd import _lib
    from rgs70_classical_challenge_amd.engine import GraphStateSampler, ctypes_ref
    from rgs70_classical_challenge_amd.instances import instance

    n = args.n
    g, u = instance(n, args.seed)
    log(f'[bench] allocating n={n} state ({8 << n >> 30} GiB)')
    s = GraphStateSampler(g, u, device=0, pipeline=args.pipeline)
    sl = s.slab
    log(f'[bench] plan {s.passes} passes: {s.plan_text}')
    stream = sl.stream
    sp = stream.cuda_stream
    rng = np.random.Generator(np.random.PCG64(args.seed + 1))
    total = args.warmup + args.steps
    uni = torch.as_tensor(rng.random((total, args.samples))).to(sl.device)
    idx = torch.empty(args.samples, dtype=torch.int64, device=sl.device)
    pdr = torch.empty(args.samples, dtype=torch.float64, device=sl.device)
    npass = s.passes
    ev = [[(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
           for _ in range(npass + 2)] for _ in range(total)]

    def step(k):
        for p in range(npass):
            ev[k][p][0].record(stream)
            _lib.call('rgs_slab_pass', ctypes_ref(s._desc), p, sp)
            ev[k][p][1].record(stream)
        ev[k][npass][0].record(stream)
        _lib.call('rgs_mass_pyramid', _lib.ptr(sl.pyramid), sl.base_bits, sp)
        ev[k][npass][1].record(stream)
        ev[k][npass + 1][0].record(stream)
        _lib.call('rgs_draw', _lib.ptr(sl.state), n, _lib.ptr(sl.pyramid), sl.mass_bits, _lib.ptr(uni[k]),
                  args.samples, s.flip, _lib.ptr(idx), _lib.ptr(pdr), sp)
        ev[k][npass + 1][1].record(stream)

    for k in range(args.warmup):
        step(k)
        torch.cuda.synchronize()
        log(f'[bench] warmup {k + 1}/{args.warmup} done')
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.warmup, total):
        step(k)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    log(f'[bench] timed {args.steps} steps in {t1 - t0:.3f} s')
    norm = s.norm()
    timed = range(args.warmup, total)
    pass_ms = [float(np.mean([ev[k][p][0].elapsed_time(ev[k][p][1]) for k in timed])) for p in range(npass + 2)]
    nbytes = pass_bytes(s.plan_text, n)
    per_pass = []
    for p, tok in enumerate(s.plan_text.split()):
        per_pass.append({'pass': tok, 'ms': round(pass_ms[p], 3), 'gbps': round(nbytes[p] / pass_ms[p] / 1e6, 1)})
    wht = [p for p, tok in enumerate(s.plan_text.split()) if '[wht]' in tok]
    wht_gbps = sum(nbytes[p] for p in wht) / sum(pass_ms[p] for p in wht) / 1e6 if wht else None
    kernels = {}
    for p, tok in enumerate(s.plan_text.split()):
        k = kernels.setdefault(kernel_of(tok, bool(s._desc.fuse_walsh)), {'launches': 0, 'ms': 0.0, 'bytes': 0})
        k['launches'] += 1
        k['ms'] += pass_ms[p]
        k['bytes'] += nbytes[p]
    for k in kernels.values():
        k['avg_ms'] = round(k['ms'] / k['launches'], 3)
        k['bytes_per_launch'] = k['bytes'] / k['launches']
        k['gbps'] = round(k['bytes'] / k['ms'] / 1e6, 1)
        k['ms'] = round(k['ms'], 3)
    dominant = max(kernels, key=lambda name: kernels[name]['ms'])
    # probabilities of the drawn samples stay on the device; sanity-check a few on the host
    drawn = idx[:16].cpu().numpy()
    assert np.all((drawn >= 0) & (drawn < (1 << n)))
    res = dict(s=s, t=t1 - t0, norm=norm, per_pass=per_pass, wht_gbps=wht_gbps, kernels=kernels,
               last_uniforms=uni[total - 1], last_index=idx.cpu().numpy(), last_p=pdr.cpu().numpy(),
               dominant=dominant, passes_gbps=sum(nbytes) / sum(pass_ms[:npass]) / 1e6,
               pyramid_ms=pass_ms[npass], draw_ms=pass_ms[npass + 1])
    return res


def run_deferred(s, res, args, torch):
    """The headline step with its last pass deferred: that pass writes the
    FP64 masses only and each draw recomputes the tile holding its leaf block
    (rgs_draw_leaves).  The last timed step's draws must equal the stored-pass
    draws of the same uniforms bit for bit."""
    from rgs70_classical_challenge_amd import _lib
    from rgs70_classical_challenge_amd.engine import ctypes_ref
    import ctypes
    sl = s.slab
    s.defer_leaves = True
    d = s._make_desc()
    passes, text = sl.describe(d)
    sp = sl.stream.cuda_stream
    nb = ctypes.c_size_t()
    _lib.call('rgs_draw_leaves_workspace', ctypes_ref(d), args.samples, ctypes.byref(nb))
    work = torch.empty(nb.value, dtype=torch.uint8, device=sl.device)
    uni = res['last_uniforms']
    idx = torch.empty(args.samples, dtype=torch.int64, device=sl.device)
    pdr = torch.empty(args.samples, dtype=torch.float64, device=sl.device)
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(3)]
    times = []

    def step():
        for p in range(passes - 1):
            _lib.call('rgs_slab_pass', ctypes_ref(d), p, sp)
        ev[0][0].record(sl.stream)
        _lib.call('rgs_slab_pass', ctypes_ref(d), passes - 1, sp)
        ev[0][1].record(sl.stream)
        ev[1][0].record(sl.stream)
        _lib.call('rgs_mass_pyramid', _lib.ptr(sl.pyramid), sl.base_bits, sp)
        ev[1][1].record(sl.stream)
        ev[2][0].record(sl.stream)
        _lib.call('rgs_draw_leaves', ctypes_ref(d), _lib.ptr(sl.pyramid), _lib.ptr(uni), args.samples,
                  _lib.ptr(work), nb.value, _lib.ptr(idx), _lib.ptr(pdr), sp)
        ev[2][1].record(sl.stream)

    step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
        torch.cuda.synchronize()
        times.append([a.elapsed_time(b) for a, b in ev])
    dt = (time.perf_counter() - t0) / args.steps
    same = bool(np.array_equal(idx.cpu().numpy(), res['last_index'])
                and np.array_equal(pdr.cpu().numpy(), res['last_p']))
    ms = np.mean(times, axis=0)
    s.defer_leaves = False
    s._desc = s._make_desc()
    return {'passes': passes, 'plan': text, 'ms_per_step': round(dt * 1e3, 2),
            'samples_per_s': round(args.samples / dt, 1), 'last_pass_ms': round(float(ms[0]), 3),
            'last_pass_gbps': round((8 << s.n) / ms[0] / 1e6, 1), 'pyramid_ms': round(float(ms[1]), 3),
            'draw_ms': round(float(ms[2]), 3), 'draws_equal_stored': same, 'norm': sl.root()}


def run_folded(s, args, torch):
    """Same state with H^n|0> folded into the init (3 passes)."""
    from rgs70_classical_challenge_amd import _lib
    from rgs70_classical_challenge_amd.engine import Gates, ctypes_ref
    sl = s.slab
    d = sl.desc(_lib.INIT_GRAPH, 2.0 ** (-s.n / 2), init_lut=s._lut, init_rows=s._rows, gates=Gates(sl, s.u))
    passes, text = sl.describe(d)
    rng = np.random.default_rng(5)
    uni = torch.as_tensor(rng.random(args.samples)).to(sl.device)
    idx = torch.empty(args.samples, dtype=torch.int64, device=sl.device)
    pdr = torch.empty(args.samples, dtype=torch.float64, device=sl.device)
    sp = sl.stream.cuda_stream

    def step():
        _lib.call('rgs_slab_pipeline', ctypes_ref(d), sp)
        _lib.call('rgs_mass_pyramid', _lib.ptr(sl.pyramid), sl.base_bits, sp)
        _lib.call('rgs_draw', _lib.ptr(sl.state), s.n, _lib.ptr(sl.pyramid), sl.mass_bits, _lib.ptr(uni),
                  args.samples, s.flip, _lib.ptr(idx), _lib.ptr(pdr), sp)
    step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    return {'passes': passes, 'plan': text, 'ms_per_step': round(dt * 1e3, 2),
            'samples_per_s': round(args.samples / dt, 1), 'norm': sl.root()}


def run_scan(torch, count=65536, seed=70, reps=5):
    """Batched rank-scan p(x) of `count` uniform proposals already resident on
    the device (packed uint64 words), one launch per batch, complex64 tables
    (the scorer) and complex128 tables (the audit); mean over `reps` launches
    timed with events on the scorer's stream."""
    from rgs70_classical_challenge_amd.instances import instance
    from rgs70_classical_challenge_amd.scan import ScanOracle, bits_to_words
    g, u = instance(70, seed)
    so = ScanOracle(g, u, 'auto')            # the autotuner's order (tuned.json)
    bits = np.random.default_rng(seed).integers(0, 2, (count, 70), dtype=np.uint8)
    words = torch.as_tensor(bits_to_words(bits).view(np.int64)).to(so.device)
    out = {'workload': f'70-vertex 3-regular, {count} uniform proposals per launch (device-resident)',
           'rmax': so.rmax, 'sum_states': so.plan.sum_states, 'order': 'autotuner (tuned.json)'}
    for prec in (64, 128):
        so.probabilities_device(words, prec)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(so.stream)
        for _ in range(reps):
            so.probabilities_device(words, prec)
        e1.record(so.stream)
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / reps
        key = '' if prec == 64 else '_c128'
        out['ms_per_launch' + key] = round(ms, 3)
        out['proposals_per_s' + key] = round(count / ms * 1e3, 1)
    out['table_gbps'] = round(so.plan.table_bytes(64) * count / out['ms_per_launch'] / 1e6, 1)
    # the reported p(x): complex64 tables, proposals with 2^n p < 0.1 rescored
    # with complex128 tables (a second launch over the flagged ones)
    so.probabilities_checked(words)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        p, flag = so.probabilities_checked(words)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) / reps * 1e3
    out['ms_per_launch_checked'] = round(ms, 3)
    out['proposals_per_s_checked'] = round(count / ms * 1e3, 1)
    out['rescored_fraction'] = round(float(flag.float().mean().item()), 5)
    p128 = so.probabilities_device(words, 128)
    p64 = so.probabilities_device(words, 64)
    rel = (torch.abs(p64 - p128) / p128).cpu().numpy()
    relc = (torch.abs(p - p128) / p128).cpu().numpy()
    out['c64_vs_c128_rel_err'] = {'median': float(np.median(rel)), 'p99': float(np.quantile(rel, 0.99)),
                                  'max': float(rel.max())}
    out['reported_vs_c128_rel_err'] = {'median': float(np.median(relc)), 'p99': float(np.quantile(relc, 0.99)),
                                       'max': float(relc.max())}
    return out


def run_scan_rgs70(torch, count=8):
    """Rank-30 scan of the reference's own 70-qubit instance with its
    optimised order (HBM tables, one proposal at a time): seconds per
    amplitude in complex64 and complex128 tables.  Reference (1 x H200):
    0.0750 s (complex64, this order) and about 0.130 s (complex128)."""
    from rgs70_classical_challenge_amd.scan import ScanOracle
    golden = os.path.join(ROOT, 'tests', 'golden')
    d = np.load(os.path.join(golden, 'rgs_graph.npz'))
    order = np.load(os.path.join(golden, 'scan_order_optimized.npy')).tolist()
    props = json.load(open(os.path.join(golden, 'rgs70_cap1_proposals.json')))['proposals'][:count]
    bits = np.array([[int(c) for c in r['bitstring_q0_left']] for r in props], np.uint8)
    want = np.array([r['p_fp64'] for r in props])
    out = {'workload': f'70-qubit instance, optimised order (max rank 30), {count} proposals',
           'reference_seconds_per_amplitude': {'complex64': 0.0750, 'complex128': 0.130}}
    for name, fp64 in (('complex64', False), ('complex128', True)):
        so = ScanOracle(d['graph'], d['measurement_unitaries'], order, fp64=fp64)
        so.probabilities(bits[:1])                      # tables, buffers, first launch
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        p = so.probabilities(bits)
        dt = (time.perf_counter() - t0) / count
        size = 16 if fp64 else 8
        out[name] = {'seconds_per_amplitude': round(dt, 5), 'rmax': so.rmax,
                     'table_gbps': round(so.plan.table_bytes(128 if fp64 else 64) / dt / 1e9, 1),
                     'max_rel_err_vs_recorded_fp64': float(np.max(np.abs(p - want) / want)),
                     'table_bytes_per_amplitude': so.plan.table_bytes(128 if fp64 else 64), 'entry_bytes': size}
        del so
        torch.cuda.empty_cache()
    return out


def run_learned(torch, n=34):
    """Learned versus uniform proposals for capped rejection (accepted
    samples/s at equal audited XEB); the full run's record minus the traces."""
    from rgs70_classical_challenge_amd.learned import run
    r = run(n=n, caps_learned=(1.6, 1.8, 2.0, 2.2, 2.5, 2.8), log=log)
    keep = ('heldout_gain_nats', 'heldout_ceiling_nats', 'train_s', 'hidden', 'steps', 'scan_rmax', 'scan_sum_states')
    out = {k: r[k] for k in keep}
    out['uniform'] = {k: r['uniform'][k] for k in ('accepted_per_s', 'proposals_per_s', 'acceptance', 'xeb',
                                                   'xeb_se', 'decision_flips_fp64', 'decision_flips_beyond_tolerance')}
    out['learned'] = {c: {k: v[k] for k in ('accepted_per_s', 'proposals_per_s', 'acceptance', 'xeb', 'xeb_se',
                                            'decision_flips_fp64', 'decision_flips_beyond_tolerance')}
                      for c, v in r['learned'].items()}
    out['equal_xeb_comparison'] = r['equal_xeb_comparison']
    out['analysis'] = r.get('analysis')
    return out


def run_batch(torch, count=256):
    """Config 4: 256 instances with n uniform in 16..30 (3-regular for even n,
    Erdos-Renyi G(n, 3/n) for odd n), 1,000 samples each, tile plans from the
    autotuner's table: one launch sequence per instance (eager), one graph per
    memory wave, the whole batch as one graph, and the native multi-state
    batch (states of one pass plan share each pass's launch), for the whole
    batch and for its launch-bound n <= 22 part."""
    from rgs70_classical_challenge_amd.batch import BatchSampler, uneven_instances
    specs = uneven_instances(count, nmin=16, nmax=30)
    res = BatchSampler(specs, samples=1000).benchmark()
    res['workload'] = f'{count} instances, n uniform in 16..30, 1000 samples each'
    small = [sp for sp in specs if sp[0] <= 22]
    sres = BatchSampler(small, samples=1000).benchmark()
    res['small_instances'] = {'n_max': 22, 'instances': len(small), 'amplitudes': sres['amplitudes'],
                              **{k: sres[k] for k in ('eager_ms', 'graph_ms', 'single_graph_ms', 'graph_speedup',
                                                      'single_graph_speedup', 'fused_ms', 'fused_graph_ms',
                                                      'fused_launches', 'fused_groups', 'fused_speedup',
                                                      'fused_graph_speedup')}}
    return res


def run_production(torch, count=6):
    """The reference's production workload on one device: the 70-qubit
    instance's 2^35-amplitude conditional state per sample, replaying its
    recorded B bits and draw uniforms (rgs70_born_records.json keeps 24 of
    them); reference: median 2.4162 s per sample on 2 x H200."""
    from rgs70_classical_challenge_amd.production import replay
    golden = os.path.join(ROOT, 'tests', 'golden')
    d = np.load(os.path.join(golden, 'rgs_graph.npz'))
    data = json.load(open(os.path.join(golden, 'rgs70_born_records.json')))
    summary, rows = replay(d['graph'], d['measurement_unitaries'], data['A_vertices'], data['B_vertices'],
                           data['records'][:count], log=log)
    summary['reference_median_seconds_per_sample'] = 2.4162
    summary['reference_hardware'] = '2 x H200, 2^34 complex64 amplitudes per GPU'
    summary['workload'] = f'70-qubit RGS instance, m=35 conditional state (256 GiB complex64), {count} recorded samples'
    return summary


def cpu_baseline(seed):
    """The oracle (NumPy restatement of the reference's CPU full-state path,
    graph_sampler.py:71-93) on the host: the n=26 state (graph phase, then one
    2x2 gate pass per qubit), its FP64 pyramid and 10,000 draws, on one core
    and with the state passes split over every core this process may use.
    Extrapolated to n=34: the state and pyramid work scale as n 2^n and 2^n,
    the draws (one descent each) as n."""
    from oracle import draw as od
    from oracle import graph_state as gs
    from rgs70_classical_challenge_amd.instances import instance
    n0, samples, n = 26, 10000, 34
    cores = len(os.sched_getaffinity(0))
    g, u = instance(n0, seed)
    uni = np.random.default_rng(1).random(samples)
    t0 = time.perf_counter()
    v = gs.full_state(g, u)
    t_state = time.perf_counter() - t0
    t0 = time.perf_counter()
    vt = gs.full_state_threads(g, u, cores)
    t_state_mt = time.perf_counter() - t0
    assert np.array_equal(v, vt)
    del vt
    t0 = time.perf_counter()
    levels = od.pyramid(od.probabilities(v))
    t_pyr = time.perf_counter() - t0
    t0 = time.perf_counter()
    od.tree_draws(levels, uni)
    t_draw = time.perf_counter() - t0
    # one write pass (16 B per complex128 amplitude) and n0 read+write gate passes
    pass_bytes = (16 + 32 * n0) * (1 << n0)
    grow = 2.0 ** (n - n0)

    def per_s(ts):
        return samples / (ts * grow * n / n0 + t_pyr * grow + t_draw * n / n0)

    one = {'value': round(per_s(t_state), 4), 'cores': 1, 'state_seconds': round(t_state, 3),
           'state_gbps_complex128': round(pass_bytes / t_state / 1e9, 2)}
    return {'value': round(per_s(t_state_mt), 4), 'unit': 'samples/s', 'cores': cores, 'kind': 'port',
            'sample': f'oracle at n={n0}: state {t_state_mt:.2f} s on {cores} threads ({t_state:.2f} s on 1), '
                      f'pyramid {t_pyr:.2f} s, {samples} draws {t_draw:.2f} s; extrapolated to n={n} '
                      f'(state x2^{n - n0}*{n}/{n0}, pyramid x2^{n - n0}, draws x{n}/{n0})',
            'state_gbps_complex128': round(pass_bytes / t_state_mt / 1e9, 2), 'one_core': one,
            'pyramid_seconds': round(t_pyr, 3), 'draw_seconds': round(t_draw, 3)}


def stream_reference(torch, gib=32):
    """Library streaming rates on the same device, as context for the pass
    rates: torch copy_ (read + write) and fill_ (write only) of gib GiB,
    best of five each (HIP events)."""
    n = (gib << 30) // 4
    a = torch.ones(n, dtype=torch.float32, device='cuda')
    b = torch.empty_like(a)

    def best(fn):
        fn()
        torch.cuda.synchronize()
        t = []
        for _ in range(5):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            fn()
            e1.record()
            torch.cuda.synchronize()
            t.append(e0.elapsed_time(e1))
        return min(t)

    copy_ms = best(lambda: b.copy_(a))
    fill_ms = best(lambda: b.fill_(0.0))
    del a, b
    torch.cuda.empty_cache()
    return {'bytes': gib << 30, 'torch_copy_ms': round(copy_ms, 3),
            'torch_copy_gbps': round(2 * (gib << 30) / copy_ms / 1e6, 1),
            'torch_fill_ms': round(fill_ms, 3), 'torch_fill_gbps': round((gib << 30) / fill_ms / 1e6, 1)}


def calibrate_copy(torch):
    """One plain device-to-device copy of 2^32 bytes for counter calibration."""
    a = torch.ones(1 << 30, dtype=torch.float32, device='cuda')
    b = torch.empty_like(a)
    b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({'calibration': 'torch copy_', 'bytes_read': 1 << 32, 'bytes_written': 1 << 32}), flush=True)


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=5)
    ap.add_argument('--warmup', type=int, default=2)
    ap.add_argument('--n', type=int, default=34)
    ap.add_argument('--samples', type=int, default=10000)
    ap.add_argument('--pipeline', default='slab', choices=['slab', 'literal', 'folded'])
    ap.add_argument('--seed', type=int, default=20261203)
    ap.add_argument('--no-cpu', action='store_true', help='skip the host baseline')
    ap.add_argument('--no-extras', action='store_true', help='skip folded / scan extras')
    ap.add_argument('--calibrate-copy', action='store_true', help='run only the counter-calibration copy')
    args = ap.parse_args()
    world = int(os.environ.get('WORLD_SIZE', '1'))
    if args.gpus != 1 or world != 1:
        print('bench.py measures one GPU only: this build runs the per-GPU slab on a single MI355X; '
              'multi-GPU sharding is a design (see the design notes), not a measured path.', file=sys.stderr)
        sys.exit(2)
    import torch
    if not torch.cuda.is_available():
        print('bench.py needs a HIP device', file=sys.stderr)
        sys.exit(1)
    if args.calibrate_copy:
        calibrate_copy(torch)
        return
    from rgs70_classical_challenge_amd import _lib
    _lib.lib()

    res = run_headline(args, torch)
    s = res['s']
    value = args.samples * args.steps / res['t']
    out = {
        'metric': METRIC,
        'value': round(value, 2),
        'unit': 'samples/s',
        'n_gpus': 1,
        'steps': args.steps,
        'warmup': args.warmup,
        'ms_per_step': round(res['t'] / args.steps * 1e3, 3),
        'higher_is_better': True,
        'scaling': 'weak',
        'vs_baseline': None,
        'dtype': 'c64',
        'data': 'synthetic',
        'config': {'workload': f'n={args.n} 3-regular graph state, QR unitaries, {args.pipeline} pipeline '
                               f'(init -> Walsh -> gates -> FP64 masses), {args.samples} exact samples per step',
                   'n': args.n, 'samples_per_step': args.samples, 'pipeline': args.pipeline, 'seed': args.seed,
                   'passes': s.passes, 'plan': s.plan_text},
        'roofline': None,
        'kernels': res['kernels'],
        'passes_gbps': round(res['passes_gbps'], 1),
        'walsh_gbps': round(res['wht_gbps'], 1) if res['wht_gbps'] else None,
        'passes': res['per_pass'],
        'pyramid_ms': round(res['pyramid_ms'], 3),
        'draw_ms': round(res['draw_ms'], 3),
        'fp64_norm': res['norm'],
    }
    dom = res['dominant']
    k = res['kernels'][dom]
    achieved = k['bytes_per_launch'] / k['avg_ms'] / 1e6
    out['roofline'] = {'bound': 'hbm', 'achieved': round(achieved, 1), 'peak': HBM_PEAK_GBPS, 'unit': 'GB/s',
                       'frac': round(achieved / HBM_PEAK_GBPS, 4), 'traffic': None, 'kernel': dom,
                       'avg_ms': k['avg_ms'], 'algorithmic_bytes_per_launch': k['bytes_per_launch']}
    traffic = os.path.join(ROOT, 'profiles', 'k_tile_traffic_n34.json')
    if os.path.exists(traffic):
        t = json.load(open(traffic))
        if t.get('plan') == s.plan_text and dom in t.get('kernels', {}):
            out['roofline']['traffic'] = t['kernels'][dom]['hbm_bytes_per_launch']
    if not args.no_extras:
        try:
            out['deferred'] = run_deferred(s, res, args, torch)
        except Exception as e:  # report, do not hide the headline
            out['deferred'] = {'error': repr(e)}
        log('[bench] deferred done')
        try:
            out['folded'] = run_folded(s, args, torch)
        except Exception as e:  # report, do not hide the headline
            out['folded'] = {'error': repr(e)}
        log('[bench] folded done')
        del s, res
        torch.cuda.empty_cache()
        try:
            out['stream_reference'] = stream_reference(torch)
        except Exception as e:
            out['stream_reference'] = {'error': repr(e)}
        log('[bench] stream reference done')
        try:
            out['scan'] = run_scan(torch)
        except Exception as e:
            out['scan'] = {'error': repr(e)}
        log('[bench] scan done')
        torch.cuda.empty_cache()
        try:
            out['scan_rgs70'] = run_scan_rgs70(torch)
        except Exception as e:
            out['scan_rgs70'] = {'error': repr(e)}
        log('[bench] rank-30 scan done')
        torch.cuda.empty_cache()
        try:
            out['batch'] = run_batch(torch)
        except Exception as e:
            out['batch'] = {'error': repr(e)}
        log('[bench] uneven batch done')
        torch.cuda.empty_cache()
        try:
            out['production'] = run_production(torch)
        except Exception as e:
            out['production'] = {'error': repr(e)}
        log('[bench] production replay done')
        torch.cuda.empty_cache()
        try:
            out['learned'] = run_learned(torch)
        except Exception as e:
            out['learned'] = {'error': repr(e)}
        log('[bench] learned vs uniform done')
    if not args.no_cpu:
        out['cpu_baseline'] = cpu_baseline(args.seed)
        log('[bench] cpu baseline done')
    else:
        out['cpu_baseline'] = None
    print(json.dumps(out), flush=True)


if __name__ == '__main__':
    main()
