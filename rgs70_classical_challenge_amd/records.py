"""Production records: the reference's JSONL wire format, its RNG replay
verification and its score statistics.

A production run writes one JSON object per line (born_production.py:37-61):
``{"metadata": ...}``, then one record per sample, then ``{"summary": ...}``.
Each sample record holds the PCG64 state before and after the sample (state
and increment as decimal strings), the 35 B bits drawn with
``rng.integers(0, 2, m)``, exactly three draw uniforms, the A index, the
qubit-0-left bitstring, the raw and normalised probabilities, the
conditional norm, the elapsed time and the engine's stage times.

``verify_group`` replays a group's RNG stream from its seed and checks every
record against it and against an optional list of complex128 scores, with the
reference verifier's failure messages (born/verify.py:58-86).
``statistics`` is the audit's summary of a score list
(analyze_born_production.py:48-70): mean, standard error, normal and seeded
bootstrap 95% intervals, and the clipped (K = 4) one-sided Hoeffding lower
bound.
"""
import copy
import json
import math
import os
import time
from pathlib import Path

import numpy as np

BOOTSTRAPS = 20000
BOOTSTRAP_SEED = 424242


def decimal_state(state):
    """PCG64 state with the 128-bit words as decimal strings (born_production.py:8-11)."""
    s = copy.deepcopy(state)
    for k in ('state', 'inc'):
        s['state'][k] = str(int(s['state'][k]))
    return s


def same_state(state, reference):
    """Equal PCG64 states whatever the integer encoding (born/verify.py:31-33)."""
    return all(int(state['state'][k]) == int(reference['state'][k]) for k in ('state', 'inc')) and \
        int(state['has_uint32']) == int(reference['has_uint32']) and \
        int(state['uinteger']) == int(reference['uinteger'])


class DrawRecorder:
    """Hands the engine's draw the generator's uniforms and keeps them
    (born_production.py:13-16)."""

    def __init__(self, rng):
        self.rng = rng
        self.uniforms = []

    def random(self):
        v = float(self.rng.random())
        self.uniforms.append(v)
        return v


def bitstring_q0_left(A, B, a_index, b_bits, n):
    bits = [0] * n
    for j, q in enumerate(A):
        bits[q] = (int(a_index) >> j) & 1
    for j, q in enumerate(B):
        bits[q] = int(b_bits[j])
    return ''.join(map(str, bits))


def group_generator(group):
    """The group's PCG64 generator, checked against its recorded initial state."""
    rng = np.random.Generator(np.random.PCG64(int(group['cli_seed_decimal'])))
    init = group.get('initial_pcg64_state_decimal128')
    if init is not None and not same_state(rng.bit_generator.state, init):
        raise ValueError('Initial RNG mismatch')
    return rng


def produce(engine, n, group, out, count=None, meta=None, log=None, setup_seconds=0.0, process_start=None):
    """Draw ``count`` (default: the group's sample_count) samples with a
    conditional engine (ConditionalSampler) on the group's RNG stream and write
    the JSONL records to ``out``, which must not exist.  ``setup_seconds`` (the
    engine's construction) and the process clock start (perf_counter) go
    into the header and footer.  Returns the records."""
    process_start = time.perf_counter() if process_start is None else process_start
    A, B = engine.A, engine.B
    count = group['sample_count'] if count is None else count
    rng = group_generator(group)
    out = Path(out)
    out.parent.mkdir(parents=True, exist_ok=True)
    if out.exists():
        raise FileExistsError('Production output already exists; do not overwrite or silently resume')
    metadata = dict(group=group, engine=type(engine).__name__, numpy=np.__version__, conditional_qubits=len(A),
                    setup_seconds=setup_seconds, process_to_ready_seconds=time.perf_counter() - process_start,
                    B_vertices=list(B), A_vertices=list(A),
                    rng_schedule='fresh B integers, then 3 uniform draws; no warmup consumes production RNG')
    metadata.update(meta or {})
    records = []
    with out.open('x') as f:
        f.write(json.dumps({'metadata': metadata}) + '\n')
        t0 = time.perf_counter()
        wall0 = time.time()
        for i in range(count):
            before = decimal_state(rng.bit_generator.state)
            start = time.perf_counter()
            b_bits = rng.integers(0, 2, len(B)).tolist()
            stages = engine.compute(b_bits)
            draws = DrawRecorder(rng)
            index, pa, norm = engine.draw(draws)
            if len(draws.uniforms) != 3:
                raise AssertionError('RNG call schedule changed')
            if not (0 <= index < (1 << len(A)) and np.isfinite(pa) and pa > 0 and abs(norm - 1) < 1e-3):
                raise ArithmeticError('Invalid production result')
            rec = dict(sample=i, group_id=group['group_id'], global_sample_id=group['global_sample_id_start'] + i,
                       bitstring_q0_left=bitstring_q0_left(A, B, index, b_bits, n), B_bits=b_bits, A_index=int(index),
                       rngstate_before=before, rngstate_after=decimal_state(rng.bit_generator.state),
                       draw_uniforms=draws.uniforms, probability_raw=pa * 2.0 ** (-len(B)),
                       probability=pa / norm * 2.0 ** (-len(B)), conditional_norm=norm,
                       elapsed_seconds=time.perf_counter() - start, stages=stages)
            f.write(json.dumps(rec) + '\n')
            f.flush()
            # the returned copy also carries the draw's bin-edge margins (not
            # part of the wire format)
            records.append(dict(rec, draw_margins=list(getattr(engine, 'last_margins', []))))
            if log:
                log(f'[records] sample {i}: {rec["elapsed_seconds"]:.3f} s')
        summary = dict(samples=count, group_id=group['group_id'], generation_start_unix=wall0,
                       generation_end_unix=time.time(), generation_seconds=time.perf_counter() - t0,
                       process_elapsed_seconds=time.perf_counter() - process_start, setup_seconds=setup_seconds,
                       all_count_completed=True)
        f.write(json.dumps({'summary': summary}) + '\n')
        f.flush()
        os.fsync(f.fileno())
    return records


def read_jsonl(path):
    return [json.loads(line) for line in Path(path).read_text().splitlines() if line.strip()]


def verify_group(records, group, A, B, n, scores=None, label='group'):
    """Replay the group's RNG stream over its sample records and check each one:
    RNG state before, the B bits, the three uniforms and the state after, the
    bitstring against A_index / B_bits (and the score's, when scores are
    given), and the global sample id.  Returns (failures, z, rel, norm_err)
    with z = 2^n * score probability, rel = relative deviation of the record's
    probability from the score's, norm_err = |conditional_norm - 1|."""
    failures = []
    try:
        rng = group_generator(group)
    except ValueError:
        return [f'{label}: initial PCG64 state mismatch'], [], [], []
    samples = [r for r in records if 'sample' in r]
    if scores is not None and len(scores) != len(samples):
        failures.append(f'{label}: count mismatch')
    z, rel, norm_err = [], [], []
    for k, rec in enumerate(samples):
        where = f'{label} sample {k}'
        if not same_state(rng.bit_generator.state, rec['rngstate_before']):
            failures.append(f'{where}: RNG state before')
        b_bits = rng.integers(0, 2, len(B)).tolist()
        uniforms = [float(rng.random()) for _ in range(3)]
        if b_bits != rec['B_bits'] or uniforms != rec['draw_uniforms'] or \
                not same_state(rng.bit_generator.state, rec['rngstate_after']):
            failures.append(f'{where}: RNG replay')
        bits = bitstring_q0_left(A, B, rec['A_index'], rec['B_bits'], n)
        score = scores[k] if scores is not None and k < len(scores) else None
        if bits != rec['bitstring_q0_left'] or (score is not None and (bits != score['bitstring_q0_left']
                                                                       or score['index'] != k)):
            failures.append(f'{where}: bit ordering or score alignment')
        if 'global_sample_id' in rec and rec['global_sample_id'] != group['global_sample_id_start'] + k:
            failures.append(f'{where}: global sample id')
        if score is not None:
            z.append(2.0 ** n * score['probability'])
            rel.append(abs(rec['probability'] - score['probability']) / score['probability'])
        norm_err.append(abs(rec['conditional_norm'] - 1))
    return failures, z, rel, norm_err


def statistics(values, resamples=BOOTSTRAPS, seed=BOOTSTRAP_SEED):
    """Audit summary of per-sample scores x (x = 2^n p - 1 for linear XEB):
    mean, standard error, normal 95% interval, bootstrap 95% interval (PCG64
    seeded, resamples drawn 128 at a time), the mean of min(x + 1, 4) - 1 and
    its one-sided Hoeffding lower bound at alpha = 0.05."""
    x = np.asarray(values, dtype=float)
    if not len(x):
        return {'n': 0}
    rng = np.random.Generator(np.random.PCG64(seed))
    boot = []
    for start in range(0, resamples, 128):
        idx = rng.integers(0, len(x), size=(min(128, resamples - start), len(x)))
        boot.extend(x[idx].mean(axis=1).tolist())
    mean = float(x.mean())
    se = float(x.std(ddof=1) / math.sqrt(len(x))) if len(x) > 1 else None
    clipped = float(np.minimum(x + 1, 4).mean() - 1)
    lower = clipped - 4 * math.sqrt(math.log(20) / (2 * len(x)))
    return {'n': int(len(x)), 'mean': mean, 'standard_error': se,
            'normal_95_CI_approximate': [mean - 1.96 * se, mean + 1.96 * se] if se is not None else None,
            'bootstrap_95_CI_approximate': np.quantile(boot, [.025, .975]).tolist(),
            'bootstrap_resamples': resamples, 'bootstrap_seed': seed,
            'min': float(x.min()), 'max': float(x.max()),
            'clipped_K4_mean': clipped, 'clipped_K4_one_sided_Hoeffding_lower_alpha_0_05': lower}


def capped_statistics(values, seed=BOOTSTRAP_SEED, resamples=BOOTSTRAPS):
    """Capped-run summary of the accepted strings' scores x = 2^n p - 1
    (analyze_challenge.py:11-15): mean, standard error, normal and seeded
    bootstrap 95% intervals (numpy default_rng, resamples drawn 256 at a
    time), and the clipped (K = 4) one-sided Hoeffding lower bound at
    alpha = 0.025 (two predeclared runs)."""
    x = np.asarray(values, dtype=np.float64)
    if not len(x):
        return {'n': 0}
    mean = float(x.mean())
    se = float(x.std(ddof=1) / np.sqrt(len(x))) if len(x) > 1 else None
    rng = np.random.default_rng(seed)
    boot = []
    for start in range(0, resamples, 256):
        boot.extend(x[rng.integers(0, len(x), size=(min(256, resamples - start), len(x)))].mean(axis=1).tolist())
    clipped = np.minimum(x + 1, 4) - 1
    return {'n': int(len(x)), 'mean': mean, 'standard_error': se,
            'normal_95_CI': [mean - 1.96 * se, mean + 1.96 * se] if se is not None else None,
            'bootstrap_95_CI': np.quantile(boot, [.025, .975]).tolist(), 'bootstrap_resamples': resamples,
            'bootstrap_seed': seed, 'min': float(x.min()), 'max': float(x.max()),
            'clipped_K4_mean': float(clipped.mean()),
            'clipped_K4_one_sided_Hoeffding_lower_alpha_0_025':
                float(clipped.mean() - 4 * np.sqrt(np.log(40) / (2 * len(x))))}
