"""Frozen-input gate of a production run and the independent audit of its logs.

* ``check_frozen`` is the gate a production worker passes before it draws
  anything (born_production.py:19-21): every file the manifest lists must
  still have its recorded SHA-256, else ``ValueError('Frozen source/input
  changed: <path>')``.
* ``run_group`` is the worker itself (born_production.py:17-61): gate, the
  group's seed entry and initial PCG64 state, the graph, the engine named by
  the manifest, then ``records.produce`` into a new JSONL file.
* ``audit`` re-reads a finished run - the seed and production manifests, the
  frozen files, the graph and partition, every group's production and
  complex128-score logs - and reports each problem as a failure code
  (analyze_born_production.py:81-345).  The codes are the reference
  auditor's names, so reports of the two can be compared line by line; the
  checks are restated here from what each code asserts.  Sizes come from the
  graph (n vertices, |A| = |B| = n / 2) instead of being fixed at 70 / 35.

The audit does not recompute probabilities: it checks RNG replay, the A/B to
bitstring mapping, the logged norms and probability identities, the
generated-versus-complex128 relative error (below 1e-4, the manifest's
threshold), the timing order, and the score statistics of the whole run.
"""
import hashlib
import json
import math
from pathlib import Path

import numpy as np

from .records import BOOTSTRAPS, decimal_state, produce, same_state, statistics

REL_ERROR_MAX = 1e-4
NORM_ERROR_MAX = 1e-3
ENGINES = {'conditional': 'ConditionalSampler',
           # the reference's engine names, both served by the same engine here
           'specialized_deferred': 'ConditionalSampler', 'deferred_fused': 'ConditionalSampler'}


def sha256(path):
    return hashlib.sha256(Path(path).read_bytes()).hexdigest()


def check_frozen(manifest_path, root='.'):
    """The production gate: (manifest, its SHA-256) if every listed file still
    hashes to its recorded digest; raises ValueError naming the first file that
    changed (FileNotFoundError if one is missing)."""
    manifest_path = Path(manifest_path)
    manifest = json.loads(manifest_path.read_text())
    files = manifest.get('files') or {}
    if not files:
        raise ValueError('manifest lists no frozen files')
    for name, digest in files.items():
        if sha256(Path(root) / name) != digest:
            raise ValueError(f'Frozen source/input changed: {name}')
    return manifest, sha256(manifest_path)


def run_group(manifest_path, group_id, out, root='.', engine=None, log=None):
    """One production worker: the gate, then ``group_id``'s samples into the
    new JSONL file ``out`` (born_production.py:17-61).  ``engine`` overrides the
    manifest's engine (an object with compute / draw / A / B, for tests)."""
    import time
    process_start = time.perf_counter()
    root = Path(root)
    manifest, digest = check_frozen(manifest_path, root)
    seeds = json.loads((root / manifest['seed_manifest']).read_text())
    group = seeds['groups'][group_id]
    if group['group_id'] != group_id:
        raise ValueError('Group mismatch')
    data = np.load(root / manifest['graph'], allow_pickle=False)
    g = data['graph']
    if manifest.get('engine') not in ENGINES:
        raise ValueError('Unvalidated engine')
    t = time.perf_counter()
    if engine is None:
        from .engine import ConditionalSampler
        engine = ConditionalSampler(g, data['measurement_unitaries'], data['A'].tolist(), data['B'].tolist())
    setup = time.perf_counter() - t
    return produce(engine, len(g), group, out, meta={'manifest_sha256': digest, 'engine': manifest.get('engine')},
                   log=log, setup_seconds=setup, process_start=process_start)


def write_scores(path, records, probabilities, seconds, setup_seconds=0.0):
    """Complex128 score log of a group in the reference's layout: a metadata
    line, then per sample its index, bitstring, probability and seconds."""
    with Path(path).open('x') as f:
        f.write(json.dumps({'metadata': {'dtype': 'complex128', 'reverse_order': True,
                                         'n': len(records[0]['bitstring_q0_left']) if records else 0,
                                         'setup_seconds': setup_seconds}}) + '\n')
        for k, (rec, p, s) in enumerate(zip(records, probabilities, seconds)):
            f.write(json.dumps({'index': k, 'bitstring_q0_left': rec['bitstring_q0_left'],
                                'probability': float(p), 'seconds': float(s)}) + '\n')


def _gf2_rank(m):
    m = np.asarray(m, np.uint8).copy()
    r = 0
    for c in range(m.shape[1]):
        nz = np.flatnonzero(m[r:, c])
        if not len(nz):
            continue
        p = r + int(nz[0])
        m[[r, p]] = m[[p, r]]
        below = np.flatnonzero(m[r + 1:, c]) + r + 1
        m[below] ^= m[r]
        r += 1
        if r == m.shape[0]:
            break
    return r


def audit(directory, seed_path, manifest_path, source_root, expected_counts, resamples=BOOTSTRAPS,
          allow_synthetic=False, global_id_offset=0):
    """Audit a finished production run; returns a report whose ``failures`` is a
    list of {'code', 'group_id'?, 'sample'?, 'detail'?} and whose
    ``validation_passed`` is True only when it is empty."""
    directory, source_root = Path(directory), Path(source_root)
    seed_path, manifest_path = Path(seed_path), Path(manifest_path)
    failures, hashes, table, groups = [], {}, [], []

    def fail(code, gid=None, sample=None, detail=None):
        item = {'code': code}
        if gid is not None:
            item['group_id'] = gid
        if sample is not None:
            item['sample'] = sample
        if detail is not None:
            item['detail'] = str(detail)
        failures.append(item)

    def check(ok, code, gid=None, sample=None, detail=None):
        if not ok:
            fail(code, gid, sample, detail)
        return ok

    def load_json(path):
        try:
            hashes[str(path)] = sha256(path)
            return json.loads(Path(path).read_text())
        except (OSError, ValueError) as e:
            fail('unreadable_json', detail=f'{path}: {e}')
            return {}

    def load_lines(path):
        try:
            hashes[str(path)] = sha256(path)
            return [json.loads(line) for line in Path(path).read_text().splitlines() if line.strip()]
        except (OSError, ValueError) as e:
            fail('unreadable_jsonl', detail=f'{path}: {e}')
            return []

    seeds, manifest = load_json(seed_path), load_json(manifest_path)
    total = sum(expected_counts)
    ngroups = len(expected_counts)
    check(allow_synthetic or not manifest.get('synthetic_fixture'), 'synthetic_fixture_rejected_in_production')
    check(seeds.get('total_samples') == total, 'seed_manifest_total_count')
    check(len(seeds.get('groups', [])) == ngroups, 'seed_manifest_group_count')
    try:
        entries = seeds.get('groups', [])
        cli = [int(e['cli_seed_decimal']) for e in entries]
        init = [e['initial_pcg64_state_decimal128']['state'] for e in entries]
        check(len(set(cli)) == len(cli) == ngroups, 'group_rng_seeds_not_distinct')
        check(len({str(s['inc']) for s in init}) == ngroups, 'group_PCG64_stream_increments_not_distinct')
        check(len({(str(s['state']), str(s['inc'])) for s in init}) == ngroups,
              'group_PCG64_initial_states_not_distinct')
    except (KeyError, TypeError, ValueError):
        fail('malformed_group_seed_or_initial_state')
    files = manifest.get('files', {})
    check(bool(files), 'empty_frozen_source_manifest')
    for name, digest in files.items():
        try:
            actual = sha256(source_root / name)
            hashes[str(source_root / name)] = actual
            check(actual == digest, 'frozen_file_hash_mismatch', detail=name)
        except OSError as e:
            fail('frozen_file_missing', detail=f'{name}: {e}')
    bound = manifest.get('seed_manifest_sha256') or files.get(manifest.get('seed_manifest', ''))
    check(bound == hashes.get(str(seed_path)), 'seed_manifest_not_bound_by_production_manifest')
    check(manifest.get('uniforms_per_sample') == 3, 'production_manifest_uniform_schedule_missing_or_not_3')
    A = B = None
    n = None
    try:
        with np.load(source_root / manifest.get('graph', 'rgs_graph.npz'), allow_pickle=False) as gr:
            A, B, adj = gr['A'].astype(int).tolist(), gr['B'].astype(int).tolist(), gr['graph']
        n = len(adj)
        part = check(len(A) == len(B) == n // 2 and sorted(A + B) == list(range(n)), 'invalid_frozen_partition')
        binary = check(adj.shape == (n, n) and np.array_equal(adj, adj.T) and not np.diag(adj).any()
                       and np.isin(adj, [0, 1]).all(), 'invalid_binary_graph')
        if part and binary:
            check(_gf2_rank(adj[np.ix_(A, B)]) == len(B), 'cross_partition_GF2_rank_not_full')
        check(manifest.get('graph') in files, 'graph_not_hash_bound')
    except (OSError, ValueError, KeyError) as e:
        fail('unreadable_graph_partition', detail=e)
    m = len(B) if B is not None else 0
    global_ids = []
    for gid, count in enumerate(expected_counts):
        first_failure = len(failures)
        start_id = global_id_offset + sum(expected_counts[:gid])
        entry = seeds.get('groups', [])[gid] if gid < len(seeds.get('groups', [])) else {}
        check(entry.get('group_id') == gid, 'seed_group_id', gid)
        check(entry.get('sample_count') == count, 'seed_group_count', gid)
        check(entry.get('global_sample_id_start') == start_id and
              entry.get('global_sample_id_end_inclusive') == start_id + count - 1, 'seed_global_id_range', gid)
        prod = load_lines(directory / f'born_group{gid}.jsonl')
        scored = load_lines(directory / f'born_group{gid}_fp64.jsonl')
        metas = [r['metadata'] for r in prod if isinstance(r, dict) and 'metadata' in r]
        ends = [r['summary'] for r in prod if isinstance(r, dict) and 'summary' in r]
        rows = [r for r in prod if isinstance(r, dict) and 'sample' in r]
        smetas = [r['metadata'] for r in scored if isinstance(r, dict) and 'metadata' in r]
        scores = [r for r in scored if isinstance(r, dict) and 'probability' in r]
        check(len(metas) == 1 and len(ends) == 1, 'production_header_footer_count', gid)
        check(len(smetas) == 1, 'scoring_header_count', gid)
        check(len(rows) == count, 'production_sample_count', gid, detail=len(rows))
        check(len(scores) == count, 'fp64_sample_count', gid, detail=len(scores))
        meta = metas[0] if metas else {}
        end = ends[0] if ends else {}
        smeta = smetas[0] if smetas else {}
        check(meta.get('group') == entry, 'logged_seed_group_not_manifest_group', gid)
        check(meta.get('manifest_sha256') == hashes.get(str(manifest_path)), 'logged_manifest_hash', gid)
        check(meta.get('engine') == manifest.get('engine'), 'logged_engine', gid)
        check(meta.get('A_vertices') == A and meta.get('B_vertices') == B and meta.get('conditional_qubits') == m,
              'logged_partition_mapping', gid)
        check('3 uniform' in meta.get('rng_schedule', ''), 'logged_rng_schedule', gid)
        check(smeta.get('dtype') == 'complex128' and smeta.get('reverse_order') is True and smeta.get('n') == n,
              'scoring_not_reverse_complex128', gid)
        check(end.get('samples') == count and end.get('group_id') == gid and end.get('all_count_completed') is True,
              'production_completion_footer', gid)
        try:
            rng = np.random.Generator(np.random.PCG64(int(entry['cli_seed_decimal'])))
            check(same_state(rng.bit_generator.state, entry['initial_pcg64_state_decimal128']), 'initial_rng_state', gid)
        except (KeyError, ValueError, TypeError) as e:
            fail('invalid_rng_manifest', gid, detail=e)
            rng = None
        sample_seconds = 0.0
        gtable = []
        for k, row in enumerate(rows):
            try:
                sid = row['global_sample_id']
                global_ids.append(sid)
                check(row['sample'] == k and row['group_id'] == gid and sid == start_id + k, 'sample_or_global_id',
                      gid, k)
                if rng is not None:
                    check(same_state(row['rngstate_before'], decimal_state(rng.bit_generator.state)), 'rng_before', gid, k)
                    want_b = rng.integers(0, 2, m).tolist()
                    want_u = [float(rng.random()) for _ in range(3)]
                    check(row['B_bits'] == want_b, 'B_rng_replay', gid, k)
                    check(row['draw_uniforms'] == want_u, 'three_uniforms_rng_replay', gid, k)
                    check(same_state(row['rngstate_after'], decimal_state(rng.bit_generator.state)), 'rng_after', gid, k)
                bits, ai, bb = row['bitstring_q0_left'], row['A_index'], row['B_bits']
                ok_bits = isinstance(bits, str) and len(bits) == n and not set(bits) - set('01')
                ok_ai = isinstance(ai, int) and 0 <= ai < 2 ** m
                ok_bb = isinstance(bb, list) and len(bb) == m and all(v in (0, 1) for v in bb)
                if check(ok_bits and ok_ai and ok_bb, 'invalid_bitstring_or_index', gid, k) and A is not None:
                    full = [0] * n
                    for j, q in enumerate(A):
                        full[q] = (ai >> j) & 1
                    for j, q in enumerate(B):
                        full[q] = bb[j]
                    check(bits == ''.join(map(str, full)), 'A_B_to_full_bitstring_mapping', gid, k)
                raw, normed, norm = (float(row[f]) for f in ('probability_raw', 'probability', 'conditional_norm'))
                if not check(all(math.isfinite(v) and v > 0 for v in (raw, normed, norm)) and raw <= 1 and normed <= 1,
                             'invalid_generated_probability_or_norm', gid, k):
                    continue
                check(abs(norm - 1) < NORM_ERROR_MAX, 'conditional_norm_threshold', gid, k)
                check(math.isclose(normed, raw / norm, rel_tol=5e-15, abs_tol=0), 'normalized_probability_identity',
                      gid, k)
                if k >= len(scores):
                    continue
                sc = scores[k]
                check(sc.get('index') == k and sc.get('bitstring_q0_left') == bits, 'fp64_index_bitstring_mapping',
                      gid, k)
                p64, secs = float(sc['probability']), float(sc['seconds'])
                if not check(math.isfinite(p64) and 0 < p64 <= 1 and math.isfinite(secs) and secs >= 0,
                             'invalid_fp64_probability_or_time', gid, k):
                    continue
                raw_rel, norm_rel = abs(raw - p64) / p64, abs(normed - p64) / p64
                check(raw_rel < REL_ERROR_MAX, 'raw_probability_relative_error_ge_1e_4', gid, k, raw_rel)
                elapsed = float(row['elapsed_seconds'])
                if not check(math.isfinite(elapsed) and elapsed >= 0, 'invalid_sample_time', gid, k):
                    continue
                sample_seconds += elapsed
                item = {'group_id': gid, 'sample': k, 'global_sample_id': sid, 'bitstring_q0_left': bits,
                        'p_raw': raw, 'p_normalized': normed, 'p_fp64': p64, 'norm_absolute_error': abs(norm - 1),
                        'raw_relative_error': raw_rel, 'normalized_relative_error': norm_rel,
                        'xeb_fp64': 2.0 ** n * p64 - 1, 'generation_sample_seconds': elapsed, 'fp64_seconds': secs}
                gtable.append(item)
                table.append(item)
            except (KeyError, TypeError, ValueError, IndexError, ZeroDivisionError) as e:
                fail('malformed_sample_record', gid, k, e)
        timing = {}
        try:
            gen, proc, setup = (float(end[f]) for f in ('generation_seconds', 'process_elapsed_seconds',
                                                        'setup_seconds'))
            if check(all(math.isfinite(v) and v >= 0 for v in (gen, proc, setup)), 'invalid_group_time', gid):
                check(proc + 1e-3 >= gen + setup, 'group_time_order', gid)
                check(gen + 1e-3 >= sample_seconds, 'generation_less_than_sample_sum', gid)
                check(math.isclose(setup, float(meta['setup_seconds']), rel_tol=0, abs_tol=1e-12),
                      'setup_header_footer_mismatch', gid)
                timing = {'generation_seconds': gen, 'process_elapsed_seconds': proc, 'setup_seconds': setup}
        except (KeyError, TypeError, ValueError) as e:
            fail('malformed_group_timing', gid, detail=e)
        groups.append({'group_id': gid, 'expected_count': count, 'observed_count': len(rows),
                       'fp64_count': len(scores), 'audit_passed': len(failures) == first_failure,
                       'score': statistics([x['xeb_fp64'] for x in gtable], resamples), 'timing': timing})
    check(global_ids == list(range(global_id_offset, global_id_offset + total)),
          'complete_ordered_global_ids_expected_offset')
    check(len(table) == total, 'complete_fp64_scored_count', detail=len(table))
    return {'validation_passed': not failures, 'expected_samples': total, 'matched_scored_samples': len(table),
            'failures': failures, 'groups': groups,
            'score': statistics([x['xeb_fp64'] for x in table], resamples),
            'errors': {k: (float(np.max([x[k] for x in table])) if table else None)
                       for k in ('norm_absolute_error', 'raw_relative_error', 'normalized_relative_error')},
            'duplicate_complete_bitstrings': len(table) - len({x['bitstring_q0_left'] for x in table}),
            'input_sha256': hashes}
