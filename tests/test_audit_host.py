"""Production gate and run audit on a small synthetic run: a frozen-file
manifest, four groups written by the production worker (run_group ->
records.produce, with a CPU stand-in engine that consumes the reference's
RNG schedule), their complex128 score logs, then the audit.  The clean run
passes; each deliberate corruption is reported under the reference
auditor's failure code (analyze_born_production.py:348-419 lists the same
corruptions for its self-test); a changed frozen file stops the worker."""
import json

import numpy as np
import pytest

from rgs70_classical_challenge_amd import audit as au
from rgs70_classical_challenge_amd.records import decimal_state

COUNTS = (2, 2, 1, 1)


class StandIn:
    """compute(B bits) / draw(rng) with three uniforms per draw; p = 1, norm = 1."""

    def __init__(self, A, B):
        self.A, self.B = A, B

    def compute(self, b_bits):
        return {'weights': 0.0}

    def draw(self, rng):
        u = [rng.random() for _ in range(3)]
        return int(u[0] * (1 << len(self.A))), 1.0, 1.0


def _run(root):
    n, m = 12, 6
    A, B = list(range(m)), list(range(m, n))
    g = np.zeros((n, n), np.uint8)
    g[np.arange(m), np.arange(m) + m] = 1
    g[0, 3] = g[7, 10] = 1
    g |= g.T
    np.savez(root / 'graph.npz', graph=g, measurement_unitaries=np.repeat(np.eye(2)[None], n, 0), A=A, B=B)
    (root / 'worker.py').write_text('# frozen source stand-in\n')
    seeds = {'total_samples': sum(COUNTS), 'groups': []}
    for gid, count in enumerate(COUNTS):
        rng = np.random.Generator(np.random.PCG64(900 + gid))
        seeds['groups'].append({'group_id': gid, 'sample_count': count, 'global_sample_id_start': sum(COUNTS[:gid]),
                                'global_sample_id_end_inclusive': sum(COUNTS[:gid + 1]) - 1,
                                'cli_seed_decimal': str(900 + gid),
                                'initial_pcg64_state_decimal128': decimal_state(rng.bit_generator.state)})
    (root / 'seeds.json').write_text(json.dumps(seeds))
    manifest = {'synthetic_fixture': True, 'graph': 'graph.npz', 'seed_manifest': 'seeds.json',
                'uniforms_per_sample': 3, 'engine': 'conditional',
                'files': {f: au.sha256(root / f) for f in ('graph.npz', 'seeds.json', 'worker.py')}}
    (root / 'manifest.json').write_text(json.dumps(manifest))
    for gid in range(len(COUNTS)):
        recs = au.run_group(root / 'manifest.json', gid, root / f'born_group{gid}.jsonl', root=root,
                            engine=StandIn(A, B))
        au.write_scores(root / f'born_group{gid}_fp64.jsonl', recs, [r['probability_raw'] for r in recs],
                        [0.01] * len(recs))
    return root


def _audit(root, **kw):
    return au.audit(root, root / 'seeds.json', root / 'manifest.json', root, COUNTS, resamples=32,
                    allow_synthetic=kw.pop('allow_synthetic', True), **kw)


def _codes(rep):
    return [f['code'] for f in rep['failures']]


def test_clean_run_passes_and_corruptions_are_named(tmp_path):
    root = _run(tmp_path)
    rep = _audit(root)
    assert rep['validation_passed'], rep['failures']
    assert rep['matched_scored_samples'] == sum(COUNTS)
    assert all(g['audit_passed'] for g in rep['groups'])
    target = root / 'born_group0.jsonl'
    original = target.read_text()
    mutations = [
        ('uniform', lambda r: r[1]['draw_uniforms'].__setitem__(0, 0.123), 'three_uniforms_rng_replay'),
        ('before', lambda r: r[1]['rngstate_before']['state'].__setitem__('state', '0'), 'rng_before'),
        ('after', lambda r: r[1]['rngstate_after']['state'].__setitem__('inc', '0'), 'rng_after'),
        ('B', lambda r: r[1]['B_bits'].__setitem__(0, 1 - r[1]['B_bits'][0]), 'B_rng_replay'),
        ('global id', lambda r: r[1].__setitem__('global_sample_id', 100), 'sample_or_global_id'),
        ('manifest hash', lambda r: r[0]['metadata'].__setitem__('manifest_sha256', '0' * 64), 'logged_manifest_hash'),
        ('count', lambda r: r.pop(1), 'production_sample_count'),
        ('norm', lambda r: r[1].__setitem__('conditional_norm', 1.01), 'conditional_norm_threshold'),
        ('bitstring', lambda r: r[1].__setitem__('bitstring_q0_left', '1' * 12), 'A_B_to_full_bitstring_mapping'),
        ('footer', lambda r: r[-1]['summary'].__setitem__('all_count_completed', False), 'production_completion_footer'),
    ]
    for name, mutate, code in mutations:
        rows = [json.loads(x) for x in original.splitlines()]
        mutate(rows)
        target.write_text(''.join(json.dumps(x) + '\n' for x in rows))
        rep = _audit(root)
        assert not rep['validation_passed'] and code in _codes(rep), (name, _codes(rep))
        target.write_text(original)
    scores = root / 'born_group0_fp64.jsonl'
    saved = scores.read_text()
    rows = [json.loads(x) for x in saved.splitlines()]
    rows[1]['bitstring_q0_left'] = '0' * 12
    scores.write_text(''.join(json.dumps(x) + '\n' for x in rows))
    assert 'fp64_index_bitstring_mapping' in _codes(_audit(root))
    rows[1]['bitstring_q0_left'] = json.loads(saved.splitlines()[1])['bitstring_q0_left']
    rows[1]['probability'] *= 1.001
    scores.write_text(''.join(json.dumps(x) + '\n' for x in rows))
    assert 'raw_probability_relative_error_ge_1e_4' in _codes(_audit(root))
    scores.write_text(saved)
    src = root / 'worker.py'
    src.write_text(src.read_text() + '# modified\n')
    assert 'frozen_file_hash_mismatch' in _codes(_audit(root))
    with pytest.raises(ValueError, match='Frozen source/input changed: worker.py'):
        au.run_group(root / 'manifest.json', 0, root / 'again.jsonl', root=root, engine=StandIn([0], [1]))
    # a synthetic fixture is refused outside the self-test
    assert 'synthetic_fixture_rejected_in_production' in _codes(_audit(root, allow_synthetic=False))


def test_worker_refuses_existing_output_and_unknown_engine(tmp_path):
    root = _run(tmp_path)
    with pytest.raises(FileExistsError):
        au.run_group(root / 'manifest.json', 0, root / 'born_group0.jsonl', root=root, engine=StandIn([0], [1]))
    m = json.loads((root / 'manifest.json').read_text())
    m['engine'] = 'unvalidated'
    (root / 'manifest.json').write_text(json.dumps(m))
    with pytest.raises(ValueError, match='Unvalidated engine'):
        au.run_group(root / 'manifest.json', 0, root / 'x.jsonl', root=root, engine=StandIn([0], [1]))
