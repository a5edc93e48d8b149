"""Production record wire format, RNG replay verification and the audit's
score statistics, checked against the reference's own records and audit
summary (work/rgs_attack/born_production.py:37-61, born/verify.py:58-86,
analyze_born_production.py:48-70)."""
import copy
import json

import numpy as np
import pytest

from rgs70_classical_challenge_amd import records as R


def _golden(golden_dir):
    d = json.load(open(f'{golden_dir}/rgs70_born_records.json'))
    return d['group'], d['A_vertices'], d['B_vertices'], d['records']


def test_recorded_group_replays_clean(golden_dir):
    group, A, B, recs = _golden(golden_dir)
    failures, _, _, norm_err = R.verify_group(recs, group, A, B, 70)
    assert failures == []
    assert max(norm_err) < 1e-6
    # the generator of the group reproduces the recorded initial state
    rng = R.group_generator(group)
    assert R.same_state(rng.bit_generator.state, recs[0]['rngstate_before'])
    assert R.decimal_state(rng.bit_generator.state)['state'] == recs[0]['rngstate_before']['state']


def test_replay_names_each_tampered_field(golden_dir):
    group, A, B, recs = _golden(golden_dir)
    cases = []
    bad = copy.deepcopy(recs)
    bad[3]['B_bits'][0] ^= 1
    cases.append((bad, ['sample 3: RNG replay', 'sample 3: bit ordering or score alignment']))
    bad = copy.deepcopy(recs)
    bad[5]['A_index'] ^= 4
    cases.append((bad, ['sample 5: bit ordering or score alignment']))
    bad = copy.deepcopy(recs)
    bad[7]['rngstate_before']['state']['state'] = str(int(bad[7]['rngstate_before']['state']['state']) + 1)
    cases.append((bad, ['sample 7: RNG state before']))
    bad = copy.deepcopy(recs)
    bad[2]['draw_uniforms'][2] = bad[2]['draw_uniforms'][2] + 2.0 ** -40
    cases.append((bad, ['sample 2: RNG replay']))
    for tampered, want in cases:
        failures, _, _, _ = R.verify_group(tampered, group, A, B, 70, label='g0')
        assert failures == [f'g0 {w}' for w in want], failures
    bad_group = dict(group, cli_seed_decimal=str(int(group['cli_seed_decimal']) + 1))
    failures, _, _, _ = R.verify_group(recs, bad_group, A, B, 70, label='g0')
    assert failures == ['g0: initial PCG64 state mismatch']


def test_scores_align_and_give_z(golden_dir):
    group, A, B, recs = _golden(golden_dir)
    scores = [dict(index=k, bitstring_q0_left=r['bitstring_q0_left'], probability=r['p_fp64'])
              for k, r in enumerate(recs)]
    failures, z, rel, _ = R.verify_group(recs, group, A, B, 70, scores=scores)
    assert failures == []
    assert np.allclose(z, [2.0 ** 70 * r['p_fp64'] for r in recs], rtol=0, atol=0)
    # the reference's complex64 engine against its complex128 rescoring
    assert max(rel) < 3.4e-6
    scores[4] = dict(scores[4], index=9)
    failures, _, _, _ = R.verify_group(recs, group, A, B, 70, scores=scores)
    assert failures == ['group sample 4: bit ordering or score alignment']


def test_statistics_match_recorded_audit(golden_dir):
    """Every field of the audit's group-0 score summary, the seeded bootstrap
    interval included, from the 432 recorded scores."""
    z = np.array(json.load(open(f'{golden_dir}/rgs70_batch1_group0_z.json'))['z'])
    rec = json.load(open(f'{golden_dir}/rgs70_batch1_summary.json'))['groups'][0]['score']
    st = R.statistics(z - 1)
    for key in ('n', 'bootstrap_resamples', 'bootstrap_seed'):
        assert st[key] == rec[key], key
    for key in ('mean', 'standard_error', 'min', 'max', 'clipped_K4_mean',
                'clipped_K4_one_sided_Hoeffding_lower_alpha_0_05'):
        assert st[key] == pytest.approx(rec[key], rel=1e-12, abs=1e-15), key
    for key in ('normal_95_CI_approximate', 'bootstrap_95_CI_approximate'):
        assert st[key] == pytest.approx(rec[key], rel=1e-12), key


def test_produce_refuses_to_overwrite(tmp_path):
    out = tmp_path / 'g.jsonl'
    out.write_text('{}\n')

    class Engine:
        A, B = [0], [1]

    group = dict(group_id=0, sample_count=1, global_sample_id_start=0, cli_seed_decimal='1')
    with pytest.raises(FileExistsError):
        R.produce(Engine(), 2, group, out)


def test_capped_statistics_match_recorded_analysis(golden_dir):
    """The capped-run analysis of the cap-8 run's first part: 863 accepted
    strings, every recorded field of its score summary from their FP64
    probabilities, bootstrap interval included."""
    d = json.load(open(f'{golden_dir}/rgs70_cap8_part0_scores.json'))
    x = 2.0 ** 70 * np.array(d['p_fp64']) - 1
    assert np.allclose(x, d['xeb_fp64'], rtol=1e-12, atol=1e-12)
    st = R.capped_statistics(x)
    rec = d['score']
    for key in ('n', 'bootstrap_resamples', 'bootstrap_seed'):
        assert st[key] == rec[key], key
    for key in ('mean', 'standard_error', 'min', 'max', 'clipped_K4_mean',
                'clipped_K4_one_sided_Hoeffding_lower_alpha_0_025'):
        assert st[key] == pytest.approx(rec[key], rel=1e-12, abs=1e-15), key
    for key in ('normal_95_CI', 'bootstrap_95_CI'):
        assert st[key] == pytest.approx(rec[key], rel=1e-12), key
