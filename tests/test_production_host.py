"""Host logic of the production replay: the recorded-uniform stream and the
bitstring assembly of a conditional sample (born_production.py:47-51)."""
import json

import pytest

from rgs70_classical_challenge_amd.production import RecordedUniforms, bitstring_of


def test_recorded_uniforms_replay_in_order_and_refuse_more():
    r = RecordedUniforms([0.25, 0.5, 0.75])
    assert [r.random(), r.random(), r.random()] == [0.25, 0.5, 0.75]
    assert r.used == 3
    with pytest.raises(RuntimeError):
        r.random()


def test_bitstrings_of_recorded_samples(golden_dir):
    data = json.load(open(f'{golden_dir}/rgs70_born_records.json'))
    A, B = data['A_vertices'], data['B_vertices']
    assert sorted(A + B) == list(range(70))
    for rec in data['records']:
        assert len(rec['draw_uniforms']) == 3 and len(rec['B_bits']) == 35
        assert 0 <= rec['A_index'] < 1 << 35
        assert bitstring_of(A, B, rec['A_index'], rec['B_bits'], 70) == rec['bitstring_q0_left']
    # a flipped A bit moves exactly one character
    rec = data['records'][0]
    s = bitstring_of(A, B, rec['A_index'] ^ 1, rec['B_bits'], 70)
    diff = [i for i, (a, b) in enumerate(zip(s, rec['bitstring_q0_left'])) if a != b]
    assert diff == [A[0]]
