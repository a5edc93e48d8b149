"""Replay of the reference's production conditional Born samples on one device.

The reference draws each 70-qubit sample by fixing 35 uniform B bits, building
the normalised 2^35-amplitude conditional state of A on two GPUs (2^34
complex64 amplitudes each) and drawing A with three uniforms
(work/rgs_attack/born_production.py:44-58).  Its records keep the B bits, the
three uniforms, the drawn A index, the complex64 probability and the
complex128 rescoring of the full string.  This module rebuilds every recorded
sample with ``ConditionalSampler`` on ONE device (the whole 2^35 complex64
state, 256 GiB, in one slab), feeds the recorded uniforms to the draw in
order, and reports per sample: the time of compute + draw, whether the drawn
index equals the recorded one, and the relative error of p at the recorded
string against the complex128 value.

    python -m rgs70_classical_challenge_amd.production graph.npz records.json [count]
"""
import json
import time

import numpy as np


class RecordedUniforms:
    """Stands in for the numpy Generator: .random() returns the recorded
    uniforms in order and refuses to invent more."""

    def __init__(self, values):
        self.values = [float(v) for v in values]
        self.used = 0

    def random(self):
        if self.used >= len(self.values):
            raise RuntimeError('draw asked for more uniforms than were recorded')
        v = self.values[self.used]
        self.used += 1
        return v


def bitstring_of(A, B, a_index, b_bits, n):
    """Qubit-0-left bitstring of a sample (born_production.py:49-51)."""
    bits = [0] * n
    for j, q in enumerate(A):
        bits[q] = (a_index >> j) & 1
    for j, q in enumerate(B):
        bits[q] = int(b_bits[j])
    return ''.join(map(str, bits))


def replay(g, u, A, B, records, tuning=None, log=None):
    from .engine import ConditionalSampler
    t = time.perf_counter()
    eng = ConditionalSampler(g, u, A, B, tuning=tuning)
    setup = time.perf_counter() - t
    m = len(A)
    rows = []
    for rec in records:
        t = time.perf_counter()
        stages = eng.compute(rec['B_bits'])
        rng = RecordedUniforms(rec['draw_uniforms'])
        index, pa, norm = eng.draw(rng)
        elapsed = time.perf_counter() - t
        p_at = eng.probability_at(int(rec['A_index'])) / norm * 2.0 ** (-m)
        row = dict(sample=rec['sample'], seconds=elapsed, stages=stages, uniforms_used=rng.used,
                   margins=list(eng.last_margins),
                   index=index, index_matches=index == int(rec['A_index']),
                   bitstring_matches=bitstring_of(A, B, index, rec['B_bits'], len(g)) == rec['bitstring_q0_left'],
                   norm=norm, norm_recorded=rec['conditional_norm'],
                   p_at_recorded=p_at, p_fp64=rec['p_fp64'], p_recorded=rec['probability'],
                   rel_err_vs_fp64=abs(p_at - rec['p_fp64']) / rec['p_fp64'],
                   rel_err_vs_recorded_c64=abs(p_at - rec['probability']) / rec['probability'])
        rows.append(row)
        if log:
            log(f'[production] sample {rec["sample"]}: {elapsed:.3f} s, index match {row["index_matches"]}, '
                f'rel err vs fp64 {row["rel_err_vs_fp64"]:.2e}')
    secs = np.array([r['seconds'] for r in rows])
    rel = np.array([r['rel_err_vs_fp64'] for r in rows])
    summary = dict(samples=len(rows), setup_seconds=setup, plan=eng.plan_text, passes=eng.passes,
                   median_seconds_per_sample=float(np.median(secs)), min_seconds=float(secs.min()),
                   max_seconds=float(secs.max()), index_matches=int(sum(r['index_matches'] for r in rows)),
                   max_rel_err_vs_fp64=float(rel.max()), median_rel_err_vs_fp64=float(np.median(rel)),
                   max_norm_diff=float(max(abs(r['norm'] - r['norm_recorded']) for r in rows)))
    return summary, rows


def main(argv):
    import sys
    d = np.load(argv[0])
    data = json.load(open(argv[1]))
    count = int(argv[2]) if len(argv) > 2 else len(data['records'])
    A, B = data['A_vertices'], data['B_vertices']
    if list(d['A']) != list(A) or list(d['B']) != list(B):
        raise ValueError('records and graph disagree on the partition')
    summary, rows = replay(d['graph'], d['measurement_unitaries'], A, B, data['records'][:count],
                           log=lambda *a: print(*a, file=sys.stderr, flush=True))
    summary['reference_median_seconds_per_sample'] = 2.4162
    summary['reference_hardware'] = '2 x H200 (2^34 complex64 per GPU)'
    print(json.dumps({'summary': summary, 'rows': rows}), flush=True)


if __name__ == '__main__':
    import sys
    main(sys.argv[1:])
