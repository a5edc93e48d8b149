"""Extract small golden fixtures from the reference's data files.

Reads only data (npz / npy / json / jsonl / the circuit's .qasm text) under /root/reference with safe
loaders (numpy.load with allow_pickle=False, json); no reference code is
imported or run.  Writes into tests/golden/.  Re-run from the repository root:

    python tests/golden/make_golden.py
"""
import json
import shutil
from pathlib import Path

import numpy as np

REF = Path('/root/reference')
WORK = REF / 'work' / 'rgs_attack'
OUT = Path(__file__).resolve().parent


def jsonl(path):
    return [json.loads(line) for line in Path(path).read_text().splitlines() if line.strip()]


def main():
    # graphs, unitaries and partitions (graph_sampler.py:170 writer format)
    for name in ('induced_12q.npz', 'induced_24q.npz', 'induced_32q.npz', 'induced_40q.npz', 'rgs_graph.npz'):
        d = np.load(WORK / name, allow_pickle=False)
        np.savez(OUT / name, **{k: d[k] for k in d.files})
    for name in ('scan_order.npy', 'scan_order_optimized.npy'):
        np.save(OUT / name, np.load(WORK / name, allow_pickle=False))

    # rank profiles of the two RGS70 orders
    width = json.loads((WORK / 'scan_width.json').read_text())
    prof = json.loads((WORK / 'order_optimization' / 'profile.json').read_text())
    (OUT / 'rgs70_profiles.json').write_text(json.dumps({
        'scan_order': {'profile': width['profile'], 'sum_states': width['sum_states'],
                       'maxrank': width['maxrank']},
        'scan_order_optimized': {'profile': prof['optimized']['profile'],
                                 'sum_states': prof['optimized']['sum_states'],
                                 'maxrank': prof['optimized']['maxrank']},
        'baseline_profile_check': prof['baseline']['profile'],
    }, indent=1))

    # complex128 70q amplitudes (reverse-order FP64 scan) with the production
    # records that drew them: B bits, A index, RNG states and uniforms.
    prod = [r for r in jsonl(REF / 'born' / 'batch1' / 'born_group0.jsonl') if 'sample' in r]
    meta = jsonl(REF / 'born' / 'batch1' / 'born_group0.jsonl')[0]['metadata']
    fp64 = [r for r in jsonl(REF / 'born' / 'batch1' / 'born_group0_fp64.jsonl') if 'index' in r]
    rows = []
    for rec, sc in list(zip(prod, fp64))[:24]:
        assert rec['bitstring_q0_left'] == sc['bitstring_q0_left']
        rows.append({k: rec[k] for k in ('sample', 'bitstring_q0_left', 'B_bits', 'A_index',
                                         'rngstate_before', 'rngstate_after', 'draw_uniforms',
                                         'probability_raw', 'probability', 'conditional_norm')}
                    | {'re': sc['re'], 'im': sc['im'], 'p_fp64': sc['probability']})
    (OUT / 'rgs70_born_records.json').write_text(json.dumps({
        'group': meta['group'], 'A_vertices': meta['A_vertices'], 'B_vertices': meta['B_vertices'],
        'records': rows}, indent=1))

    # capped uniform-rejection run: metadata + first proposals (RNG replay) and the
    # FP64 audit of the same proposals (p_fp32, p_fp64, margins).
    cap = jsonl(WORK / 'cap1_archive' / 'capped1726.jsonl')
    cmeta = cap[0]['metadata']
    props = [r for r in cap if 'proposal' in r][:64]
    audit = {r['proposal']: r for r in jsonl(WORK / 'cap1_full_audit_artifact' / 'cap1_all_fp64_audit.jsonl')
             if 'proposal' in r}
    (OUT / 'rgs70_cap1_proposals.json').write_text(json.dumps({
        'metadata': {k: cmeta[k] for k in ('n', 'seed', 'cap')},
        'proposals': [{k: p[k] for k in ('proposal', 'bitstring_q0_left', 'probability_raw',
                                         'acceptance_probability', 'acceptance_uniform', 'accepted')}
                      | {'p_fp64': audit[p['proposal']]['p_fp64'],
                         'margin_fp64': audit[p['proposal']]['margin_fp64']} for p in props],
        'audit_summary': json.loads((WORK / 'cap1_full_audit_artifact' /
                                     'cap1_all_fp64_audit.jsonl.summary.json').read_text()),
    }, indent=1))

    # 32 inputs scored with both scan orders and a complex128 reference
    bench = json.loads((WORK / 'cap1_full_audit_artifact' / 'scan_order_benchmark.json').read_text())
    (OUT / 'rgs70_scan_order_benchmark.json').write_text(json.dumps({
        'summary': bench['summary'],
        'rows': [{k: r[k] for k in ('index', 'bitstring_q0_left', 'old_probability', 'new_probability',
                                    'reference_probability')} for r in bench['rows']]}, indent=1))

    # recorded statistics for the XEB formulas (born/verify.py:36-41)
    summ = json.loads((REF / 'born' / 'batch1' / 'analysis' / 'summary.json').read_text())
    z = [2.0 ** 70 * r['probability'] for r in fp64]
    (OUT / 'rgs70_batch1_group0_z.json').write_text(json.dumps({'z': z}))
    shutil.copyfile(REF / 'born' / 'batch1' / 'analysis' / 'summary.json', OUT / 'rgs70_batch1_summary.json')
    # capped run (cap 8, part 0): the FP64 scores of its accepted strings and
    # the recorded score summary of the capped-run analysis (analyze_challenge.py)
    import csv
    with open(WORK / 'cap8_archive' / 'analysis0' / 'accepted_scores.csv') as f:
        acc = [dict(r) for r in csv.DictReader(f)]
    cap8 = json.loads((WORK / 'cap8_archive' / 'analysis0' / 'summary.json').read_text())
    (OUT / 'rgs70_cap8_part0_scores.json').write_text(json.dumps({
        'p_fp64': [float(r['p_fp64']) for r in acc], 'xeb_fp64': [float(r['xeb_fp64']) for r in acc],
        'cap': cap8['cap'], 'proposals': cap8['proposals'], 'accepted': cap8['accepted'],
        'score': cap8['score']}))
    # the 70-qubit circuit file (input of the QASM front end, gzip-compressed)
    # and the instance block of the reference's structural audit of it
    import gzip
    qasm = REF / 'work' / 'qat' / 'data' / 'classically-verifiable-problems' / 'circuit-models' / \
        'random_graph_sampling' / 'nq70_depth70_checks27_basis_fstate.qasm'
    with gzip.GzipFile(OUT / 'rgs70_circuit.qasm.gz', 'wb', mtime=0) as f:
        f.write(qasm.read_bytes())
    inst = json.loads((WORK / 'audit.json').read_text())['instance']
    (OUT / 'rgs70_audit_instance.json').write_text(json.dumps(
        {k: inst[k] for k in ('n', 'graph_edges', 'partition_A', 'partition_B', 'cross_rank',
                              'has_exact_uniform_B_marginal', 'local_unitary_residual')}, indent=1))
    print('wrote fixtures to', OUT, 'summary keys', sorted(summ))


if __name__ == '__main__':
    main()
