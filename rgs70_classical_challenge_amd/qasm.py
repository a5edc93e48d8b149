"""QASM front end: a Clifford-plus-magic-tail circuit file -> (graph, per-qubit
measurement unitaries), the inputs of every sampler engine.

Restates ``Tableau`` and ``parse_qasm`` of the reference
(work/rgs_attack/graph_sampler.py:11-57, 111-132):

* The Clifford part of the circuit (every gate before the first ``tdg``) is
  simulated on a stabilizer tableau started at |0...0>.
* The remaining single-qubit gates (``h``, ``tdg``, ``rz``) form the magic
  tail; each qubit's tail is multiplied into one 2x2 matrix.
* The stabilizer state is brought to graph form by local Cliffords T_q,
  |stab> = (x)_q T_q^dagger |G> up to a global phase, and the measurement
  unitary of qubit q is tail_q T_q^dagger, so the circuit's output state is
  (x)_q u_q |G>.

Representation (this module's own): the n generators are stored column-wise
as Python integers - ``x[a]`` and ``z[a]`` hold bit i = the X and Z part of
generator i on qubit a, and ``r`` holds the sign bits - so a gate on qubit a
updates every generator with a few big-integer operations.  Generators use
the Hermitian convention (-1)^r prod_a P(x_a, z_a) with P(1, 1) = Y.  The
conversion to graph form works on rows (bit a of ``X[i]`` = x of generator i
on qubit a).

The graph-form conversion follows the same pivot rule as the reference
(column j takes the first generator at or below row j with an X on qubit j,
else qubit j gets a Hadamard), because the graph a circuit maps to depends on
those choices and the reference's recorded graph and unitaries
(rgs_graph.npz) are the parity target.
"""
import re
from pathlib import Path

import numpy as np

HADAMARD = np.array([[1, 1], [1, -1]], np.complex128) / np.sqrt(2.0)
S_GATE = np.diag([1.0, 1j])
Z_GATE = np.diag([1.0, -1.0]).astype(np.complex128)
TDG_GATE = np.diag([1.0, np.exp(-1j * np.pi / 4)])


def _pc(v):
    return bin(v).count('1')


class Tableau:
    """Stabilizer tableau of an n-qubit state (generator i starts as Z_i)."""

    def __init__(self, n):
        self.n = n
        self.mask = (1 << n) - 1
        self.x = [0] * n
        self.z = [1 << a for a in range(n)]
        self.r = 0

    # -- Clifford gates (conjugation of every generator) -------------------
    def h(self, a):
        self.r ^= self.x[a] & self.z[a]
        self.x[a], self.z[a] = self.z[a], self.x[a]

    def s(self, a):
        self.r ^= self.x[a] & self.z[a]
        self.z[a] ^= self.x[a]

    def sdg(self, a):
        self.r ^= self.x[a] & ~self.z[a] & self.mask
        self.z[a] ^= self.x[a]

    def zgate(self, a):
        self.r ^= self.x[a]

    def xgate(self, a):
        self.r ^= self.z[a]

    def sx(self, a):
        # sqrt(X) = H S H
        self.h(a)
        self.s(a)
        self.h(a)

    def cx(self, a, b):
        xa, za, xb, zb = self.x[a], self.z[a], self.x[b], self.z[b]
        self.r ^= xa & zb & ~(xb ^ za) & self.mask
        self.x[b] = xb ^ xa
        self.z[a] = za ^ zb

    def cz(self, a, b):
        self.h(b)
        self.cx(a, b)
        self.h(b)

    GATES = {'h': 1, 's': 1, 'sdg': 1, 'z': 1, 'x': 1, 'sx': 1, 'cx': 2, 'cz': 2}

    def gate(self, name, qs):
        """Apply a named gate (the reference's gate names) to qubits qs."""
        if name not in self.GATES or len(qs) != self.GATES[name]:
            raise ValueError(f'unsupported Clifford gate {name} on {qs}')
        fn = {'z': self.zgate, 'x': self.xgate}.get(name) or getattr(self, name)
        fn(*qs)

    # -- graph form ----------------------------------------------------------
    def rows(self):
        """(X rows, Z rows, sign bits) with bit a of X[i] = x of generator i on qubit a."""
        n = self.n
        X = [0] * n
        Z = [0] * n
        for a in range(n):
            xa, za = self.x[a], self.z[a]
            for i in range(n):
                X[i] |= ((xa >> i) & 1) << a
                Z[i] |= ((za >> i) & 1) << a
        return X, Z, [(self.r >> i) & 1 for i in range(n)]

    def to_graph(self):
        """(G uint8 [n][n], T complex128 [n][2][2]) with |G> = (x)_q T_q |stab>
        up to a global phase.  Consumes the tableau's state."""
        n = self.n
        X, Z, R = self.rows()
        local = np.repeat(np.eye(2, dtype=np.complex128)[None], n, axis=0)

        def on_column(j, kind):
            bit = 1 << j
            for i in range(n):
                xi, zi = (X[i] >> j) & 1, (Z[i] >> j) & 1
                if kind == 'h':
                    R[i] ^= xi & zi
                    X[i] = (X[i] & ~bit) | (zi << j)
                    Z[i] = (Z[i] & ~bit) | (xi << j)
                elif kind == 'sdg':
                    R[i] ^= xi & (1 - zi)
                    Z[i] ^= xi << j
                else:  # 'z'
                    R[i] ^= xi

        def multiply_into(i, j):
            # generator i <- generator i * generator j, sign from the i-power count
            xn, zn = X[i] ^ X[j], Z[i] ^ Z[j]
            e = (2 * R[i] + 2 * R[j] + 2 * _pc(Z[i] & X[j]) + _pc(X[i] & Z[i]) + _pc(X[j] & Z[j])
                 - _pc(xn & zn)) % 4
            if e % 2:
                raise ArithmeticError('generators do not commute')
            R[i], X[i], Z[i] = e // 2, xn, zn

        for j in range(n):
            piv = next((i for i in range(j, n) if (X[i] >> j) & 1), None)
            if piv is None:
                on_column(j, 'h')
                local[j] = HADAMARD @ local[j]
                piv = next((i for i in range(j, n) if (X[i] >> j) & 1), None)
            if piv is None:
                raise RuntimeError('no graph pivot')
            X[j], X[piv] = X[piv], X[j]
            Z[j], Z[piv] = Z[piv], Z[j]
            R[j], R[piv] = R[piv], R[j]
            for i in range(n):
                if i != j and (X[i] >> j) & 1:
                    multiply_into(i, j)
        if any(X[i] != 1 << i for i in range(n)):
            raise ArithmeticError('X part is not the identity after elimination')
        G = np.array([[(Z[i] >> a) & 1 for a in range(n)] for i in range(n)], np.uint8)
        if not np.array_equal(G, G.T):
            raise ArithmeticError('Z part is not symmetric')
        for j in range(n):
            if G[j, j]:
                on_column(j, 'sdg')
                local[j] = S_GATE.conj().T @ local[j]
        for j in range(n):
            if R[j]:
                on_column(j, 'z')
                local[j] = Z_GATE @ local[j]
        if any(R):
            raise ArithmeticError('signs left after the local Z corrections')
        G = np.array([[(Z[i] >> a) & 1 for a in range(n)] for i in range(n)], np.uint8)
        return G, local


_LINE = re.compile(r'(\w+)(?:\(([^)]*)\))? (q\[.*);')


def parse_qasm(path):
    """(graph uint8 [n][n], measurement unitaries complex128 [n][2][2], the
    tableau, the gate list) of an OpenQASM 2 file: Clifford gates up to the
    first ``tdg``, then a single-qubit magic tail of h / tdg / rz.  Same
    return values as the reference's parse_qasm."""
    text = Path(path).read_text() if not str(path).endswith('.gz') else _read_gz(path)
    m = re.search(r'qreg q\[(\d+)\]', text)
    if not m:
        raise ValueError('no qreg declaration')
    n = int(m.group(1))
    tab = Tableau(n)
    tail = np.repeat(np.eye(2, dtype=np.complex128)[None], n, axis=0)
    in_tail = False
    gates = []
    for line in text.splitlines():
        mm = _LINE.match(line)
        if not mm:
            continue
        name, arg, ops = mm.groups()
        if name == 'qreg':
            continue
        qs = [int(v) for v in re.findall(r'q\[(\d+)\]', ops)]
        if any(not 0 <= q < n for q in qs):
            raise ValueError(f'qubit out of range in: {line}')
        in_tail = in_tail or name == 'tdg'
        gates.append((name, arg, qs))
        if not in_tail:
            tab.gate(name, qs)
            continue
        if len(qs) != 1:
            raise ValueError('entangling gate after the magic tail began')
        if name == 'h':
            mat = HADAMARD
        elif name == 'tdg':
            mat = TDG_GATE
        elif name == 'rz':
            theta = float(arg)
            mat = np.diag([np.exp(-0.5j * theta), np.exp(0.5j * theta)])
        else:
            raise ValueError(f'gate {name} not allowed in the magic tail')
        tail[qs[0]] = mat @ tail[qs[0]]
    g, t = tab.to_graph()
    u = tail @ np.conj(np.transpose(t, (0, 2, 1)))
    return g, u, tab, gates


def _read_gz(path):
    import gzip
    with gzip.open(path, 'rt') as f:
        return f.read()


def cut_rank(g, A, B):
    """GF(2) rank of G[A, B] (the native elimination of rgs_gf2_rank)."""
    from .engine import gf2_rank
    return gf2_rank(np.asarray(g, np.uint8)[np.ix_(A, B)])


def main(argv):
    """python -m rgs70_classical_challenge_amd.qasm circuit.qasm[.gz] [out.npz]:
    convert, print the instance summary, and write graph / unitaries / a
    balanced full-rank partition (first of up to 1000 seeded random splits)."""
    import json
    g, u, _, gates = parse_qasm(argv[0])
    n = len(g)
    rng = np.random.default_rng(20261030)
    for _ in range(1000):
        perm = rng.permutation(n)
        A, B = perm[:n // 2].tolist(), perm[n // 2:].tolist()
        r = cut_rank(g, A, B)
        if r == n // 2:
            break
    out = {'n': n, 'gates': len(gates), 'graph_edges': int(g.sum() // 2), 'partition_A': A, 'partition_B': B,
           'cross_rank': r, 'local_unitary_residual': float(np.max(np.abs(u @ u.conj().transpose(0, 2, 1) - np.eye(2))))}
    if len(argv) > 1:
        np.savez(argv[1], graph=g, measurement_unitaries=u, A=A, B=B)
    print(json.dumps(out), flush=True)


if __name__ == '__main__':
    import sys
    main(sys.argv[1:])
