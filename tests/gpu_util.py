import numpy as np

from qompsim import QuantumCircuit, gate
from qompsim import workloads as W


def rand_state(n, rng):
    v = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    return v / np.linalg.norm(v)


def rand_unitary(d, rng):
    q, r = np.linalg.qr(rng.normal(size=(d, d)) + 1j * rng.normal(size=(d, d)))
    return q * (np.diag(r) / np.abs(np.diag(r)))


def random_circuit(n, count, rng, max_k=3):
    """Mixed gate set: dense 1..max_k-qubit (with controls), named gates,
    diagonal and controlled-phase gates, Pauli rotations."""
    c = QuantumCircuit(n)
    for _ in range(count):
        r = rng.random()
        q = rng.choice(n, size=min(n, max_k + 2), replace=False).tolist()
        if r < 0.3:
            k = int(rng.integers(1, min(max_k, n) + 1))
            g = gate.DenseMatrix(q[:k], rand_unitary(1 << k, rng))
            if len(q) > k and rng.random() < 0.4:
                g.add_control_qubit(q[k], int(rng.integers(0, 2)))
        elif r < 0.4 and n >= 2:
            g = gate.CNOT(q[0], q[1])
        elif r < 0.5:
            g = gate.H(q[0])
        elif r < 0.6:
            g = gate.RZ(q[0], rng.uniform(0, 6.3))
        elif r < 0.7 and n >= 2:
            g = W.cphase(q[0], q[1], rng.uniform(0, 6.3))
        elif r < 0.8 and n >= 2:
            g = gate.PauliRotation(q[:2], [3, 1], rng.uniform(0, 6.3))
        elif r < 0.9:
            g = gate.DiagonalMatrix(q[:1], np.exp(1j * rng.uniform(0, 6.3, 2)))
        else:
            g = gate.RY(q[0], rng.uniform(0, 6.3))
        c.add_gate(g)
    return c
