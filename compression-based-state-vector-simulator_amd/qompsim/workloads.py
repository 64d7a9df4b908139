"""Circuits of the benchmark configurations (seeded, synthetic).

GHZ follows benchmark/circuit_factory.py's ghz (H(0) then a CNOT ladder); the
QFT, HEA, QAOA and TFIM circuits are the ones the measurement configurations
define: QFT-with-swaps on an RY*RZ product state (unitary DFT of the input in
the index = sum_q 2^q b_q convention), an RY/RZ + CZ-chain ansatz with a
transverse-field Ising cost, QAOA-MaxCut on random 3-regular graphs, and a
transverse-field Ising Trotter evolution.
"""
import numpy as np

from . import gate as _g
from ._lib import COMMUTE_Z
from .circuit import ParametricQuantumCircuit, QuantumCircuit
from .observable import Observable


def ghz(n):
    c = QuantumCircuit(n)
    c.add_H_gate(0)
    for i in range(n - 1):
        c.add_CNOT_gate(i, i + 1)
    return c


def cphase(control, target, phi):
    """Controlled phase diag(1, e^{i phi}) on target, conditioned on control."""
    return _g.QuantumGate("CP", [target], np.diag([1.0, np.exp(1j * phi)]), controls=[control],
                          commute=[COMMUTE_Z])


def qft_gates(n, swaps=True):
    gates = []
    for j in range(n - 1, -1, -1):
        gates.append(_g.H(j))
        for k in range(j - 1, -1, -1):
            gates.append(cphase(k, j, np.pi / 2 ** (j - k)))
    if swaps:
        for q in range(n // 2):
            gates.append(_g.SWAP(q, n - 1 - q))
    return gates


def product_layer(n, seed):
    rng = np.random.default_rng(seed)
    th = rng.uniform(0, 2 * np.pi, n)
    ph = rng.uniform(0, 2 * np.pi, n)
    gates = []
    for q in range(n):
        gates.append(_g.RY(q, th[q]))
        gates.append(_g.RZ(q, ph[q]))
    return gates, th, ph


def qft_on_product(n, seed=2):
    """RY(theta_q) RZ(phi_q) per qubit then QFT with swaps: 2n + n + n(n-1)/2 + n//2 gates."""
    c = QuantumCircuit(n)
    layer, _, _ = product_layer(n, seed)
    for g in layer + qft_gates(n):
        c.add_gate(g)
    return c


def hea(n, layers=8, seed=3):
    c = ParametricQuantumCircuit(n)
    rng = np.random.default_rng(seed)
    th = rng.uniform(0, 2 * np.pi, 2 * n * layers)
    k = 0
    for _ in range(layers):
        for q in range(n):
            c.add_parametric_RY_gate(q, th[k]); k += 1
            c.add_parametric_RZ_gate(q, th[k]); k += 1
        for q in range(n - 1):
            c.add_CZ_gate(q, q + 1)
    return c


def tfim_observable(n, hx=0.7):
    obs = Observable(n)
    for i in range(n - 1):
        obs.add_operator(1.0, f"Z {i} Z {i + 1}")
    for i in range(n):
        obs.add_operator(hx, f"X {i}")
    return obs


def random_3regular(n, rng):
    """Random 3-regular simple graph by the pairing model with rejection."""
    while True:
        pts = np.repeat(np.arange(n), 3)
        rng.shuffle(pts)
        edges = set()
        ok = True
        for a, b in zip(pts[::2], pts[1::2]):
            if a == b or (min(a, b), max(a, b)) in edges:
                ok = False
                break
            edges.add((min(a, b), max(a, b)))
        if ok:
            return sorted(edges)


def qaoa_maxcut(n, edges, gammas, betas):
    c = ParametricQuantumCircuit(n)
    for q in range(n):
        c.add_H_gate(q)
    for gm, bt in zip(gammas, betas):
        for (a, b) in edges:
            c.add_parametric_multi_Pauli_rotation_gate([a, b], [3, 3], 2 * gm)
        for q in range(n):
            c.add_parametric_RX_gate(q, 2 * bt)
    return c


def qaoa_maxcut_shared(n, edges, gammas, betas):
    """QAOA MaxCut with one parameter per layer angle: gate parameter index
    2l holds RZZ(2 gamma_l) of every edge, 2l+1 RX(2 beta_l) of every qubit,
    so gradients accumulate per shared parameter (d/dgamma_l = 2 * d/dangle)."""
    c = ParametricQuantumCircuit(n)
    for q in range(n):
        c.add_H_gate(q)
    for layer, (gm, bt) in enumerate(zip(gammas, betas)):
        for (a, b) in edges:
            c.add_parametric_multi_Pauli_rotation_gate([a, b], [3, 3], 2 * gm)
            c.gate_list[-1]._param_index = 2 * layer
        for q in range(n):
            c.add_parametric_RX_gate(q, 2 * bt)
            c.gate_list[-1]._param_index = 2 * layer + 1
    return c


def qaoa_batch(count=512, sizes=(12, 14, 16, 18, 20, 22, 24), p=4, seed=4):
    """Seeded instances: random 3-regular graphs, node counts round-robin over
    `sizes`, gamma, beta ~ U[0, pi)."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        n = sizes[i % len(sizes)]
        edges = random_3regular(n, rng)
        gm = rng.uniform(0, np.pi, p)
        bt = rng.uniform(0, np.pi, p)
        out.append((n, edges, gm, bt))
    return out


def maxcut_observable(n, edges):
    obs = Observable(n)
    for (a, b) in edges:
        obs.add_operator(0.5, "")
        obs.add_operator(-0.5, f"Z {a} Z {b}")
    return obs


def tfim_trotter(n, steps=12, jzz=0.2, hx=0.3, excited=1):
    c = QuantumCircuit(n)
    c.add_X_gate(excited)
    for _ in range(steps):
        for i in range(n - 1):
            c.add_multi_Pauli_rotation_gate([i, i + 1], [3, 3], jzz)
        for q in range(n):
            c.add_RX_gate(q, hx)
    return c
