"""GPU parity of the uncompressed LDS-pass path against the oracle
(tolerance 1e-12 on complex128 amplitudes)."""
import os

import numpy as np
import pytest

import pyoracle as O
from gpu_util import rand_state, random_circuit
from qompsim import Observable, QuantumCircuit, QuantumState, inner_product
from qompsim import workloads as W

pytestmark = pytest.mark.gpu
TOL = 1e-12


@pytest.fixture
def tile_bits():
    old = os.environ.get("QOMPSIM_TILE_BITS")
    yield lambda v: os.environ.__setitem__("QOMPSIM_TILE_BITS", str(v))
    if old is None:
        os.environ.pop("QOMPSIM_TILE_BITS", None)
    else:
        os.environ["QOMPSIM_TILE_BITS"] = old


@pytest.mark.parametrize("n,max_k", [(1, 1), (2, 2), (3, 3), (5, 3), (8, 3), (11, 3), (14, 3), (10, 2), (13, 2),
                                     (17, 2)])
def test_random_circuits_match_oracle(n, max_k, rng):
    # max_k <= 2 on >= 10 qubits runs the wave-resident engine, the rest the LDS-tile engine
    c = random_circuit(n, 60, rng, max_k=min(max_k, n))
    psi = rand_state(n, rng)
    s = QuantumState(n)
    s.load(psi)
    c.update_quantum_state(s)
    ref = O.apply_circuit(psi, c.gate_list)
    assert np.max(np.abs(s.get_vector() - ref)) < TOL


@pytest.mark.parametrize("tb", [5, 7, 10, 12, 13])
def test_tile_sizes(tb, tile_bits, rng):
    tile_bits(tb)
    n = 15
    c = random_circuit(n, 80, rng, max_k=3)
    psi = rand_state(n, rng)
    s = QuantumState(n)
    s.load(psi)
    c.update_quantum_state(s)
    assert np.max(np.abs(s.get_vector() - O.apply_circuit(psi, c.gate_list))) < TOL


@pytest.mark.parametrize("n", [10, 12, 16])
def test_engines_agree(n, rng, monkeypatch):
    c = random_circuit(n, 120, rng, max_k=2)
    psi = rand_state(n, rng)
    s1, s2 = QuantumState(n), QuantumState(n)
    s1.load(psi)
    s2.load(psi)
    c.update_quantum_state(s1)
    monkeypatch.setenv("QOMPSIM_ENGINE", "lds")
    c.update_quantum_state(s2)
    assert np.max(np.abs(s1.get_vector() - s2.get_vector())) < TOL


def test_five_qubit_dense_gate(rng):
    from gpu_util import rand_unitary
    from qompsim import QuantumCircuit, gate

    n = 12
    c = QuantumCircuit(n)
    c.add_gate(gate.DenseMatrix([9, 0, 4, 11, 2], rand_unitary(32, rng)))
    g = gate.DenseMatrix([3, 7, 1, 10], rand_unitary(16, rng))
    g.add_control_qubit(5, 1)
    c.add_gate(g)
    psi = rand_state(n, rng)
    s = QuantumState(n)
    s.load(psi)
    c.update_quantum_state(s)
    assert np.max(np.abs(s.get_vector() - O.apply_circuit(psi, c.gate_list))) < TOL


def test_qft20_is_dft(rng):
    n = 20
    psi = rand_state(n, rng)
    s = QuantumState(n)
    s.load(psi)
    c = W.qft_on_product(n)
    c.gate_list = W.qft_gates(n)
    c.update_quantum_state(s)
    assert np.max(np.abs(s.get_vector() - np.fft.ifft(psi, norm="ortho"))) < TOL


def test_ghz_and_basis(rng):
    s = QuantumState(16)
    W.ghz(16).update_quantum_state(s)
    v = s.get_vector()
    assert abs(v[0] - 2 ** -0.5) < 1e-15 and abs(v[-1] - 2 ** -0.5) < 1e-15
    s.set_computational_basis(12345)
    v = s.get_vector()
    assert v[12345] == 1 and np.count_nonzero(v) == 1


def test_reductions_match_oracle(rng):
    n = 16
    a, b = rand_state(n, rng) * 1.3, rand_state(n, rng)
    sa, sb = QuantumState(n), QuantumState(n)
    sa.load(a)
    sb.load(b)
    assert abs(sa.get_squared_norm() - O.norm(a)) < 1e-13
    assert abs(inner_product(sa, sb) - O.inner(a, b)) < 1e-13
    obs = Observable(n)
    obs.add_operator(0.3, "X 0 Y 9 Z 15")
    obs.add_operator(-1.1, "Z 1 Z 3")
    obs.add_operator(0.7, "X 12")
    obs.add_operator(0.2, "")
    assert abs(obs.get_expectation_value(sa) - O.expectation(a, obs).real) < 1e-13
    out = QuantumState(n)
    obs.apply_to_state(None, sa, out)
    assert np.max(np.abs(out.get_vector() - O.apply_pauli_sum(a, obs))) < TOL


def test_persistent_raw_passes_many_tiles_per_wave(rng, monkeypatch):
    """26 qubits: the persistent RAW passes walk several tiles per wave, each
    tile's rows staged in LDS during the previous tile; the wave engine must
    agree with the LDS-tile engine (which stages nothing)."""
    n = 26
    c = random_circuit(n, 40, rng, max_k=2)
    psi = rand_state(n, rng)
    s1, s2 = QuantumState(n), QuantumState(n)
    s1.load(psi)
    s2.load(psi)
    c.update_quantum_state(s1)
    monkeypatch.setenv("QOMPSIM_ENGINE", "lds")
    c.update_quantum_state(s2)
    assert np.max(np.abs(s1.get_vector() - s2.get_vector())) < TOL


def test_real_dense_path_bit_identical(rng, monkeypatch):
    """Real one-target matrices (RY, H, X, real dense with zero and negative
    entries) and matrices with real diagonal and imaginary off-diagonal
    entries (RX, i X) take the wave engine's cheaper paths on every lane bit
    and register slot; they must give the same bits as the complex paths
    (QOMPSIM_REAL_DENSE=0) and match the oracle."""
    from qompsim import QuantumCircuit
    from qompsim import gate

    n = 14
    c = QuantumCircuit(n)
    for layer in range(3):
        for q in range(n):
            c.add_gate(gate.RY(q, rng.uniform(0, 6.3)))
            c.add_gate(gate.RZ(q, rng.uniform(0, 6.3)))
        for q in range(0, n, 3):
            c.add_gate(gate.H(q))
            c.add_gate(gate.X((q + 1) % n))
            c.add_gate(gate.DenseMatrix([(q + 2) % n], np.array([[0.0, -1.0], [1.0, 0.0]])))
            c.add_gate(gate.RX((q + 1) % n, rng.uniform(0, 6.3)))
            c.add_gate(gate.DenseMatrix([q], np.array([[0.0, 1j], [1j, 0.0]])))
            c.add_gate(gate.DenseMatrix([(q + 2) % n], np.array([[0.6, -0.8j], [0.8j, -0.6]])))
        for q in range(n - 1):
            c.add_gate(gate.CZ(q, q + 1))
    psi = rand_state(n, rng)
    s1, s2 = QuantumState(n), QuantumState(n)
    s1.load(psi)
    s2.load(psi)
    c.update_quantum_state(s1)
    monkeypatch.setenv("QOMPSIM_REAL_DENSE", "0")
    c.update_quantum_state(s2)
    v1, v2 = s1.get_vector(), s2.get_vector()
    assert np.array_equal(v1.view(np.uint64), v2.view(np.uint64))
    assert np.max(np.abs(v1 - O.apply_circuit(psi, c.gate_list))) < TOL


def test_diagonal_gates_gathered_within_passes(rng, monkeypatch):
    """RY RZ pairs per qubit and a CZ chain (an HEA layer): the diagonal gates
    of a pass move past the commuting dense gates into one elementwise op.
    The state must match the original order (QOMPSIM_DIAG_GATHER=0) to
    rounding and the oracle; the adjoint gradient too."""
    from qompsim.adjoint import adjoint_gradient

    n = 14
    c = W.hea(n, layers=3, seed=11)
    psi = rand_state(n, rng)
    s1, s2 = QuantumState(n), QuantumState(n)
    s1.load(psi)
    s2.load(psi)
    c.update_quantum_state(s1)
    obs = W.tfim_observable(n)
    e1, g1 = adjoint_gradient(c, obs)
    monkeypatch.setenv("QOMPSIM_DIAG_GATHER", "0")
    c.update_quantum_state(s2)
    e2, g2 = adjoint_gradient(c, obs)
    v1, v2 = s1.get_vector(), s2.get_vector()
    assert np.max(np.abs(v1 - v2)) < TOL
    assert np.max(np.abs(v1 - O.apply_circuit(psi, c.gate_list))) < TOL
    assert abs(e1 - e2) < TOL and np.max(np.abs(np.asarray(g1) - np.asarray(g2))) < TOL


@pytest.mark.parametrize("n", [12, 16])
def test_product_matrices_split(n, monkeypatch):
    """Two-qubit matrices that factor as (A x B) D or D (A x B) (the
    optimizer's merge of RZZ with RX gates, and of phases after one-qubit
    gates) are applied as the diagonal and two one-qubit gates: the oracle's
    amplitudes within 1e-12, and within 1e-13 of the 4x4 products
    (QOMPSIM_SPLIT_PRODUCTS=0)."""
    from qompsim import QuantumCircuitOptimizer

    rng = np.random.default_rng(n)
    c = QuantumCircuit(n)
    for i in range(n - 1):
        c.add_multi_Pauli_rotation_gate([i, i + 1], [3, 3], float(rng.uniform(0, 6.3)))
    for q in range(n):
        c.add_RX_gate(q, float(rng.uniform(0, 6.3)))
    for q in range(0, n - 1, 2):
        c.add_RY_gate(q, float(rng.uniform(0, 6.3)))
        c.add_gate(W.cphase(q, q + 1, float(rng.uniform(0, 6.3))))
    QuantumCircuitOptimizer().optimize(c, 2)
    qft = W.qft_on_product(n, seed=n)  # H merged with controlled phases: selected one-qubit matrices
    QuantumCircuitOptimizer().optimize(qft, 2)
    for circ in (c, qft):
        monkeypatch.delenv("QOMPSIM_SPLIT_PRODUCTS", raising=False)
        psi = rand_state(n, rng)
        s1 = QuantumState(n)
        s1.load(psi)
        circ.update_quantum_state(s1)
        monkeypatch.setenv("QOMPSIM_SPLIT_PRODUCTS", "0")
        s2 = QuantumState(n)
        s2.load(psi)
        circ.update_quantum_state(s2)
        ref = O.apply_circuit(psi, circ.gate_list)
        assert np.max(np.abs(s1.get_vector() - ref)) < 1e-12
        assert np.max(np.abs(s1.get_vector() - s2.get_vector())) < 1e-13
