"""Phase tables of the wave engine's elementwise ops: the phase terms of a run
of diagonal gates become a per-lane table per tile, per-slot lane and tile
factors and a per-row table, built on the device once per program.  Results
must match the per-term loop (QOMPSIM_PHASE_TABLES=0) to rounding and the
oracle within 1e-12 on complex128 amplitudes; the compressed stage loop
keeps the oracle's null bitmap."""
import re

import numpy as np
import pytest

import pyoracle as O
from gpu_util import rand_state, random_circuit
from qompsim import CompressedQuantumState, QuantumCircuit, QuantumCircuitOptimizer, QuantumState, gate
from qompsim import workloads as W

pytestmark = pytest.mark.gpu
TOL = 1e-12


def _table_ops(err):
    return sum(int(m) for m in re.findall(r"phase tables: ops=(\d+)", err))


def _mixed_diagonal_circuit(n, rng):
    """QFT (controlled phases between lane, slot and base qubits), a layer of
    RZ (terms requiring a bit to be 0), unit diagonal matrices on three
    qubits, doubly controlled phases (a slot term with lane and base
    conditions: not separable, the per-term loop) and a non-unit diagonal
    (a diagonal table beside the terms)."""
    c = QuantumCircuit(n)
    for g in W.qft_gates(n, swaps=False):
        c.add_gate(g)
    for q in range(n):
        c.add_RZ_gate(q, rng.uniform(0, 6.3))
    for _ in range(4):
        q = rng.choice(n, size=3, replace=False).tolist()
        c.add_gate(gate.DiagonalMatrix(q, np.exp(1j * rng.uniform(-3.1, 3.1, 8))))
    for _ in range(3):
        q = rng.choice(n, size=3, replace=False).tolist()
        g = W.cphase(q[0], q[1], rng.uniform(0, 6.3))
        g.add_control_qubit(q[2], 1)
        c.add_gate(g)
    c.add_gate(gate.DiagonalMatrix([1, n - 1], np.array([1.0, 0.5j, -0.25, 0.75 + 0.1j])))
    for g in W.qft_gates(n, swaps=True):
        c.add_gate(g)
    return c


@pytest.mark.parametrize("n", [12, 14, 17])
def test_tables_match_term_loop_and_oracle(n, rng, monkeypatch, capfd):
    monkeypatch.setenv("QOMPSIM_PLAN_LOG", "1")
    c = _mixed_diagonal_circuit(n, rng)
    psi = rand_state(n, rng)
    s1 = QuantumState(n)
    s1.load(psi)
    c.update_quantum_state(s1)
    assert _table_ops(capfd.readouterr().err) > 0
    monkeypatch.setenv("QOMPSIM_PHASE_TABLES", "0")
    s2 = QuantumState(n)
    s2.load(psi)
    c.update_quantum_state(s2)
    assert _table_ops(capfd.readouterr().err) == 0
    v1, v2 = s1.get_vector(), s2.get_vector()
    assert np.max(np.abs(v1 - v2)) < TOL
    assert np.max(np.abs(v1 - O.apply_circuit(psi, c.gate_list))) < TOL


def test_tables_tile_uniform_factor(rng, monkeypatch, capfd):
    """Diagonal runs conditioned on base qubits only (controlled phases and
    RZ among the high qubits, between H layers on the low ones): the
    per-tile factor is one entry per tile, read as a uniform value."""
    n = 15
    monkeypatch.setenv("QOMPSIM_PLAN_LOG", "1")
    c = QuantumCircuit(n)
    for layer in range(3):
        for q in range(4):
            c.add_H_gate(q)
        for a in range(11, n):
            c.add_RZ_gate(a, rng.uniform(0, 6.3))
            for b in range(a + 1, n):
                c.add_gate(W.cphase(a, b, rng.uniform(0, 6.3)))
    psi = rand_state(n, rng)
    s1, s2 = QuantumState(n), QuantumState(n)
    s1.load(psi)
    s2.load(psi)
    c.update_quantum_state(s1)
    assert _table_ops(capfd.readouterr().err) > 0
    monkeypatch.setenv("QOMPSIM_PHASE_TABLES", "0")
    c.update_quantum_state(s2)
    assert np.max(np.abs(s1.get_vector() - s2.get_vector())) < TOL
    assert np.max(np.abs(s1.get_vector() - O.apply_circuit(psi, c.gate_list))) < TOL


@pytest.mark.parametrize("tb", ["9", "10"])
def test_tables_compressed_mixed_circuit(tb, rng, monkeypatch, capfd):
    """The mixed diagonal circuit through the compressed stage loop on 512-
    and 1024-amplitude wave tiles (DEC->ENC stages and stages of several
    passes): tables against the per-term loop and the oracle's stage loop
    (amplitudes within 1e-6, the quantisation step at eps_rel 1e-3)."""
    n, B, I = 16, 10, 5
    monkeypatch.setenv("QOMPSIM_WAVE_TILE_BITS", tb)
    monkeypatch.setenv("QOMPSIM_PLAN_LOG", "1")
    c = _mixed_diagonal_circuit(n, rng)
    v = rand_state(n, rng)
    out = []
    for tables in ("1", "0"):
        monkeypatch.setenv("QOMPSIM_PHASE_TABLES", tables)
        cs = CompressedQuantumState(n, B, eps_rel=1e-3, eps_abs=1e-7, verbose=False)
        cs.load(v)
        cs.partition_and_register_circuit(c, I)
        cs.apply_circuit()
        ops = _table_ops(capfd.readouterr().err)
        assert (ops > 0) if tables == "1" else (ops == 0)
        out.append(cs)
    oc = O.OracleCompressedState(n, B, 1e-3, 1e-7, 1)
    oc.load(v)
    oc.partition_and_register_circuit(c, I)
    oc.apply_circuit()
    ref = oc.get_vector(normalise=False)
    for cs in out:
        assert np.array_equal(cs.get_null_bitmap(), oc.null_bitmap())
        assert np.max(np.abs(cs.get_raw_vector() - ref)) < 1e-6


@pytest.mark.parametrize("n", [10, 13, 15])
def test_tables_random_circuits(n, rng, monkeypatch):
    """Random two-target circuits with controlled phases, RZ, Pauli rotations
    and diagonal matrices among dense gates."""
    c = random_circuit(n, 150, rng, max_k=2)
    psi = rand_state(n, rng)
    s = QuantumState(n)
    s.load(psi)
    c.update_quantum_state(s)
    assert np.max(np.abs(s.get_vector() - O.apply_circuit(psi, c.gate_list))) < TOL


@pytest.mark.parametrize("n", [22, 27])
def test_tables_large_group_qft(n, rng, monkeypatch, capfd):
    """QFT-n with an HEA layer on 1024-amplitude tiles: 2^12 tiles per group
    (a per-tile lane table of 2^18 entries) and 2^17 tiles (past the lane
    table's limit: terms on one lane bit and base bits take the per-tile
    lane-bit angles, the rest the per-term loop); tables against the
    per-term loop."""
    monkeypatch.setenv("QOMPSIM_PLAN_LOG", "1")
    c = W.qft_on_product(n, seed=5)
    for q in range(n):
        c.add_RY_gate(q, rng.uniform(0, 6.3))
        c.add_RZ_gate(q, rng.uniform(0, 6.3))
    for q in range(n - 1):
        c.add_CZ_gate(q, q + 1)
    s1, s2 = QuantumState(n), QuantumState(n)
    c.update_quantum_state(s1)
    assert _table_ops(capfd.readouterr().err) > 0
    monkeypatch.setenv("QOMPSIM_PHASE_TABLES", "0")
    c.update_quantum_state(s2)
    assert np.max(np.abs(s1.get_vector() - s2.get_vector())) < TOL


def test_tables_adjoint_gradient(monkeypatch):
    """The adjoint sweep's wave kernels (single-wave 512-amplitude tiles with
    psi and lambda rows, and the paired 1024-amplitude form) with tables: the
    HEA gradient against the per-term loop and the oracle's backprop."""
    from qompsim.adjoint import adjoint_gradient

    n = 14
    c = W.hea(n, layers=3, seed=17)
    obs = W.tfim_observable(n)
    e1, g1 = adjoint_gradient(c, obs)
    monkeypatch.setenv("QOMPSIM_PHASE_TABLES", "0")
    e2, g2 = adjoint_gradient(c, obs)
    assert abs(e1 - e2) < TOL and np.max(np.abs(np.asarray(g1) - np.asarray(g2))) < TOL
    oe, og = O.backprop(c, obs)
    assert abs(e1 - oe) < 1e-10 and np.max(np.abs(np.asarray(g1) - og)) < 1e-10


def test_tables_compressed_qft_vs_oracle():
    """QFT on a product state through the compressed stage loop (DEC->ENC and
    DEC->RAW->ENC stages on the wave engine): null bitmap bit-exact,
    amplitudes within 1e-6 of the oracle's stage loop, and each stage's
    compressed bytes within 1e-3 of the oracle's.  An H followed by a pi/4
    controlled phase multiplies quantised amplitudes by (1/sqrt 2)^2: the
    real and imaginary parts land on quantisation ties up to rounding, and
    the side of each tie follows the order of the floating-point operations.
    About 1% of one stage's codes here differ from the oracle's by one step
    (with tables and with the per-term loop alike), which moves a few
    chunks' widths."""
    n, B, I = 18, 12, 4
    c = W.qft_on_product(n, seed=9)
    QuantumCircuitOptimizer().optimize(c, 2)
    cs = CompressedQuantumState(n, B, eps_rel=1e-3, eps_abs=1e-7, verbose=False)
    oc = O.OracleCompressedState(n, B, 1e-3, 1e-7, 1)
    cs.partition_and_register_circuit(c, I)
    oc.partition_and_register_circuit(c, I)
    cs.apply_circuit()
    oc.apply_circuit()
    assert np.array_equal(cs.get_null_bitmap(), oc.null_bitmap())
    h, oh = np.array(cs.get_bytes_history(), float), np.array(oc.bytes_history(), float)
    assert h.shape == oh.shape and np.all(np.abs(h - oh) <= 1e-3 * oh)
    assert np.max(np.abs(cs.get_raw_vector() - oc.get_vector(normalise=False))) < 1e-6
