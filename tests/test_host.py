"""Host-side checks without a GPU: the C ABI library loads and exports every
entry point include/qsim.h declares; the C++ optimizer keeps the circuit
unitary and reproduces the derived GHZ fusion; Python surface conventions."""
import os
import re

import numpy as np
import pytest

import pyoracle as O
from qompsim import QuantumCircuit, QuantumCircuitOptimizer, _lib, gate
from qompsim import workloads as W

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol():
    hdr = open(os.path.join(REPO, "include", "qsim.h")).read()
    declared = set(re.findall(r"\b(qsim_[a-z0-9_]+)\s*\(", hdr))
    assert len(declared) > 30
    for name in declared:
        assert hasattr(_lib.lib, name), name
    assert set(_lib.EXPORTED) == declared


def test_gate_record_layout():
    assert _lib.GATE_DTYPE.itemsize == 184
    assert _lib.GATE_DTYPE.fields["matrix_offset"][1] == 176


def test_error_path_reports_message():
    with pytest.raises(ValueError, match="block_size"):
        from qompsim import CompressedQuantumState

        CompressedQuantumState(10, 3, verbose=False)


def _rand_circuit(n, rng, count=40):
    c = QuantumCircuit(n)
    makers = [lambda: gate.H(int(rng.integers(n))), lambda: gate.RX(int(rng.integers(n)), rng.uniform(0, 6)),
              lambda: gate.RZ(int(rng.integers(n)), rng.uniform(0, 6)), lambda: gate.T(int(rng.integers(n))),
              lambda: gate.CNOT(*map(int, rng.choice(n, 2, replace=False))),
              lambda: gate.CZ(*map(int, rng.choice(n, 2, replace=False))),
              lambda: gate.SWAP(*map(int, rng.choice(n, 2, replace=False))),
              lambda: W.cphase(*map(int, rng.choice(n, 2, replace=False)), rng.uniform(0, 6))]
    for _ in range(count):
        c.add_gate(makers[int(rng.integers(len(makers)))]())
    return c


@pytest.mark.parametrize("block", [2, 3])
def test_optimizer_preserves_unitary(rng, block):
    n = 6
    for _ in range(3):
        c = _rand_circuit(n, rng)
        psi = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
        before = O.apply_circuit(psi, c.gate_list)
        QuantumCircuitOptimizer().optimize(c, block)
        after = O.apply_circuit(psi, c.gate_list)
        assert np.max(np.abs(before - after)) < 1e-12
        for g in c.gate_list:
            assert len(g.get_target_index_list()) + len(g.get_control_index_list()) <= max(block, 2)


def test_optimizer_ghz_derivation():
    c = W.ghz(20)
    QuantumCircuitOptimizer().optimize(c, 2)
    names = [g.get_name() for g in c.gate_list]
    assert names == ["DenseMatrix"] + ["CNOT"] * 18
    assert c.gate_list[0].get_target_index_list() == [0, 1]
    assert [g.get_control_index_list() for g in c.gate_list[1:]] == [[i] for i in range(1, 19)]


def test_optimizer_never_merges_parametric():
    c = W.hea(4, layers=1)
    n0 = c.get_gate_count()
    QuantumCircuitOptimizer().optimize(c, 2)
    assert sum(g.is_parametric() for g in c.gate_list) == 8
    assert c.get_gate_count() <= n0


def test_python_merge_matches_optimizer_unitary(rng):
    g = gate.merge(gate.H(0), gate.CNOT(0, 1))
    psi = rng.normal(size=4) + 1j * rng.normal(size=4)
    assert np.allclose(O.apply_circuit(psi, [g]), O.apply_circuit(psi, [gate.H(0), gate.CNOT(0, 1)]))
    assert g.get_target_index_list() == [0, 1]


def test_qft_gate_count():
    c = W.qft_on_product(32, seed=2)
    assert c.get_gate_count() == 608


def test_commuting_partition_plan_is_valid_and_shorter():
    """Host-side stage packing: every stage's high set fits max(I, 2), each
    stage lists gates in circuit order, it needs no more stages than
    partition_circuit, and applying the stages in turn equals the circuit
    (oracle, random state)."""
    import numpy as np

    import pyoracle as O
    from qompsim import QuantumCircuitOptimizer
    from qompsim import workloads as W
    from qompsim.state import partition_commuting

    c = W.qft_on_product(32, seed=2)
    QuantumCircuitOptimizer().optimize(c, 2)
    st = partition_commuting(c, 20, 4)
    assert sorted(i for s in st for i in s) == list(range(len(c.gate_list)))
    for s in st:
        assert s == sorted(s)
        hi = set()
        for i in s:
            g = c.gate_list[i]
            hi |= {q for q in list(g.get_target_index_list()) + list(g.get_control_index_list()) if q >= 20}
        assert len(hi) <= 4
    assert len(st) <= 24  # partition_circuit gives 29 on this circuit
    n = 11
    c2 = W.qft_on_product(n, seed=3)
    QuantumCircuitOptimizer().optimize(c2, 2)
    st2 = partition_commuting(c2, 6, 2)
    rng = np.random.default_rng(2)
    v = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    v /= np.linalg.norm(v)
    ref = O.apply_circuit(v, c2.gate_list)
    got = O.apply_circuit(v, [c2.gate_list[i] for s in st2 for i in s])
    assert np.max(np.abs(got - ref)) < 1e-12


def test_qulacs_import_surface():
    """The reference harness's imports (script/run_sim.py:4-6) resolve to this
    package's classes."""
    import qulacs
    from qulacs import CompressedQuantumState, QuantumState
    from qulacs.circuit import QuantumCircuitOptimizer
    from qulacs.gate import CNOT, H
    from qulacs.state import inner_product

    import qompsim

    assert QuantumState is qompsim.QuantumState
    assert CompressedQuantumState is qompsim.CompressedQuantumState
    assert QuantumCircuitOptimizer is qompsim.QuantumCircuitOptimizer
    assert inner_product is qompsim.inner_product
    assert qulacs.check_build_for_mpi() is False
    c = qulacs.QuantumCircuit(3)
    c.add_gate(H(0))
    c.add_gate(CNOT(0, 1))
    assert [g.get_name() for g in c.gate_list] == ["H", "CNOT"]
