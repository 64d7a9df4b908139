"""Region store (store="regions", QSIM_STORE_REGIONS): the stage loop with its
payload in regions mapped on demand and released as they are consumed.  The
codec, the gates and the null bookkeeping are those of the double-buffered
store, so every result must be bit-identical to it; only where the payload
sits, and how much device memory it holds, differ."""
import numpy as np
import pytest

from gpu_util import rand_state, random_circuit
from qompsim import CompressedQuantumState, QuantumCircuit, QuantumCircuitOptimizer
from qompsim import workloads as W
from qompsim.observable import Observable

pytestmark = pytest.mark.gpu


def _pair(n, B, **kw):
    kw.setdefault("eps_rel", 1e-3)
    kw.setdefault("eps_abs", 1e-7)
    return (CompressedQuantumState(n, B, store="double", verbose=False, **kw),
            CompressedQuantumState(n, B, store="regions", verbose=False, **kw))


def _tfim_step(n, first):
    """One Trotter step of the transverse-field Ising chain (the first one
    starts with X on qubit 1)."""
    c = QuantumCircuit(n)
    if first:
        c.add_X_gate(1)
    for i in range(n - 1):
        c.add_multi_Pauli_rotation_gate([i, i + 1], [3, 3], 0.2)
    for q in range(n):
        c.add_RX_gate(q, 0.3)
    QuantumCircuitOptimizer().optimize(c, 2)
    return c


def _same(a, b):
    assert np.array_equal(a.get_raw_vector().view(np.uint64), b.get_raw_vector().view(np.uint64))
    assert np.array_equal(a.get_null_bitmap(), b.get_null_bitmap())
    assert np.array_equal(a.get_block_bytes(), b.get_block_bytes())
    assert a.get_bytes_history() == b.get_bytes_history()


@pytest.mark.parametrize("n,B,I,seed", [(16, 6, 4, 1), (18, 8, 3, 2), (14, 5, 2, 3)])
def test_regions_match_double_store(n, B, I, seed):
    c = random_circuit(n, 150, np.random.default_rng(seed), max_k=2)
    QuantumCircuitOptimizer().optimize(c, 2)
    a, b = _pair(n, B)
    for cs in (a, b):
        cs.partition_and_register_circuit(c, I)
        cs.apply_circuit()
    _same(a, b)


def test_regions_loaded_state_qft(rng):
    n, B = 16, 7
    v = rand_state(n, rng)
    v[: 3 << B] = 0  # three null blocks at load
    c = W.qft_on_product(n, seed=4)
    QuantumCircuitOptimizer().optimize(c, 2)
    a, b = _pair(n, B)
    for cs in (a, b):
        cs.load(v)
        cs.partition_and_register_circuit(c, 4)
        cs.apply_circuit()
    _same(a, b)


def test_regions_multipass_lds_engine(monkeypatch):
    """Stages of several passes through the staging buffer (64-amplitude LDS
    tiles) write their last pass into the regions too."""
    monkeypatch.setenv("QOMPSIM_ENGINE", "lds")
    n, B = 13, 6
    c = random_circuit(n, 120, np.random.default_rng(77), max_k=2)
    a, b = _pair(n, B, tile_bits=6)
    for cs in (a, b):
        cs.partition_and_register_circuit(c, 3)
        cs.apply_circuit()
    assert max(b.get_stage_passes()) > 1
    _same(a, b)


def test_regions_successive_circuits_and_readouts():
    """Trotter steps registered one after another (the first stage of each
    step consumes the last stage's layout), then the read-out paths."""
    n, B = 18, 8
    a, b = _pair(n, B)
    for s in range(3):
        c = _tfim_step(n, s == 0)
        for cs in (a, b):
            cs.partition_and_register_circuit(c, 4)
            cs.apply_circuit()
        _same(a, b)
    obs = Observable(n)
    obs.add_operator(1.0, "Z 0")
    obs.add_operator(0.5, "X 3 X 12")
    assert a.get_expectation_value(obs) == b.get_expectation_value(obs)
    assert a.get_squared_norm() == b.get_squared_norm()
    assert a.sampling_blockwise(500, 9) == b.sampling_blockwise(500, 9)


def test_regions_basis_and_reset():
    a, b = _pair(12, 6)
    for cs in (a, b):
        cs.set_computational_basis(2345)
    _same(a, b)
    assert b.get_nonnull_blocks() == [2345 >> 6]
    st = b.get_store_bytes()
    assert st["arena"] > 0 and st["arena_peak"] >= st["arena"]


def test_regions_checkpoint_resume_is_exact(tmp_path):
    """Region store checkpoint files: Trotter steps run, saved after step 2,
    restored into a fresh region store and continued, end bit-identical to
    an uninterrupted run; a region store refuses a double-store file."""
    n, B = 18, 8
    full = CompressedQuantumState(n, B, eps_rel=1e-3, eps_abs=1e-7, store="regions", verbose=False)
    steps = [_tfim_step(n, s == 0) for s in range(4)]
    for c in steps:
        full.partition_and_register_circuit(c, 4)
        full.apply_circuit()
    a = CompressedQuantumState(n, B, eps_rel=1e-3, eps_abs=1e-7, store="regions", verbose=False)
    for c in steps[:2]:
        a.partition_and_register_circuit(c, 4)
        a.apply_circuit()
    path = str(tmp_path / "r.qckp")
    a.save_checkpoint(path, next_stage=2)
    b = CompressedQuantumState(n, B, eps_rel=1e-3, eps_abs=1e-7, store="regions", verbose=False)
    assert b.load_checkpoint(path) == 2
    assert np.array_equal(b.get_raw_vector().view(np.uint64), a.get_raw_vector().view(np.uint64))
    assert np.array_equal(b.get_null_bitmap(), a.get_null_bitmap())
    for c in steps[2:]:
        b.partition_and_register_circuit(c, 4)
        b.apply_circuit()
    _same_state(b, full)
    d = CompressedQuantumState(n, B, eps_rel=1e-3, eps_abs=1e-7, store="double", verbose=False)
    with pytest.raises(ValueError, match="store kind"):
        d.load_checkpoint(path)


def _same_state(a, b):
    assert np.array_equal(a.get_raw_vector().view(np.uint64), b.get_raw_vector().view(np.uint64))
    assert np.array_equal(a.get_null_bitmap(), b.get_null_bitmap())
    assert np.array_equal(a.get_block_bytes(), b.get_block_bytes())


def test_regions_peak_below_two_states():
    """Two Trotter steps at 24 qubits: the region store's largest mapping
    stays well below the two whole arenas a double-buffered stage needs
    (input and output of the largest stage)."""
    n, B = 24, 10
    b = CompressedQuantumState(n, B, eps_rel=1e-3, eps_abs=1e-7, store="regions", verbose=False)
    for s in range(2):
        c = _tfim_step(n, s == 0)
        b.partition_and_register_circuit(c, 4)
        b.apply_circuit()
    big = max(b.get_bytes_history())
    st = b.get_store_bytes()
    assert st["arena_peak"] < 1.6 * big, (st, big)


def test_regions_match_double_store_qft26_packed():
    """QFT-26 with commutation-aware stage packing (multi-pass stages, folded
    SWAPs, 1024 blocks): both stores give the same bits."""
    n, B = 26, 16
    c = W.qft_on_product(n, seed=8)
    QuantumCircuitOptimizer().optimize(c, 2)
    a, b = _pair(n, B)
    for cs in (a, b):
        cs.partition_and_register_circuit(c, 4, reorder=True)
        cs.apply_circuit()
    _same(a, b)
    assert max(b.get_stage_passes()) > 1


@pytest.mark.parametrize("n,B", [(10, 10), (12, 9), (13, 10)])
def test_regions_few_block_qubits(n, B):
    """Fewer block qubits than key bits (down to a single block, one region):
    the same bits as the double-buffered store."""
    c = random_circuit(n, 80, np.random.default_rng(n * 31 + B), max_k=2)
    QuantumCircuitOptimizer().optimize(c, 2)
    a, b = _pair(n, B)
    for cs in (a, b):
        cs.partition_and_register_circuit(c, 2)
        cs.apply_circuit()
    _same(a, b)
    b.set_computational_basis(5)
    a.set_computational_basis(5)
    _same(a, b)


@pytest.mark.parametrize("engine", ["wave", "lds"])
def test_regions_retries_read_the_batch_copy(engine, monkeypatch):
    """Every batch's first encode overflows (QOMPSIM_REGION_TIGHT=1): the
    retry reads the batch's input descriptors from their copy (the region
    store's single descriptor table is encoded in place), and the result is
    bit-identical to the double-buffered store, for single-pass stages and
    for stages of several passes through the staging buffer."""
    monkeypatch.setenv("QOMPSIM_REGION_TIGHT", "1")
    if engine == "lds":
        monkeypatch.setenv("QOMPSIM_ENGINE", "lds")
    n, B = 15, 7
    c = W.qft_on_product(n, seed=12)
    QuantumCircuitOptimizer().optimize(c, 2)
    kw = {"tile_bits": 6} if engine == "lds" else {}
    a, b = _pair(n, B, **kw)
    for cs in (a, b):
        cs.partition_and_register_circuit(c, 4)
        cs.apply_circuit()
    _same(a, b)
    if engine == "lds":
        assert max(b.get_stage_passes()) > 1
    # one descriptor table and a batch's copy, against the double store's two
    assert b.get_store_bytes()["descriptors"] <= a.get_store_bytes()["descriptors"]
