"""GPU parity of the compressed stage loop: codec bit-exact against the
oracle's restatement, the contract bound on every double, the GHZ-20 stage
and null-bitmap fixture bit-exact, and compressed runs against the oracle's
stage loop and the exact state."""
import json
import os

import numpy as np
import pytest

import pyoracle as O
from gpu_util import rand_state, random_circuit
from qompsim import CompressedQuantumState, QuantumCircuit, QuantumCircuitOptimizer, QuantumState, gate
from qompsim import workloads as W

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")


@pytest.mark.parametrize("er,ea,guard", [(1e-3, 1e-7, 1), (1e-3, 1e-7, 0), (1e-2, 0.0, 0), (1e-6, 1e-12, 2)])
def test_codec_bit_exact_vs_oracle(rng, er, ea, guard):
    n, B = 14, 9
    x = rng.normal(scale=10 ** rng.uniform(-6, 0, size=1 << n), size=1 << n) + 1j * rng.normal(scale=1e-4, size=1 << n)
    x[: 1 << B] = 0  # first block exactly zero -> null
    x[5000:5100] = 0
    x[7000] = 1.0
    cs = CompressedQuantumState(n, B, eps_rel=er, eps_abs=ea, guard_bits=guard, verbose=False)
    cs.load(x)
    got = cs.get_raw_vector()
    ref, words, _ = O.codec_roundtrip(x.view(np.float64), er, ea, guard)
    ref = ref.view(np.complex128)
    assert np.array_equal(got.view(np.uint64), ref.view(np.uint64)) or np.array_equal(got, ref)
    xd = x.view(np.float64)
    assert np.all(np.abs(got.view(np.float64) - xd) <= er * np.abs(xd) + ea)
    null = cs.get_null_bitmap()
    assert null[0] and not null[1:].any()
    nonnull_chunks = (1 << (B - 5)) * int((~null).sum())
    assert int(cs.get_block_bytes().sum()) == 8 * (words + nonnull_chunks)


def test_codec_negative_zero_chunks(rng):
    """All-zero chunks made of -0.0 carry no payload and must not disturb the
    chunk packed after them (regression: their sign bits were once OR-ed into
    the next chunk's first payload word)."""
    n, B = 14, 10  # one block per group: L = 10 runs the wave-tile engine
    x = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    x *= 2.0 ** -7
    for c0 in (64, 96, 1000, 4096):  # whole 32-amplitude chunks of -0.0
        x[32 * c0: 32 * c0 + 32] = complex(-0.0, -0.0)
    x[32 * 300 + 5] = complex(-0.0, 0.25)
    cs = CompressedQuantumState(n, B, eps_rel=1e-3, eps_abs=1e-7, guard_bits=1, verbose=False)
    cs.load(x)
    got = cs.get_raw_vector()
    ref, _, _ = O.codec_roundtrip(x.view(np.float64), 1e-3, 1e-7, 1)
    assert np.array_equal(got.view(np.float64), ref)
    # the same through a DEC -> ENC stage that leaves the state unchanged
    c = QuantumCircuit(n)
    c.add_gate(gate.DenseMatrix([0], np.eye(2)))
    cs.register_circuit_stages(c, [[0]], [[[b] for b in range(1 << (n - B))]])
    cs.apply_circuit()
    assert np.array_equal(cs.get_raw_vector().view(np.float64), ref)


def test_ghz20_abs_bound_cpu_contract():
    """GHZ-20 under the CPU path's contract: pure absolute bound 1e-7 (SZ3 ABS
    mode, substituted by this codec with eps_rel = 0), all-zero threshold
    1e-6, fidelity mode against the uncompressed state."""
    fx = json.load(open(os.path.join(GOLD, "ghz20_b12_i4.json")))
    c = W.ghz(20)
    QuantumCircuitOptimizer().optimize(c, 2)
    cs = CompressedQuantumState(20, 12, eps_rel=0.0, eps_abs=1e-7, zero_mode="threshold", zero_threshold=1e-6,
                                verbose=False)
    cs.partition_and_register_circuit(c, 4)
    cs.apply_circuit()
    assert cs.get_nonnull_blocks() == fx["final_nonnull_blocks"]
    sv = QuantumState(20)
    c.update_quantum_state(sv)
    exact = sv.get_vector()
    raw = cs.get_raw_vector()
    assert np.max(np.abs(raw.view(np.float64) - exact.view(np.float64))) <= 1e-7
    assert cs.fidelity_with(sv) >= 1 - 1e-12
    oc = O.OracleCompressedState(20, 12, 0.0, 1e-7, 1, "threshold", 1e-6)
    oc.partition_and_register_circuit(c, 4)
    oc.apply_circuit()
    assert np.array_equal(cs.get_null_bitmap(), oc.null_bitmap())


@pytest.mark.parametrize("mode", ["exact", "threshold"])
def test_ghz20_fixture(mode):
    fx = json.load(open(os.path.join(GOLD, "ghz20_b12_i4.json")))
    c = W.ghz(20)
    QuantumCircuitOptimizer().optimize(c, 2)
    cs = CompressedQuantumState(20, 12, eps_rel=1e-3, eps_abs=1e-7, zero_mode=mode, zero_threshold=1e-6,
                                verbose=False)
    cs.partition_and_register_circuit(c, 4, verbose=1)
    assert [len(s) for s in cs.get_stages()] == fx["stage_sizes"]
    groups = cs.get_block_groups()
    assert [len(g) for g in groups] == fx["stage_num_groups"]
    cs.apply_circuit()
    assert cs.get_nonnull_blocks() == fx["final_nonnull_blocks"]
    oc = O.OracleCompressedState(20, 12, 1e-3, 1e-7, 1, mode, 1e-6)
    oc.partition_and_register_circuit(c, 4)
    oc.apply_circuit()
    assert np.array_equal(cs.get_null_bitmap(), oc.null_bitmap())
    v = cs.get_vector()
    for k, (re, im) in fx["final_amplitudes"].items():
        assert abs(v[int(k)] - complex(re, im)) < 1e-6
    assert np.max(np.abs(v - oc.get_vector())) < 1e-12
    assert cs.get_bytes_history() == oc.bytes_history()


@pytest.mark.parametrize("n,B,I,tb", [(16, 8, 4, 0), (16, 8, 2, 6), (18, 10, 4, 7), (14, 6, 3, 5)])
def test_compressed_qft_vs_oracle(n, B, I, tb):
    c = W.qft_on_product(n, seed=2)
    QuantumCircuitOptimizer().optimize(c, 2)
    kw = dict(eps_rel=1e-3, eps_abs=1e-7, guard_bits=1, verbose=False)
    if tb:
        kw["tile_bits"] = tb
    cs = CompressedQuantumState(n, B, **kw)
    cs.partition_and_register_circuit(c, I)
    cs.apply_circuit()
    oc = O.OracleCompressedState(n, B, 1e-3, 1e-7, 1)
    oc.partition_and_register_circuit(c, I)
    assert cs.get_stages() == [s[0] for s in oc.stages()]
    oc.apply_circuit()
    exact = O.apply_circuit(O.zero_state(n), c.gate_list)
    v = cs.get_vector()
    assert abs(np.vdot(exact, v)) > 0.99999
    # same algorithm, rounding differences only: within a few quantisation steps
    assert np.max(np.abs(v - oc.get_vector())) < 1e-6
    assert np.array_equal(cs.get_null_bitmap(), oc.null_bitmap())
    s = QuantumState(n)
    s.load(exact)
    assert abs(cs.fidelity_with(s) - abs(np.vdot(exact, v))) < 1e-12
    assert len(cs.get_stage_stats()) == cs.get_num_stages()


def test_multipass_random_circuit(monkeypatch):
    # LDS tiles of 2^6 amplitudes on every stage: stages with high qubits
    # need several passes through the HBM staging buffer
    monkeypatch.setenv("QOMPSIM_ENGINE", "lds")
    n, B = 13, 7
    c = random_circuit(n, 120, np.random.default_rng(77), max_k=2)
    # near-lossless codec settings: the relative budget would otherwise allow
    # errors of eps_rel * |x| per double
    cs = CompressedQuantumState(n, B, eps_rel=1e-12, eps_abs=1e-10, guard_bits=2, tile_bits=6, verbose=False)
    cs.partition_and_register_circuit(c, 3)
    assert max(cs.get_stage_passes()) > 1
    cs.apply_circuit()
    exact = O.apply_circuit(O.zero_state(n), c.gate_list)
    v = cs.get_vector()
    assert abs(np.vdot(exact, v)) > 1 - 1e-12
    oc = O.OracleCompressedState(n, B, 1e-12, 1e-10, 2)
    oc.partition_and_register_circuit(c, 3)
    oc.apply_circuit()
    assert np.max(np.abs(v - oc.get_vector())) < 1e-9
    assert np.array_equal(cs.get_null_bitmap(), oc.null_bitmap())


def test_register_circuit_stages_explicit(rng):
    n, B = 12, 7
    c = W.qft_on_product(n, seed=3)
    cs = CompressedQuantumState(n, B, verbose=False)
    cs.partition_and_register_circuit(c, 2)
    stages = cs.get_stages()
    groups = cs.get_block_groups()
    cs2 = CompressedQuantumState(n, B, verbose=False)
    cs2.register_circuit_stages(c, stages, groups)
    cs.apply_circuit()
    cs2.apply_circuit()
    assert np.array_equal(cs.get_raw_vector(), cs2.get_raw_vector())
    bad = [[[b] for b in range(1 << (n - B))] for _ in stages]  # one block per group splits gate pairs
    cs3 = CompressedQuantumState(n, B, verbose=False)
    with pytest.raises(IndexError):
        cs3.register_circuit_stages(c, stages, bad)


def test_basis_and_empty_stage():
    cs = CompressedQuantumState(10, 6, verbose=False)
    cs.set_computational_basis(777)
    v = cs.get_raw_vector()
    assert v[777] == 1.0 and np.count_nonzero(v) == 1
    assert cs.get_nonnull_blocks() == [777 >> 6]


@pytest.mark.parametrize("n,B,I", [(16, 10, 4), (18, 11, 3)])
def test_compressed_engines_agree(n, B, I, monkeypatch):
    c = W.qft_on_product(n, seed=5)
    QuantumCircuitOptimizer().optimize(c, 2)
    res = []
    for eng in ("wave", "lds"):
        if eng == "lds":
            monkeypatch.setenv("QOMPSIM_ENGINE", "lds")
        cs = CompressedQuantumState(n, B, eps_rel=1e-3, eps_abs=1e-7, verbose=False)
        cs.partition_and_register_circuit(c, I)
        cs.apply_circuit()
        res.append((cs.get_vector(), cs.get_null_bitmap()))
    assert np.array_equal(res[0][1], res[1][1])
    assert np.max(np.abs(res[0][0] - res[1][0])) < 1e-6
    exact = O.apply_circuit(O.zero_state(n), c.gate_list)
    assert abs(np.vdot(exact, res[0][0])) > 0.99999


def test_compressed_expectation_matches_uncompressed(rng):
    n, B = 14, 9
    c = random_circuit(n, 80, rng, max_k=2)
    sv = QuantumState(n)
    c.update_quantum_state(sv)
    cs = CompressedQuantumState(n, B, eps_rel=1e-12, eps_abs=1e-12, guard_bits=2, verbose=False)
    cs.load(sv.get_vector())
    obs = W.tfim_observable(n)
    obs.add_operator(0.25, f"Z 0 Z {n - 1}")
    obs.add_operator(-0.5, f"Y 3 Z 12")
    got = cs.get_expectation_value(obs)
    ref = obs.get_expectation_value(sv)
    assert abs(got - ref) < 1e-9


def test_tfim_light_cone_compressed_vs_uncompressed():
    """Reduced cfg5: after 12 Trotter steps <Z_0> depends only on qubits
    0..12, so a 16-qubit uncompressed run pins the 20-qubit compressed one
    (and the 20-qubit uncompressed run pins the light-cone argument)."""
    from qompsim.observable import Observable

    z0 = Observable(20)
    z0.add_operator(1.0, "Z 0")
    c20 = W.tfim_trotter(20)
    QuantumCircuitOptimizer().optimize(c20, 2)
    cs = CompressedQuantumState(20, 12, eps_rel=1e-3, eps_abs=1e-7, verbose=False)
    cs.partition_and_register_circuit(c20, 4)
    cs.apply_circuit()
    zc = cs.get_expectation_value(z0) / cs.get_squared_norm()
    sv20 = QuantumState(20)
    c20.update_quantum_state(sv20)
    z20 = z0.get_expectation_value(sv20)
    z16_obs = Observable(16)
    z16_obs.add_operator(1.0, "Z 0")
    sv16 = QuantumState(16)
    W.tfim_trotter(16).update_quantum_state(sv16)
    z16 = z16_obs.get_expectation_value(sv16)
    assert abs(z20 - z16) < 1e-12
    assert abs(zc - z16) < 2e-3
    assert abs(cs.get_squared_norm() - 1.0) < 1e-2


def test_arena_overflow_regrowth_is_exact(rng, monkeypatch):
    """A tiny starting arena forces overflow retries with measured
    sub-arena sizes; the result must equal a run that never overflowed."""
    n, B = 16, 10
    c = random_circuit(n, 60, rng, max_k=2)
    QuantumCircuitOptimizer().optimize(c, 2)
    ref = CompressedQuantumState(n, B, verbose=False)
    ref.partition_and_register_circuit(c, 3)
    ref.apply_circuit()
    monkeypatch.setenv("QOMPSIM_ARENA_KIB", "16")
    cs = CompressedQuantumState(n, B, verbose=False)
    cs.partition_and_register_circuit(c, 3)
    cs.apply_circuit()
    assert np.array_equal(cs.get_raw_vector().view(np.uint64), ref.get_raw_vector().view(np.uint64))
    assert cs.get_bytes_history() == ref.get_bytes_history()


@pytest.mark.parametrize("n,B", [(14, 8), (18, 12)])
def test_sampling_blockwise_matches_oracle(n, B, rng):
    """Same loaded state (bit-exact codec on both sides), same seeds: the GPU
    draws the oracle's samples; null blocks are never sampled."""
    v = rand_state(n, rng)
    v[(np.arange(1 << n) >> B) % 4 == 1] = 0
    v /= np.linalg.norm(v)
    cs = CompressedQuantumState(n, B, eps_rel=1e-3, eps_abs=1e-7, guard_bits=1, verbose=False)
    cs.load(v)
    oc = O.OracleCompressedState(n, B, 1e-3, 1e-7, 1)
    oc.load(v)
    for shots, seed in ((1000, 0), (3, 77), (0, 5), (20000, 123456789)):
        got = cs.sampling_blockwise(shots, seed)
        assert len(got) == shots
        assert got == oc.sampling_blockwise(shots, seed), (shots, seed)
        assert all(((x >> B) % 4) != 1 for x in got)


def test_sampling_blockwise_after_circuit():
    """GHZ-16 run compressed: every sample is 0 or 2^n - 1, about half each."""
    n, B = 16, 8
    c = QuantumCircuit(n)
    c.add_gate(gate.H(0))
    for i in range(n - 1):
        c.add_gate(gate.CNOT(i, i + 1))
    cs = CompressedQuantumState(n, B, verbose=False)
    cs.partition_and_register_circuit(c, 4)
    cs.apply_circuit()
    s = cs.sampling_blockwise(4000, 9)
    assert set(s) <= {0, (1 << n) - 1}
    assert abs(s.count(0) - 2000) < 5 * np.sqrt(1000)


def test_sampling_blockwise_zero_mass():
    cs = CompressedQuantumState(10, 5, verbose=False)
    cs.load(np.zeros(1 << 10, dtype=complex))
    with pytest.raises(ValueError):
        cs.sampling_blockwise(4, 1)


def test_checkpoint_file_resume_is_exact(tmp_path):
    """Run half the stages, save, restore into a fresh store, run the rest:
    bit-identical to an uninterrupted run (vector, bitmap, bytes)."""
    n, B = 18, 10
    c = W.qft_on_product(n, seed=5)
    QuantumCircuitOptimizer().optimize(c, 2)
    kw = dict(eps_rel=1e-3, eps_abs=1e-7, guard_bits=1, verbose=False)
    full = CompressedQuantumState(n, B, **kw)
    full.partition_and_register_circuit(c, 4)
    full.apply_circuit()
    ns = full.get_num_stages()
    half = ns // 2
    a = CompressedQuantumState(n, B, **kw)
    a.partition_and_register_circuit(c, 4)
    a.run_stages(0, half)
    path = str(tmp_path / "store.qckp")
    a.save_checkpoint(path, next_stage=half)
    b = CompressedQuantumState(n, B, **kw)
    b.partition_and_register_circuit(c, 4)
    assert b.load_checkpoint(path) == half
    assert np.array_equal(b.get_raw_vector(), a.get_raw_vector())
    b.run_stages(half, ns)
    assert np.array_equal(b.get_raw_vector(), full.get_raw_vector())
    assert np.array_equal(b.get_null_bitmap(), full.get_null_bitmap())
    assert sum(b.get_block_bytes()) == sum(full.get_block_bytes())


def test_checkpoint_file_rejects_other_store(tmp_path):
    n = 12
    a = CompressedQuantumState(n, 6, verbose=False)
    path = str(tmp_path / "s.qckp")
    a.save_checkpoint(path)
    with pytest.raises(ValueError, match="block size"):
        CompressedQuantumState(n, 7, verbose=False).load_checkpoint(path)
    with pytest.raises(ValueError, match="codec"):
        CompressedQuantumState(n, 6, eps_rel=1e-4, verbose=False).load_checkpoint(path)
    bad = tmp_path / "bad.qckp"
    bad.write_bytes(b"not a checkpoint" * 8)
    with pytest.raises(ValueError, match="magic"):
        a.load_checkpoint(str(bad))


def _ckpt_layout(path):
    """(header fields, offset of the null flags) of a checkpoint file."""
    import struct

    raw = open(path, "rb").read()
    magic, ver, store, n, B, nb, cpb = struct.unpack_from("<8s4I2Q", raw, 0)
    aux0, aux1 = struct.unpack_from("<2Q", raw, 88)
    if store == 0:
        layout = (aux0 + 1) * 8 + aux0 * 16 * 8
    else:
        layout = 4 * 4 + 8 + 2 * aux1 * 8
    return dict(n=n, B=B, nb=nb, cpb=cpb, aux0=aux0, aux1=aux1), 128 + layout


@pytest.mark.parametrize("store", ["double", "regions"])
def test_checkpoint_load_rejects_truncated_and_corrupt_files(tmp_path, store):
    """A truncated file and a file whose descriptor points past the payload
    are refused before the store changes: the state afterwards is the one
    before the call."""
    n, B = 14, 8
    c = W.qft_on_product(n, seed=5)
    QuantumCircuitOptimizer().optimize(c, 2)
    a = CompressedQuantumState(n, B, store=store, verbose=False)
    a.partition_and_register_circuit(c, 4)
    a.apply_circuit()
    path = tmp_path / "a.qckp"
    a.save_checkpoint(str(path), next_stage=3)
    good = path.read_bytes()
    b = CompressedQuantumState(n, B, store=store, verbose=False)
    b.set_computational_basis(77)
    before = b.get_raw_vector()
    (tmp_path / "t.qckp").write_bytes(good[:-9])
    with pytest.raises(RuntimeError, match="size"):
        b.load_checkpoint(str(tmp_path / "t.qckp"))
    assert np.array_equal(b.get_raw_vector(), before)
    hdr, null_off = _ckpt_layout(str(path))
    nullf = np.frombuffer(good, dtype=np.uint8, count=hdr["nb"], offset=null_off)
    blk = int(np.flatnonzero(nullf == 0)[0])
    doff = null_off + hdr["nb"] + blk * hdr["cpb"] * 8
    bad = bytearray(good)
    d = int.from_bytes(bad[doff:doff + 8], "little")
    d = (d & ~((1 << 40) - 1)) | ((1 << 40) - 1)  # payload offset far past the store
    bad[doff:doff + 8] = d.to_bytes(8, "little")
    (tmp_path / "c.qckp").write_bytes(bytes(bad))
    with pytest.raises(ValueError, match="corrupt descriptor"):
        b.load_checkpoint(str(tmp_path / "c.qckp"))
    bad = bytearray(good)
    bad[doff + 5] = 70  # code width 70
    (tmp_path / "w.qckp").write_bytes(bytes(bad))
    with pytest.raises(ValueError, match="corrupt descriptor"):
        b.load_checkpoint(str(tmp_path / "w.qckp"))
    assert np.array_equal(b.get_raw_vector(), before)
    assert b.load_checkpoint(str(path)) == 3
    assert np.array_equal(b.get_raw_vector(), a.get_raw_vector())


def test_compressed_tile_shapes_agree(monkeypatch):
    """Single-pass stages on 512-amplitude (default) and 1024-amplitude tiles:
    same stage results up to the rounding of the phase factors, same bitmap."""
    n, B = 16, 10
    c = W.qft_on_product(n, seed=4)
    QuantumCircuitOptimizer().optimize(c, 2)
    out = {}
    for bits in ("9", "10"):
        monkeypatch.setenv("QOMPSIM_WAVE_TILE_BITS", bits)
        cs = CompressedQuantumState(n, B, eps_rel=1e-12, eps_abs=1e-12, guard_bits=2, verbose=False)
        cs.partition_and_register_circuit(c, 4)
        cs.apply_circuit()
        out[bits] = (cs.get_raw_vector(), cs.get_null_bitmap())
    assert np.max(np.abs(out["9"][0] - out["10"][0])) < 1e-11
    assert np.array_equal(out["9"][1], out["10"][1])
    exact = O.apply_circuit(O.zero_state(n), c.gate_list)
    assert abs(np.vdot(exact, out["9"][0])) > 1 - 1e-10


def test_commuting_partition_runs_compressed():
    """Stages from the commutation-aware packing: fewer stages, and the
    compressed run keeps the fidelity bar against the exact state."""
    n, B = 16, 8
    c = W.qft_on_product(n, seed=6)
    QuantumCircuitOptimizer().optimize(c, 2)
    ref = CompressedQuantumState(n, B, verbose=False)
    ref.partition_and_register_circuit(c, 4)
    cs = CompressedQuantumState(n, B, verbose=False)
    cs.partition_and_register_circuit(c, 4, reorder=True)
    assert cs.get_num_stages() < ref.get_num_stages()
    cs.apply_circuit()
    exact = O.apply_circuit(O.zero_state(n), c.gate_list)
    assert abs(np.vdot(exact, cs.get_vector())) > 0.99999
    # near-lossless codec: the reordered stages reproduce the exact state
    cs2 = CompressedQuantumState(n, B, eps_rel=1e-12, eps_abs=1e-12, guard_bits=2, verbose=False)
    cs2.partition_and_register_circuit(c, 4, reorder=True)
    cs2.apply_circuit()
    assert np.max(np.abs(cs2.get_vector() - exact)) < 1e-9


def test_trailing_swaps_folded_into_encode(monkeypatch):
    """Trailing SWAPs on qubits >= 5 become the encoding pass's chunk
    permutation: fewer passes, and the same state, null flags and bytes as
    applying them as gates (QOMPSIM_STORE_SWAPS=0)."""
    n, B = 18, 10
    c = W.qft_on_product(n, seed=6)  # ends with n/2 SWAPs, some on qubits 0..4
    QuantumCircuitOptimizer().optimize(c, 2)
    res = {}
    for mode in ("1", "0"):
        monkeypatch.setenv("QOMPSIM_STORE_SWAPS", mode)
        cs = CompressedQuantumState(n, B, eps_rel=1e-3, eps_abs=1e-7, verbose=False)
        cs.partition_and_register_circuit(c, 4)
        cs.apply_circuit()
        res[mode] = (cs.get_raw_vector(), cs.get_null_bitmap(), cs.get_bytes_history(), sum(cs.get_stage_passes()))
    assert np.array_equal(res["1"][0].view(np.uint64), res["0"][0].view(np.uint64))
    assert np.array_equal(res["1"][1], res["0"][1])
    assert res["1"][2] == res["0"][2]
    assert res["1"][3] < res["0"][3]
    exact = O.apply_circuit(O.zero_state(n), c.gate_list)
    assert abs(np.vdot(exact, res["1"][0] / np.linalg.norm(res["1"][0]))) > 0.9999


def test_swap_only_stage_moves_blocks_and_null_flags():
    """A stage of SWAPs between a block qubit and an in-block qubit >= 5 on a
    state with null blocks: the data and the null flags follow the swapped
    positions."""
    n, B = 14, 8
    x = np.zeros(1 << n, dtype=np.complex128)
    rng = np.random.default_rng(12)
    x[: 3 << B] = rng.normal(size=3 << B) + 1j * rng.normal(size=3 << B)  # blocks 0..2, rest null
    x /= np.linalg.norm(x)
    c = QuantumCircuit(n)
    c.add_gate(gate.SWAP(6, 9))
    c.add_gate(gate.SWAP(7, 12))
    cs = CompressedQuantumState(n, B, eps_rel=1e-12, eps_abs=1e-14, verbose=False)
    cs.load(x)
    before = cs.get_raw_vector()
    cs.partition_and_register_circuit(c, 4)
    assert cs.get_stage_passes() == [1]
    cs.apply_circuit()
    idx = np.arange(1 << n)
    def swap_bits(v, a, b):
        d = ((v >> a) ^ (v >> b)) & 1
        return v ^ ((d << a) | (d << b))
    perm = swap_bits(swap_bits(idx, 6, 9), 7, 12)
    expect = np.empty_like(before)
    expect[perm] = before
    # one more decode/encode round trip of each chunk: within the codec bound
    assert np.max(np.abs(cs.get_raw_vector() - expect)) < 1e-12
    nz = sorted({int(b) for b in (perm[np.flatnonzero(before)] >> B)})
    assert cs.get_nonnull_blocks() == sorted(set(nz) | {0, 1, 2})


@pytest.mark.parametrize("store", ["double", "regions"])
def test_tile_swaps_in_encode(store, monkeypatch):
    """Trailing SWAPs with a qubit inside the codec chunk (an in-chunk qubit
    with a block qubit, two in-chunk qubits, lane qubit 5 with a chunk qubit)
    become the encoder's tile permutation: one op-free pass whose output is
    the permuted state, null flags following the data."""
    n, B = 14, 8
    x = np.zeros(1 << n, dtype=np.complex128)
    rng = np.random.default_rng(21)
    x[: 3 << B] = rng.normal(size=3 << B) + 1j * rng.normal(size=3 << B)  # blocks 0..2, rest null
    x /= np.linalg.norm(x)
    pairs = [(1, 9), (3, 12), (0, 2), (5, 4)]
    c = QuantumCircuit(n)
    for a, b in pairs:
        c.add_gate(gate.SWAP(a, b))
    cs = CompressedQuantumState(n, B, eps_rel=1e-12, eps_abs=1e-14, store=store, verbose=False)
    cs.load(x)
    before = cs.get_raw_vector()
    cs.partition_and_register_circuit(c, 4)
    assert cs.get_stage_passes() == [1]
    cs.apply_circuit()
    idx = np.arange(1 << n)
    def swap_bits(v, a, b):
        d = ((v >> a) ^ (v >> b)) & 1
        return v ^ ((d << a) | (d << b))
    perm = idx
    for a, b in pairs:
        perm = swap_bits(perm, a, b)
    expect = np.empty_like(before)
    expect[perm] = before
    assert np.max(np.abs(cs.get_raw_vector() - expect)) < 1e-12
    nz = sorted({int(b) for b in (perm[np.flatnonzero(before)] >> B)})
    assert cs.get_nonnull_blocks() == sorted(set(nz) | {0, 1, 2})
    # the same circuit with the swaps applied as gates gives the same bits
    monkeypatch.setenv("QOMPSIM_STORE_SWAPS", "0")
    cs2 = CompressedQuantumState(n, B, eps_rel=1e-12, eps_abs=1e-14, store=store, verbose=False)
    cs2.load(x)
    cs2.partition_and_register_circuit(c, 4)
    cs2.apply_circuit()
    assert np.array_equal(cs2.get_raw_vector().view(np.uint64), cs.get_raw_vector().view(np.uint64))
    assert cs2.get_nonnull_blocks() == cs.get_nonnull_blocks()


@pytest.mark.parametrize("store", ["double", "regions"])
def test_hbm_budget(store):
    """hbm_budget bounds what the store holds: below the fixed descriptor and
    staging bytes the constructor refuses; a budget the run fits in gives the
    unbudgeted result; one it does not fit in raises RuntimeError."""
    from qompsim import _lib

    n, B = 18, 10
    c = W.qft_on_product(n, seed=9)
    QuantumCircuitOptimizer().optimize(c, 2)
    kw = dict(eps_rel=1e-3, eps_abs=1e-7, scratch_bytes=1 << 22, store=store, verbose=False)
    with pytest.raises(ValueError, match="budget"):
        CompressedQuantumState(n, B, hbm_budget=1 << 20, **kw)
    ref = CompressedQuantumState(n, B, **kw)
    ref.partition_and_register_circuit(c, 4)
    ref.apply_circuit()
    big = CompressedQuantumState(n, B, hbm_budget=64 << 20, **kw)
    big.partition_and_register_circuit(c, 4)
    big.apply_circuit()
    assert np.array_equal(big.get_raw_vector().view(np.uint64), ref.get_raw_vector().view(np.uint64))
    st = big.get_store_bytes()
    assert st["arena_peak"] + st["descriptors"] + (1 << 22) <= 64 << 20
    fixed = 2 * (1 << (n - B)) * (1 << (B - 5)) * 8 + (1 << 22)  # descriptor tables + staging
    with pytest.raises(_lib.QsimError, match="budget"):
        # 256 KiB of payload room: the region store's first mapping of its
        # regions, or the double store's arenas during the run, do not fit
        small = CompressedQuantumState(n, B, hbm_budget=fixed + (256 << 10), **kw)
        small.partition_and_register_circuit(c, 4)
        small.apply_circuit()
