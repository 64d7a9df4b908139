"""Pins of the CPU oracle (no GPU): the reference's own compiled csim kernels,
closed forms, kron-built unitaries, brute force, the derived GHZ-20 fixture
and the codec contract.  A plausible mistake in the oracle (wrong matrix bit
order, sign of a rotation, dropped control, wrong stage boundary, sticky-null
rule) fails one of these."""
import json
import os

import numpy as np
import pytest

import pyoracle as O
from qompsim import Observable, ParametricQuantumCircuit, QuantumCircuit, QuantumCircuitOptimizer, gate
from qompsim import workloads as W

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def rand_state(n, rng):
    v = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    return v / np.linalg.norm(v)


def rand_unitary(d, rng):
    q, r = np.linalg.qr(rng.normal(size=(d, d)) + 1j * rng.normal(size=(d, d)))
    return q * (np.diag(r) / np.abs(np.diag(r)))


def random_gates(n, count, rng, max_k=3, controls=True):
    gs = []
    for _ in range(count):
        k = int(rng.integers(1, max_k + 1))
        qs = rng.choice(n, size=min(n, k + 2), replace=False)
        g = gate.DenseMatrix(list(qs[:k]), rand_unitary(1 << k, rng))
        if controls and len(qs) > k and rng.random() < 0.4:
            g.add_control_qubit(int(qs[k]), int(rng.integers(0, 2)))
        gs.append(g)
    return gs


def kron_unitary(n, g):
    """Full 2^n matrix of a gate by explicit basis action (independent of the oracle loops)."""
    t, c, cv, m = g.get_target_index_list(), g.get_control_index_list(), g.get_control_value_list(), g.get_matrix()
    D = 1 << n
    U = np.zeros((D, D), dtype=complex)
    for col in range(D):
        if any(((col >> q) & 1) != v for q, v in zip(c, cv)):
            U[col, col] = 1
            continue
        cg = sum(((col >> q) & 1) << i for i, q in enumerate(t))
        for rg in range(1 << len(t)):
            row = col
            for i, q in enumerate(t):
                row = (row & ~(1 << q)) | (((rg >> i) & 1) << q)
            U[row, col] += m[rg, cg]
    return U


# ------------------------------------------------------------- gate kernels
@pytest.mark.skipif(not os.path.exists(O.REF_SO), reason="reference csim library (oracle/_ref) not built")
def test_oracle_matches_reference_csim(rng):
    ref = O.RefCsim()
    for n in (3, 7, 14):
        psi = rand_state(n, rng)
        gs = random_gates(n, 25, rng, max_k=min(3, n - 1))
        a = O.apply_circuit(psi, gs)
        b = psi.copy()
        for g in gs:
            ref.apply_gate(b, g)
        assert np.max(np.abs(a - b)) < 1e-13


@pytest.mark.skipif(not os.path.exists(O.REF_SO), reason="reference csim library (oracle/_ref) not built")
def test_named_rotations_match_reference_csim(rng):
    ref = O.RefCsim()
    n = 5
    psi = rand_state(n, rng)
    for fac, kern in ((gate.RX, ref.rx), (gate.RY, ref.ry), (gate.RZ, ref.rz)):
        a = O.apply_circuit(psi, [fac(2, 0.37)])
        b = psi.copy()
        # csim RX/RY/RZ_gate(angle) apply exp(+i angle P/2) (update_ops_pauli_single.cpp:54-125)
        kern(2, 0.37, b.ctypes.data, b.size)
        assert np.max(np.abs(a - b)) < 1e-14
    a = O.apply_circuit(psi, [gate.CNOT(1, 3), gate.CZ(0, 4), gate.H(2)])
    b = psi.copy()
    ref.cnot(1, 3, b.ctypes.data, b.size)
    ref.cz(0, 4, b.ctypes.data, b.size)
    ref.h(2, b.ctypes.data, b.size)
    assert np.max(np.abs(a - b)) < 1e-14


def test_oracle_matches_kron_unitaries(rng):
    n = 5
    psi = rand_state(n, rng)
    gs = random_gates(n, 12, rng)
    U = np.eye(1 << n, dtype=complex)
    for g in gs:
        U = kron_unitary(n, g) @ U
    assert np.max(np.abs(O.apply_circuit(psi, gs) - U @ psi)) < 1e-13


def test_rotation_convention():
    # R_P(theta) = exp(+i theta P/2): RX(pi)|0> = i|1>, RY(pi)|0> = -|1>, RZ(pi)|0> = i|0>
    z = O.zero_state(1)
    assert np.allclose(O.apply_circuit(z, [gate.RX(0, np.pi)]), [0, 1j])
    assert np.allclose(O.apply_circuit(z, [gate.RY(0, np.pi)]), [0, -1])
    assert np.allclose(O.apply_circuit(z, [gate.RZ(0, np.pi)]), [1j, 0])
    assert np.allclose(O.apply_circuit(z, [gate.RotX(0, np.pi)]), [0, -1j])


def test_ghz_closed_form():
    n = 12
    v = O.apply_circuit(O.zero_state(n), W.ghz(n).gate_list)
    ref = np.zeros(1 << n, dtype=complex)
    ref[0] = ref[-1] = 1 / np.sqrt(2)
    assert np.max(np.abs(v - ref)) < 1e-15


def test_basis_permutations_exact(rng):
    n = 8
    for _ in range(5):
        x = int(rng.integers(0, 1 << n))
        v = np.zeros(1 << n, dtype=complex)
        v[x] = 1
        c, t = rng.choice(n, 2, replace=False)
        out = O.apply_circuit(v, [gate.CNOT(int(c), int(t)), gate.X(3), gate.SWAP(0, 5)])
        y = x ^ ((1 << int(t)) if (x >> int(c)) & 1 else 0)
        y ^= 1 << 3
        b0, b5 = (y >> 0) & 1, (y >> 5) & 1
        y = (y & ~0b100001) | (b0 << 5) | (b5 << 0)
        assert out[y] == 1 and np.count_nonzero(out) == 1


def test_qft_is_unitary_dft_fixture():
    d = np.load(os.path.join(GOLD, "qft_product_n10.npz"))
    c = W.qft_on_product(10, seed=2)
    v = O.apply_circuit(O.zero_state(10), c.gate_list)
    assert np.max(np.abs(v - d["output"])) < 1e-13
    layer = O.apply_circuit(O.zero_state(10), W.product_layer(10, 2)[0])
    assert np.max(np.abs(layer - d["input"])) < 1e-14


def test_qft_vs_numpy_fft(rng):
    for n in (6, 13):
        psi = rand_state(n, rng)
        out = O.apply_circuit(psi, W.qft_gates(n))
        assert np.max(np.abs(out - np.fft.ifft(psi, norm="ortho"))) < 1e-12


def test_norm_inner_compensated(rng):
    psi = rand_state(16, rng)
    phi = rand_state(16, rng)
    assert abs(O.norm(psi) - 1.0) < 1e-14
    assert abs(O.inner(psi, phi) - np.vdot(psi, phi)) < 1e-14


def dense_pauli(n, term):
    P = np.eye(1, dtype=complex)
    ops = {0: np.eye(2), 1: np.array([[0, 1], [1, 0]]), 2: np.array([[0, -1j], [1j, 0]]), 3: np.diag([1, -1])}
    ids = dict(zip(term.index_list, term.pauli_ids))
    for q in range(n):
        P = np.kron(ops[ids.get(q, 0)], P)
    return term.coef * P


def test_expectation_and_apply_bruteforce(rng):
    n = 5
    obs = Observable(n)
    obs.add_operator(0.3, "X 0 Y 2 Z 4")
    obs.add_operator(-1.1, "Z 1 Z 3")
    obs.add_operator(0.7, "Y 1")
    obs.add_operator(0.25, "")
    H = sum(dense_pauli(n, t) for t in obs.terms)
    psi = rand_state(n, rng)
    assert abs(O.expectation(psi, obs) - np.vdot(psi, H @ psi)) < 1e-14
    assert np.max(np.abs(O.apply_pauli_sum(psi, obs) - H @ psi)) < 1e-14


def test_backprop_vs_finite_differences():
    n = 4
    c = W.hea(n, layers=2, seed=5)
    obs = W.tfim_observable(n)
    e, g = O.backprop(c, obs)
    th = c.get_parameters()
    h = 1e-6
    for i in range(c.get_parameter_count()):
        cp = c.copy()
        cp.set_parameter(i, th[i] + h)
        ep = O.expectation(O.apply_circuit(O.zero_state(n), cp.gate_list), obs).real
        cp.set_parameter(i, th[i] - h)
        em = O.expectation(O.apply_circuit(O.zero_state(n), cp.gate_list), obs).real
        assert abs((ep - em) / (2 * h) - g[i]) < 1e-8
    assert abs(e - O.expectation(O.apply_circuit(O.zero_state(n), c.gate_list), obs).real) < 1e-14


# -------------------------------------------------------------------- codec
ADVERSARIAL = np.array([0.0, -0.0, 1.0, -1.0, 1 / np.sqrt(2), -1 / np.sqrt(2), 5e-324, -5e-324, 1e-7, -1e-7,
                        1e-4, 9.99e-5, 1.0 - 1e-16, 2.2e-308, 3e-8, 1e-30], dtype=np.float64)


@pytest.mark.parametrize("er,ea,guard", [(1e-3, 1e-7, 1), (1e-3, 1e-7, 0), (1e-2, 0.0, 0), (1e-6, 1e-12, 2)])
def test_codec_error_bound(rng, er, ea, guard):
    x = np.concatenate([ADVERSARIAL, rng.normal(scale=1e-3, size=48), rng.normal(scale=1e-5, size=64),
                        rng.uniform(-1, 1, size=64), np.zeros(64)])
    y, words, widths = O.codec_roundtrip(x, er, ea, guard)
    assert np.all(np.abs(y - x) <= er * np.abs(x) + ea)
    assert np.all(y[x == 0] == 0.0)
    assert widths[-1] == 0
    assert words == int(widths.astype(np.int64).sum())


def test_codec_compresses_small_amplitudes(rng):
    x = rng.normal(scale=1.5e-5, size=64 * 64)
    _, words, _ = O.codec_roundtrip(x, 1e-3, 1e-7, 1)
    assert words * 64 < x.size * 64 / 4  # under a quarter of the raw bits


# -------------------------------------------------- stage loop restatement
def test_ghz20_stages_and_bitmap_fixture():
    fx = json.load(open(os.path.join(GOLD, "ghz20_b12_i4.json")))
    c = W.ghz(20)
    QuantumCircuitOptimizer().optimize(c, 2)
    assert c.get_gate_count() == fx["optimized_gate_count"]
    assert c.gate_list[0].get_target_index_list() == fx["optimized_first_gate_targets"]
    for mode in ("exact", "threshold"):
        cs = O.OracleCompressedState(20, 12, eps_rel=1e-3, eps_abs=1e-7, zero_mode=mode, zero_threshold=1e-6)
        assert cs.partition_and_register_circuit(c, 4) == 3
        st = cs.stages()
        assert [len(s[0]) for s in st] == fx["stage_sizes"]
        assert [s[2] for s in st] == fx["stage_num_groups"]
        for (ids, b2g, ng), outer in zip(st, fx["stage_outer_qubits"]):
            # group id bit i <-> outer qubit i (compressed_state.hpp:645-655)
            for b in (0, 5, 77, 200, 255):
                assert b2g[b] == sum(((b >> (q - 12)) & 1) << i for i, q in enumerate(outer))
        cs.apply_circuit()
        nonnull = sorted(np.nonzero(~cs.null_bitmap())[0].tolist())
        assert nonnull == fx["final_nonnull_blocks"]
        v = cs.get_vector()
        for k, (re, im) in fx["final_amplitudes"].items():
            assert abs(v[int(k)] - complex(re, im)) < 1e-6


def test_oracle_stage_loop_fidelity_and_bound(rng):
    n, B = 14, 8
    c = W.qft_on_product(n, seed=7)
    QuantumCircuitOptimizer().optimize(c, 2)
    cs = O.OracleCompressedState(n, B, eps_rel=1e-3, eps_abs=1e-7, guard=1)
    cs.partition_and_register_circuit(c, 2)
    cs.apply_circuit()
    exact = O.apply_circuit(O.zero_state(n), c.gate_list)
    v = cs.get_vector()
    fid = abs(np.vdot(exact, v))
    assert fid > 0.99999
    assert len(cs.bytes_history()) == len(cs.stages()) + 1


# ---------------------------------------------------------- block sampling
def test_sampling_blockwise_basis_and_null_blocks():
    """A basis state samples only its index; indices of null blocks never
    appear; the count is exact and the draw is a function of the seed."""
    n, B = 11, 5
    cs = O.OracleCompressedState(n, B)
    v = np.zeros(1 << n, dtype=complex)
    v[613] = 1.0
    cs.load(v)
    assert cs.sampling_blockwise(257, 5) == [613] * 257
    assert cs.sampling_blockwise(0, 5) == []
    rng = np.random.default_rng(3)
    w = rand_state(n, rng)
    w[(np.arange(1 << n) >> B) % 3 == 0] = 0  # every third block all-zero
    cs.load(w / np.linalg.norm(w))
    s1 = cs.sampling_blockwise(4000, 11)
    assert len(s1) == 4000 and s1 == cs.sampling_blockwise(4000, 11)
    assert s1 != cs.sampling_blockwise(4000, 12)
    assert all(((x >> B) % 3) != 0 for x in s1)


def test_sampling_blockwise_distribution():
    """Empirical frequencies match |decoded amplitude|^2 within 5 sigma; the
    draw is not a permutation of a fixed order (first half and second half
    agree in distribution)."""
    n, B = 7, 5
    rng = np.random.default_rng(8)
    w = rand_state(n, rng) * np.sqrt(np.linspace(0.2, 2.0, 1 << n))
    w /= np.linalg.norm(w)
    cs = O.OracleCompressedState(n, B, eps_rel=1e-12, eps_abs=1e-14, guard=2)
    cs.load(w)
    p = np.abs(cs.get_vector(normalise=False)) ** 2
    p /= p.sum()
    N = 200000
    s = np.array(cs.sampling_blockwise(N, 2024))
    f = np.bincount(s, minlength=1 << n) / N
    sig = np.sqrt(p * (1 - p) / N)
    assert np.all(np.abs(f - p) < 5 * sig + 1e-12)
    h1 = np.bincount(s[: N // 2], minlength=1 << n) / (N // 2)
    assert np.max(np.abs(h1 - p)) < 6 * np.max(np.sqrt(p * (1 - p) / (N // 2)))


def test_sampling_blockwise_zero_mass():
    n, B = 8, 5
    cs = O.OracleCompressedState(n, B)
    cs.load(np.zeros(1 << n, dtype=complex))
    with pytest.raises(ValueError):
        cs.sampling_blockwise(10, 1)
    with pytest.raises(ValueError):
        O.OracleCompressedState(8, 4)
