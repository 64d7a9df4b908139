"""GPU adjoint gradients (qsim_sv_backprop) against the oracle's restated
backprop (src/vqcsim/parametric_circuit.cpp:180-308) and central finite
differences (h=1e-5) on the uncompressed FP64 path, plus the torch.autograd
wrapper."""
import numpy as np
import pytest

import pyoracle as O
from qompsim import QuantumState, _lib
from qompsim import gate as G
from qompsim import workloads as W
from qompsim.adjoint import adjoint_gradient
from qompsim.circuit import ParametricQuantumCircuit
from qompsim.observable import Observable

pytestmark = pytest.mark.gpu


def _fd(c, obs, i, h=1e-5):
    th = c.get_parameters()
    vals = []
    for d in (h, -h):
        cp = c.copy()
        cp.set_parameter(i, th[i] + d)
        s = QuantumState(c.get_qubit_count())
        cp.update_quantum_state(s)
        vals.append(obs.get_expectation_value(s).real)
    return (vals[0] - vals[1]) / (2 * h)


def _mixed_circuit(n, rng, layers=3):
    """Parametric X/Y/Z and multi-qubit Pauli rotations between fixed gates."""
    c = ParametricQuantumCircuit(n)
    for _ in range(layers):
        for q in range(n):
            c.add_gate(G.H(q)) if rng.random() < 0.3 else None
            kind = rng.integers(3)
            ang = rng.uniform(0, 2 * np.pi)
            (c.add_parametric_RX_gate, c.add_parametric_RY_gate, c.add_parametric_RZ_gate)[kind](q, ang)
        for q in range(0, n - 1, 2):
            c.add_gate(G.CNOT(q, q + 1))
        qs = [int(x) for x in rng.choice(n, size=3, replace=False)]
        c.add_parametric_multi_Pauli_rotation_gate(qs, [int(x) for x in rng.integers(1, 4, size=3)],
                                                   rng.uniform(0, 2 * np.pi))
        a, b = (int(x) for x in rng.choice(n, size=2, replace=False))
        c.add_parametric_multi_Pauli_rotation_gate([a, b], [3, 3], rng.uniform(0, 2 * np.pi))
    return c


def _mixed_observable(n, rng):
    obs = Observable(n)
    for _ in range(6):
        qs = rng.choice(n, size=2, replace=False)
        ps = rng.choice(["X", "Y", "Z"], size=2)
        obs.add_operator(float(rng.normal()), f"{ps[0]} {qs[0]} {ps[1]} {qs[1]}")
    for q in range(n):
        obs.add_operator(0.3, f"X {q}")
    return obs


def test_adjoint_matches_oracle_and_fd():
    n = 8
    c = W.hea(n, layers=2, seed=4)
    obs = W.tfim_observable(n)
    e, g = adjoint_gradient(c, obs)
    e0, g0 = O.backprop(c, obs)
    assert abs(e - e0) < 1e-12
    assert np.max(np.abs(g - g0)) < 1e-10
    th = c.get_parameters()
    for i in range(0, len(th), 5):
        assert abs(_fd(c, obs, i) - g[i]) < 1e-6


@pytest.mark.parametrize("n", [9, 12])
def test_adjoint_mixed_paulis_vs_oracle(n, rng):
    c = _mixed_circuit(n, rng)
    obs = _mixed_observable(n, rng)
    e, g = adjoint_gradient(c, obs)
    e0, g0 = O.backprop(c, obs)
    assert abs(e - e0) < 1e-11
    assert np.max(np.abs(g - g0)) < 1e-10
    for i in (0, len(g) // 2, len(g) - 1):
        assert abs(_fd(c, obs, i) - g[i]) < 1e-6


@pytest.mark.parametrize("tb", ["4", "6", "8", "11"])
def test_adjoint_tile_sizes(tb, rng, monkeypatch):
    """Small tiles force many passes, gradient points on base bits and
    several points per pass."""
    monkeypatch.setenv("QOMPSIM_ADJ_TILE_BITS", tb)
    monkeypatch.setenv("QOMPSIM_ADJ_ENGINE", "lds")
    n = 12
    c = _mixed_circuit(n, rng, layers=2)
    obs = _mixed_observable(n, rng)
    e, g = adjoint_gradient(c, obs)
    e0, g0 = O.backprop(c, obs)
    assert abs(e - e0) < 1e-11
    assert np.max(np.abs(g - g0)) < 1e-10


def test_adjoint_from_given_state_and_final_state(rng):
    n = 10
    c = W.hea(n, layers=2, seed=7)
    obs = W.tfim_observable(n)
    v = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    v /= np.linalg.norm(v)
    e, g = adjoint_gradient(c, obs, state=v)
    e0, g0 = O.backprop(c, obs, init=v)
    assert abs(e - e0) < 1e-11
    assert np.max(np.abs(g - g0)) < 1e-10


def test_backprop_rejects_non_rotation_parametric():
    n = 5
    c = ParametricQuantumCircuit(n)
    c.add_parametric_RY_gate(0, 0.3)
    rec, pool = _lib.pack_gates(c.gate_list)
    rec["pauli_ids"][0][:] = 0  # a parametric gate without a Pauli axis
    obs = W.tfim_observable(n)
    xm, zm, ny, cr, ci = obs._arrays()
    s = QuantumState(n)
    grads = np.zeros(1)
    import ctypes
    e = ctypes.c_double()
    with pytest.raises(ValueError, match="backprop"):
        _lib.call("qsim_sv_backprop", s._h, _lib.ptr(rec), 1, _lib.ptr(pool), pool.size, len(obs.terms), _lib.ptr(xm),
                  _lib.ptr(zm), _lib.ptr(ny), _lib.ptr(cr), _lib.ptr(ci), 1, ctypes.byref(e), _lib.ptr(grads))


def test_torch_autograd_function():
    import torch

    from qompsim.torch_adjoint import expectation

    n = 8
    c = W.hea(n, layers=2, seed=5)
    obs = W.tfim_observable(n)
    th = torch.tensor(c.get_parameters(), dtype=torch.float64, requires_grad=True)
    E = expectation(th, c, obs)
    (3.0 * E).backward()
    e0, g0 = O.backprop(c, obs)
    assert abs(E.item() - e0) < 1e-12
    assert np.max(np.abs(th.grad.numpy() - 3.0 * g0)) < 1e-10


def test_torch_autograd_function_with_register():
    """Two steps through one reused working state (`register`): each equals
    the oracle at its parameters, and the register holds U|0...0> after."""
    import torch

    from qompsim.torch_adjoint import expectation

    n = 10
    c = W.hea(n, layers=2, seed=7)
    obs = W.tfim_observable(n)
    reg = QuantumState(n)
    for shift in (0.0, 0.25):
        th = torch.tensor(np.asarray(c.get_parameters()) + shift, dtype=torch.float64, requires_grad=True)
        E = expectation(th, c, obs, register=reg)
        E.backward()
        e0, g0 = O.backprop(c, obs)
        assert abs(E.item() - e0) < 1e-12
        assert np.max(np.abs(th.grad.numpy() - g0)) < 1e-10
    ref = O.apply_circuit(O.zero_state(n), c.gate_list)
    assert np.max(np.abs(reg.get_vector() - ref)) < 1e-12


@pytest.mark.parametrize("n", [10, 13, 16])
def test_adjoint_engines_agree(n, rng, monkeypatch):
    """Paired-wave, single-wave (512-amplitude tiles) and LDS-tile reverse
    sweeps give the oracle's gradients."""
    c = _mixed_circuit(n, rng, layers=3)
    obs = _mixed_observable(n, rng)
    e0, g0 = O.backprop(c, obs)
    got = {}
    for eng in ("", "pair9", "wave9", "lds"):
        monkeypatch.setenv("QOMPSIM_ADJ_ENGINE", eng)
        got[eng] = adjoint_gradient(c, obs)
    for eng, (e, g) in got.items():
        assert abs(e - e0) < 1e-11, eng
        assert np.max(np.abs(g - g0)) < 1e-10, eng


def _two_qubit_circuit(n, rng, layers=3):
    """Parametric X/Y/Z rotations, two-qubit Pauli rotations (ZZ, XX, XY, YZ)
    and fixed H, CNOT and CZ gates: every generator and gate has at most two
    X/Y qubits, the shape the lambda-only reverse sweep takes."""
    c = ParametricQuantumCircuit(n)
    for _ in range(layers):
        for q in range(n):
            if rng.random() < 0.3:
                c.add_gate(G.H(q))
            kind = rng.integers(3)
            (c.add_parametric_RX_gate, c.add_parametric_RY_gate, c.add_parametric_RZ_gate)[kind](
                q, rng.uniform(0, 2 * np.pi))
        for q in range(0, n - 1, 2):
            c.add_gate(G.CNOT(q, q + 1))
        for q in range(1, n - 1, 3):
            c.add_gate(G.CZ(q, q + 1))
        for ps in ([3, 3], [1, 1], [1, 2], [2, 3]):
            a, b = (int(x) for x in rng.choice(n, size=2, replace=False))
            c.add_parametric_multi_Pauli_rotation_gate([a, b], ps, rng.uniform(0, 2 * np.pi))
    return c


def test_paired_points_all_placements(rng, monkeypatch):
    """Two-qubit Pauli rotation points on the paired kernel (12 qubits: the
    1024-amplitude wave tile; every gate has at most two non-diagonal
    targets) for qubit pairs on lane bits, register rows and base bits and
    every ordered Pauli pair: the Z/Y signs split into a row fold, a uniform
    part and a lane part, and the flips into a row flip and a lane flip.
    Gradients within 1e-10 of the oracle's backprop."""
    monkeypatch.delenv("QOMPSIM_ADJ_ENGINE", raising=False)
    n = 12
    c = ParametricQuantumCircuit(n)
    for q in range(n):
        c.add_parametric_RY_gate(q, rng.uniform(0, 2 * np.pi))
        c.add_parametric_RX_gate(q, rng.uniform(0, 2 * np.pi))
    for q in range(n - 1):
        c.add_gate(G.CNOT(q, q + 1))
    pairs = [(0, 1), (2, 7), (6, 8), (3, 11), (9, 10), (5, 10), (7, 4), (11, 6)]
    for i, (a, b) in enumerate(pairs):
        for pa, pb in ((1, 2), (2, 3), (3, 1), (1, 1), (2, 2), (3, 3), (2, 1), (3, 2)):
            c.add_parametric_multi_Pauli_rotation_gate([a, b], [pa, pb], rng.uniform(0, 2 * np.pi))
        c.add_gate(G.H((i * 5) % n))
    obs = _mixed_observable(n, rng)
    e, g = adjoint_gradient(c, obs)
    e0, g0 = O.backprop(c, obs)
    assert abs(e - e0) < 1e-11
    assert np.max(np.abs(g - g0)) < 1e-10


@pytest.mark.parametrize("n", [12, 16])
def test_shared_parameter_z_points_vs_oracle(n, rng, monkeypatch):
    """QAOA MaxCut with one parameter per layer angle on the paired kernel:
    a cost layer's RZZ points form a run of Z-only points of one parameter,
    which the kernel evaluates from one Walsh transform of the run's
    products; the mixer's RX points take the flip path.  Gradients within
    1e-10 of the oracle's backprop."""
    monkeypatch.delenv("QOMPSIM_ADJ_ENGINE", raising=False)
    p = 3
    edges = W.random_3regular(n, rng)
    c = W.qaoa_maxcut_shared(n, edges, rng.uniform(0, np.pi, p), rng.uniform(0, np.pi, p))
    obs = W.maxcut_observable(n, edges)
    e, g = adjoint_gradient(c, obs)
    e0, g0 = O.backprop(c, obs)
    assert abs(e - e0) < 1e-11
    assert np.max(np.abs(g[: 2 * p] - g0[: 2 * p])) < 1e-10


@pytest.mark.parametrize("n", [10, 12, 15])
def test_lambda_only_sweep_matches_oracle(n, rng, monkeypatch, capfd):
    """The lambda-only reverse sweep (QOMPSIM_ADJ_ENGINE=lam: psi kept at
    reverse pass starts) gives the oracle's energy and gradients, the paired
    sweep's gradients, and leaves the final state in the register; the HEA and
    a circuit of two-qubit Pauli rotations both run on it."""
    monkeypatch.setenv("QOMPSIM_PLAN_LOG", "1")
    for c, obs in ((_two_qubit_circuit(n, rng), _mixed_observable(n, rng)),
                   (W.hea(n, layers=3, seed=n), W.tfim_observable(n))):
        e0, g0 = O.backprop(c, obs)
        monkeypatch.setenv("QOMPSIM_ADJ_ENGINE", "lam")
        s = QuantumState(n)
        e, g = adjoint_gradient(c, obs, register=s)
        assert "lambda-only sweep" in capfd.readouterr().err
        assert abs(e - e0) < 1e-11
        assert np.max(np.abs(g - g0)) < 1e-10
        exact = O.apply_circuit(O.zero_state(n), c.gate_list)
        assert np.max(np.abs(s.get_vector() - exact)) < 1e-12
        monkeypatch.setenv("QOMPSIM_ADJ_ENGINE", "pair")
        e2, g2 = adjoint_gradient(c, obs)
        assert "lambda-only sweep" not in capfd.readouterr().err
        assert np.max(np.abs(g2 - g)) < 1e-10


def _energy(c, obs, th):
    cp = c.copy()
    cp.set_parameters(th)
    s = QuantumState(c.get_qubit_count())
    cp.update_quantum_state(s)
    return obs.get_expectation_value(s).real


def test_adjoint_vs_parameter_shift():
    """Pauli rotations: dE/dtheta = (E(theta + pi/2) - E(theta - pi/2)) / 2."""
    n = 10
    c = W.hea(n, layers=3, seed=11)
    obs = W.tfim_observable(n)
    e, g = adjoint_gradient(c, obs)
    th = np.asarray(c.get_parameters())
    for i in (0, 7, len(th) // 2, len(th) - 1):
        tp, tm = th.copy(), th.copy()
        tp[i] += np.pi / 2
        tm[i] -= np.pi / 2
        shift = (_energy(c, obs, tp) - _energy(c, obs, tm)) / 2
        assert abs(shift - g[i]) < 1e-10, i


@pytest.mark.parametrize("seg", [1, 9, 83, 10**6])
def test_checkpointed_adjoint_matches_oracle(seg, rng):
    """forward_with_checkpoints + adjoint: segments of one gate, a few gates,
    about one layer, and the whole circuit give the oracle's gradients."""
    from qompsim.adjoint import CheckpointedAdjoint

    n = 12
    c = W.hea(n, layers=3, seed=5)
    obs = W.tfim_observable(n)
    adj = CheckpointedAdjoint(c, seg, block_size=8)
    adj.forward_with_checkpoints()
    e, g = adj.adjoint(obs)
    e0, g0 = O.backprop(c, obs)
    assert abs(e - e0) < 1e-11
    assert np.max(np.abs(g - g0)) < 1e-10
    k, nbytes = adj.checkpoint_info()
    assert k == max(1, -(-len(c.gate_list) // seg))
    assert nbytes > 0
    # a second adjoint call on the same forward run gives the same result
    e2, g2 = adj.adjoint(obs)
    assert e2 == e and np.array_equal(g2, g)


def test_checkpointed_forward_reuses_stores(rng):
    """Repeated forwards on one handle reuse its checkpoint stores (same qubit
    count, block size and codec; a load resets a store), with new parameters,
    more segments and then fewer: the oracle's gradients after each forward,
    and as many checkpoints as segments."""
    from qompsim.adjoint import CheckpointedAdjoint

    n = 12
    c = W.hea(n, layers=3, seed=6)
    obs = W.tfim_observable(n)
    adj = CheckpointedAdjoint(c, 40, block_size=8)
    for seg in (40, 40, 9, 83):
        adj.segment_gates = seg
        th = c.get_parameters()
        for i in range(len(th)):
            c.set_parameter(i, th[i] + rng.uniform(-0.3, 0.3))
        adj.forward_with_checkpoints()
        e, g = adj.adjoint(obs)
        e0, g0 = O.backprop(c, obs)
        assert abs(e - e0) < 1e-11
        assert np.max(np.abs(g - g0)) < 1e-10
        assert adj.checkpoint_info()[0] == -(-len(c.gate_list) // seg)


def test_checkpointed_adjoint_lossy_codec_bound(rng):
    """Checkpoints with an absolute bound of 1e-7 per double: each segment
    restarts from a state within that bound, so the gradient error stays of
    the bound's order (here well under 1e-4) while the plain sweep is exact."""
    from qompsim.adjoint import CheckpointedAdjoint, checkpoint_codec

    n = 12
    c = _mixed_circuit(n, rng, layers=3)
    obs = _mixed_observable(n, rng)
    adj = CheckpointedAdjoint(c, 20, block_size=8, codec=checkpoint_codec(eps_rel=0.0, eps_abs=1e-7))
    adj.forward_with_checkpoints()
    e, g = adj.adjoint(obs)
    e0, g0 = O.backprop(c, obs)
    assert abs(e - e0) < 1e-11  # lambda comes from the exact forward state
    err = np.max(np.abs(g - g0))
    assert err < 1e-4


def test_torch_checkpointed_expectation():
    import torch

    from qompsim.torch_adjoint import expectation

    n = 10
    c = W.hea(n, layers=4, seed=6)
    obs = W.tfim_observable(n)
    th = torch.tensor(c.get_parameters(), dtype=torch.float64, requires_grad=True)
    E = expectation(th, c, obs, segment_gates=2 * (3 * n - 1))
    (2.0 * E).backward()
    e0, g0 = O.backprop(c, obs)
    assert abs(E.item() - e0) < 1e-11
    assert np.max(np.abs(th.grad.numpy() - 2.0 * g0)) < 1e-10


def test_checkpointed_adjoint_requires_forward():
    from qompsim import _lib
    from qompsim.adjoint import CheckpointedAdjoint

    n = 8
    c = W.hea(n, layers=1, seed=1)
    adj = CheckpointedAdjoint(c, 5)
    with pytest.raises(ValueError, match="forward"):
        adj.adjoint(W.tfim_observable(n))
