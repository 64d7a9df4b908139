"""GPU parity at the shapes of the measured workloads, scaled to sizes the
oracle finishes in seconds:

- the transverse-field Ising Trotter evolution (X on qubit 1, then per step
  RZZ(0.2) on the chain and RX(0.3) on every qubit, one circuit per step) on
  both payload stores against the oracle's stage loop;
- QFT on a product state with 2^20-amplitude blocks (the block size of the
  32-qubit run) against the oracle's stage loop and the unitary DFT;
- the 8-layer hardware-efficient ansatz gradient against the oracle's
  backprop, through torch.autograd and with compressed checkpoints every two
  layers;
- the launch-splitting and persistent-grid paths of the kernels, forced at
  small sizes, bit-identical to the default launches.
"""
import numpy as np
import pytest

import pyoracle as O
from qompsim import CompressedQuantumState, QuantumCircuit, QuantumCircuitOptimizer, QuantumState
from qompsim import workloads as W
from qompsim.observable import Observable

pytestmark = pytest.mark.gpu


def _tfim_step(n, first, jzz=0.2, hx=0.3):
    c = QuantumCircuit(n)
    if first:
        c.add_X_gate(1)
    for i in range(n - 1):
        c.add_multi_Pauli_rotation_gate([i, i + 1], [3, 3], jzz)
    for q in range(n):
        c.add_RX_gate(q, hx)
    QuantumCircuitOptimizer().optimize(c, 2)
    return c


@pytest.mark.parametrize("store", ["double", "regions"])
def test_tfim_trotter_vs_oracle(store):
    """TFIM Trotter evolution, 12 steps, B=12, I=4, eps_rel 1e-3: the null
    bitmap after every step and the bytes history bit-exact, amplitudes
    within 1e-6 of the oracle's stage loop, <Z_0> and the norm against the
    oracle's decoded state and the exact evolution."""
    n, B, I, steps = 20, 12, 4, 12
    cs = CompressedQuantumState(n, B, eps_rel=1e-3, eps_abs=1e-7, store=store, verbose=False)
    oc = O.OracleCompressedState(n, B, 1e-3, 1e-7, 1)
    exact = O.zero_state(n)
    for s in range(steps):
        c = _tfim_step(n, s == 0)
        cs.partition_and_register_circuit(c, I)
        oc.partition_and_register_circuit(c, I)
        cs.apply_circuit()
        oc.apply_circuit()
        exact = O.apply_circuit(exact, c.gate_list)
        assert np.array_equal(cs.get_null_bitmap(), oc.null_bitmap()), f"bitmap after step {s + 1}"
    assert cs.get_bytes_history() == oc.bytes_history()
    raw = cs.get_raw_vector()
    oraw = oc.get_vector(normalise=False)
    assert np.max(np.abs(raw - oraw)) < 1e-6
    # compensated GPU norm of the store against the oracle's decoded vector
    assert abs(cs.get_squared_norm() - O.norm(oraw)) < 1e-9
    z0 = Observable(n)
    z0.add_operator(1.0, "Z 0")
    zc = cs.get_expectation_value(z0) / cs.get_squared_norm()
    zo = O.expectation(oraw, z0).real / O.norm(oraw)
    ze = O.expectation(exact, z0).real
    assert abs(zc - zo) < 1e-6
    # the codec's error accumulated over 12 steps (measured about 1e-4)
    assert abs(zc - ze) < 2e-3
    assert abs(cs.get_squared_norm() - 1.0) < 1e-2
    assert abs(np.vdot(exact, cs.get_vector())) > 0.999


def _product_state(n, seed):
    layer, _, _ = W.product_layer(n, seed)
    psi = np.ones(1, dtype=np.complex128)
    for q in range(n):
        v = np.array([1.0, 0.0], dtype=np.complex128)
        for g in layer[2 * q: 2 * q + 2]:
            v = np.asarray(g.get_matrix()) @ v
        psi = np.kron(v, psi)  # qubit q is bit q of the index
    return psi


def test_qft_block20_vs_oracle():
    """QFT-25 on the seeded product state with 2^20-amplitude blocks (16 MiB
    per block, 32 blocks, 5 stages at I=4): stages, bitmap and bytes history
    equal to the oracle's, amplitudes within 1e-6 of its stage loop, and
    fidelity against the unitary DFT of the input (numpy FFT) >= 0.9999."""
    n, B, I = 25, 20, 4
    c = W.qft_on_product(n, seed=2)
    QuantumCircuitOptimizer().optimize(c, 2)
    cs = CompressedQuantumState(n, B, eps_rel=1e-3, eps_abs=1e-7, verbose=False)
    cs.partition_and_register_circuit(c, I)
    oc = O.OracleCompressedState(n, B, 1e-3, 1e-7, 1)
    oc.partition_and_register_circuit(c, I)
    assert cs.get_stages() == [s[0] for s in oc.stages()]
    assert cs.get_num_stages() > 1
    cs.apply_circuit()
    oc.apply_circuit()
    assert np.array_equal(cs.get_null_bitmap(), oc.null_bitmap())
    assert cs.get_bytes_history() == oc.bytes_history()
    v = cs.get_vector()
    assert np.max(np.abs(v - oc.get_vector())) < 1e-6
    exact = np.fft.ifft(_product_state(n, 2), norm="ortho")
    fid = abs(np.vdot(exact, v))
    assert fid >= 0.9999, fid
    s = QuantumState(n)
    s.load(exact)
    assert abs(cs.fidelity_with(s) - fid) < 1e-10


def test_hea_8_layers_gradient_vs_oracle():
    """HEA-20, 8 layers (320 parameters), TFIM energy: the adjoint gradient
    through torch.autograd within 1e-10 of the oracle's backprop, and the
    checkpointed sweep (compressed checkpoints every two layers) within 1e-9."""
    import torch

    from qompsim.torch_adjoint import expectation

    n, layers = 20, 8
    c = W.hea(n, layers=layers, seed=3)
    obs = W.tfim_observable(n)
    e0, g0 = O.backprop(c, obs)
    theta = torch.tensor(c.get_parameters(), dtype=torch.float64, requires_grad=True)
    e = expectation(theta, c, obs)
    e.backward()
    assert abs(float(e) - e0) < 1e-10
    assert np.max(np.abs(theta.grad.numpy() - g0)) < 1e-10
    theta2 = torch.tensor(c.get_parameters(), dtype=torch.float64, requires_grad=True)
    seg = 2 * (3 * n - 1)  # two layers: 2n rotations and n - 1 CZ per layer
    e2 = expectation(theta2, c, obs, segment_gates=seg, block_size=16)
    e2.backward()
    assert abs(float(e2) - e0) < 1e-9
    assert np.max(np.abs(theta2.grad.numpy() - g0)) < 1e-9


def _run_all_paths(n, B):
    """A compressed QFT (single-pass DEC->ENC stages, multi-pass stages
    through the staging buffer) and the uncompressed sweep of the same
    circuit: raw bits of every result."""
    c = W.qft_on_product(n, seed=7)
    QuantumCircuitOptimizer().optimize(c, 2)
    cs = CompressedQuantumState(n, B, eps_rel=1e-3, eps_abs=1e-7, scratch_bytes=1 << 22, verbose=False)
    cs.partition_and_register_circuit(c, 4)
    cs.apply_circuit()
    s = QuantumState(n)
    c.update_quantum_state(s)
    return (cs.get_raw_vector(), cs.get_null_bitmap(), cs.get_bytes_history(), s.get_vector(),
            max(cs.get_stage_passes()))


def test_launch_split_and_persistent_stride_are_exact(monkeypatch):
    """Launches split into many dispatches (as 32-qubit stages are, at 2^23
    work-groups per dispatch) and persistent grids much smaller than the
    tile count (every wave walking many tiles) give the bits of the default
    launches."""
    n, B = 20, 12
    ref = _run_all_paths(n, B)
    assert ref[4] > 1  # a multi-pass stage ran through the staging buffer
    monkeypatch.setenv("QOMPSIM_MAX_LAUNCH_GROUPS", "3")
    monkeypatch.setenv("QOMPSIM_PERSISTENT_GROUPS", "5")
    got = _run_all_paths(n, B)
    assert np.array_equal(got[0].view(np.uint64), ref[0].view(np.uint64))
    assert np.array_equal(got[1], ref[1])
    assert got[2] == ref[2]
    assert np.array_equal(got[3].view(np.uint64), ref[3].view(np.uint64))
