"""Adjoint (reverse-sweep) gradients of <H> for parametric circuits.

Semantics of ParametricQuantumCircuit::backprop
(src/vqcsim/parametric_circuit.cpp:180-308): psi = U|0>, lambda = 2 H psi,
then for gates from last to first, g_i = Re<lambda| R_P(pi) psi>/2 for a
parametric rotation (dR/dtheta = R(theta+pi)/2) and both states are uncomputed
with G_i^dagger.  Every state update runs on the GPU.
"""
import numpy as np

from . import gate as _g
from .circuit import QuantumCircuit
from .state import QuantumState, inner_product


def adjoint_gradient(circuit, obs, state=None):
    n = circuit.get_qubit_count()
    psi = QuantumState(n)
    if state is not None:
        psi.load(state)
    circuit.update_quantum_state(psi)
    energy = obs.get_expectation_value(psi)
    lam = QuantumState(n)
    obs.apply_to_state(None, psi, lam)
    lam.multiply_coef(2.0)
    grads = np.zeros(circuit.get_parameter_count())
    tmp = QuantumState(n)
    one = QuantumCircuit(n)
    for gate in reversed(circuit.gate_list):
        if gate.is_parametric():
            tmp.load(psi)
            rpi = _g.PauliRotation(gate.get_target_index_list(), gate.get_pauli_id_list(), np.pi)
            one.gate_list = [rpi]
            one.update_quantum_state(tmp)
            grads[gate._param_index] = inner_product(lam, tmp).real / 2.0
        one.gate_list = [gate.get_inverse()]
        one.update_quantum_state(lam)
        one.update_quantum_state(psi)
    return float(np.real(energy)), grads
