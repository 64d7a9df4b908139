"""Adjoint (reverse-sweep) gradients of <H> for parametric circuits.

Semantics of ParametricQuantumCircuit::backprop
(src/vqcsim/parametric_circuit.cpp:180-308): psi = U|psi_0>, lambda = 2 H psi,
then for gates from last to first, g_i = Re<lambda| R_P(pi) psi>/2 for a
parametric rotation R_P(theta) = exp(+i theta P/2) (dR/dtheta = R(theta+pi)/2)
and both states are uncomputed with G_i^dagger.  The whole computation runs in
libqsim_hip.so (qsim_sv_backprop): the forward passes, lambda = 2 H psi in one
sweep, and the reverse sweep in fused passes over psi and lambda together.
"""
import ctypes

import numpy as np

from ._lib import call, pack_gates, ptr
from .state import QuantumState


def adjoint_gradient(circuit, obs, state=None):
    """(<H>, dE/dtheta) for `circuit` applied to `state` (|0...0> when None)."""
    n = circuit.get_qubit_count()
    psi = QuantumState(n)
    if state is not None:
        psi.load(state)
    npar = circuit.get_parameter_count() if hasattr(circuit, "get_parameter_count") else 0
    grads = np.zeros(max(npar, 1), dtype=np.float64)
    rec, pool = pack_gates(circuit.gate_list)
    xm, zm, ny, cr, ci = obs._arrays()
    energy = ctypes.c_double()
    call("qsim_sv_backprop", psi._h, ptr(rec), len(circuit.gate_list), ptr(pool), pool.size, len(obs.terms),
         ptr(xm), ptr(zm), ptr(ny), ptr(cr), ptr(ci), npar, ctypes.byref(energy), ptr(grads))
    return float(energy.value), grads[:npar]
