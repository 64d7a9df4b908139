"""Adjoint (reverse-sweep) gradients of <H> for parametric circuits.

Semantics of ParametricQuantumCircuit::backprop
(src/vqcsim/parametric_circuit.cpp:180-308): psi = U|psi_0>, lambda = 2 H psi,
then for gates from last to first, g_i = Re<lambda| R_P(pi) psi>/2 for a
parametric rotation R_P(theta) = exp(+i theta P/2) (dR/dtheta = R(theta+pi)/2)
and both states are uncomputed with G_i^dagger.  The whole computation runs in
libqsim_hip.so (qsim_sv_backprop): the forward passes, lambda = 2 H psi in one
sweep, and the reverse sweep in fused passes over psi and lambda together.

CheckpointedAdjoint splits the same computation into forward_with_checkpoints
and adjoint calls (qsim_adj_forward / qsim_adj_backward): the forward stores
the state at every segment start in a compressed store, and the reverse sweep
restarts each segment from its checkpoint.
"""
import ctypes

import numpy as np

from ._lib import CsConfig, call, pack_gates, ptr
from .state import QuantumState


def adjoint_gradient(circuit, obs, state=None, register=None):
    """(<H>, dE/dtheta) for `circuit` applied to `state` (|0...0> when None).
    `register`, a QuantumState, is used as the working state when given and
    holds U|state> afterwards."""
    n = circuit.get_qubit_count()
    psi = register if register is not None else QuantumState(n)
    if state is not None:
        psi.load(state)
    elif register is not None:
        psi.set_zero_state()
    npar = circuit.get_parameter_count() if hasattr(circuit, "get_parameter_count") else 0
    grads = np.zeros(max(npar, 1), dtype=np.float64)
    rec, pool = pack_gates(circuit.gate_list)
    xm, zm, ny, cr, ci = obs._arrays()
    energy = ctypes.c_double()
    call("qsim_sv_backprop", psi._h, ptr(rec), len(circuit.gate_list), ptr(pool), pool.size, len(obs.terms),
         ptr(xm), ptr(zm), ptr(ny), ptr(cr), ptr(ci), npar, ctypes.byref(energy), ptr(grads))
    return float(energy.value), grads[:npar]


def checkpoint_codec(eps_rel=0.0, eps_abs=1e-13, guard_bits=1):
    """Codec settings for adjoint checkpoints.  The default bound (absolute
    1e-13 per double) keeps the gradient error near the FP64 level; looser
    bounds trade gradient accuracy for checkpoint size."""
    cfg = CsConfig()
    call("qsim_cs_default_config", ctypes.byref(cfg))
    cfg.eps_rel, cfg.eps_abs, cfg.guard_bits = float(eps_rel), float(eps_abs), int(guard_bits)
    return cfg


class CheckpointedAdjoint:
    """forward_with_checkpoints(k) and adjoint(obs) over one GPU state.

    `segment_gates` gates form one segment; `block_size` is the block size of
    the compressed checkpoint stores."""

    def __init__(self, circuit, segment_gates, block_size=20, codec=None):
        self.circuit = circuit
        self.n = circuit.get_qubit_count()
        self.segment_gates = int(segment_gates)
        self.block_size = int(block_size)
        self.codec = codec if codec is not None else checkpoint_codec()
        h = ctypes.c_void_p()
        call("qsim_adj_create", ctypes.byref(h))
        self._h = h
        self.state = QuantumState(self.n)

    def __del__(self):
        if getattr(self, "_h", None):
            try:
                call("qsim_adj_destroy", self._h)
            except Exception:
                pass
            self._h = None

    def forward_with_checkpoints(self, state=None):
        """Apply the circuit to `state` (|0...0> when None), keeping checkpoints."""
        if state is None:
            self.state.set_zero_state()
        else:
            self.state.load(state)
        rec, pool = pack_gates(self.circuit.gate_list)
        call("qsim_adj_forward", self._h, self.state._h, ptr(rec), len(self.circuit.gate_list), ptr(pool), pool.size,
             self.segment_gates, self.block_size, ctypes.byref(self.codec))
        return self.state

    def adjoint(self, obs):
        """(<H>, dE/dtheta) from the stored forward run."""
        npar = self.circuit.get_parameter_count() if hasattr(self.circuit, "get_parameter_count") else 0
        grads = np.zeros(max(npar, 1), dtype=np.float64)
        xm, zm, ny, cr, ci = obs._arrays()
        energy = ctypes.c_double()
        call("qsim_adj_backward", self._h, self.state._h, len(obs.terms), ptr(xm), ptr(zm), ptr(ny), ptr(cr), ptr(ci),
             npar, ctypes.byref(energy), ptr(grads))
        return float(energy.value), grads[:npar]

    def checkpoint_info(self):
        """(number of checkpoints, compressed bytes held)."""
        k, b = ctypes.c_uint64(), ctypes.c_uint64()
        call("qsim_adj_checkpoint_bytes", self._h, ctypes.byref(k), ctypes.byref(b))
        return int(k.value), int(b.value)
