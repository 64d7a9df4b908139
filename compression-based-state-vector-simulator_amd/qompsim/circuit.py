"""QuantumCircuit, QuantumCircuitOptimizer and ParametricQuantumCircuit.

QuantumCircuit.update_quantum_state (src/cppsim/circuit.cpp:23-34) hands the
whole gate list to libqsim_hip.so, which fuses it into LDS-tiled passes over
the HBM-resident state.  The optimizer is the C++ greedy fusion of
src/cppsim/circuit_optimizer.cpp:113-168 inside the library.
"""
import ctypes

import numpy as np

from . import _lib
from . import gate as _g
from ._lib import call, lib, ptr


class QuantumCircuit:
    def __init__(self, qubit_count):
        self._n = int(qubit_count)
        self.gate_list = []

    def get_qubit_count(self):
        return self._n

    def get_gate_count(self):
        return len(self.gate_list)

    def get_gate(self, index):
        return self.gate_list[index]

    def add_gate(self, gate, position=None):
        qs = gate.get_target_index_list() + gate.get_control_index_list()
        if any(q >= self._n or q < 0 for q in qs):
            raise ValueError(f"gate {gate.get_name()} acts outside the {self._n}-qubit register")
        if position is None:
            self.gate_list.append(gate)
        else:
            self.gate_list.insert(position, gate)

    def remove_gate(self, index):
        del self.gate_list[index]

    def copy(self):
        c = type(self).__new__(type(self))
        c.__dict__.update(self.__dict__)
        c.gate_list = [g.copy() for g in self.gate_list]
        return c

    def update_quantum_state(self, state, start=0, end=None):
        """Apply gates [start, end) to a QuantumState (src/cppsim/circuit.cpp:23-34)."""
        if state.get_qubit_count() != self._n:
            raise ValueError("InvalidQubitCountException: circuit and state qubit counts differ")
        gates = self.gate_list[start:end]
        if not gates:
            return
        rec, pool = _lib.pack_gates(gates)
        call("qsim_sv_apply", state._h, ptr(rec), len(gates), ptr(pool), pool.size)

    # ---- add_* helpers (python/cppsim_wrapper.cpp:1287-1600)
    def add_X_gate(self, i): self.add_gate(_g.X(i))
    def add_Y_gate(self, i): self.add_gate(_g.Y(i))
    def add_Z_gate(self, i): self.add_gate(_g.Z(i))
    def add_H_gate(self, i): self.add_gate(_g.H(i))
    def add_S_gate(self, i): self.add_gate(_g.S(i))
    def add_Sdag_gate(self, i): self.add_gate(_g.Sdag(i))
    def add_T_gate(self, i): self.add_gate(_g.T(i))
    def add_Tdag_gate(self, i): self.add_gate(_g.Tdag(i))
    def add_sqrtX_gate(self, i): self.add_gate(_g.sqrtX(i))
    def add_sqrtXdag_gate(self, i): self.add_gate(_g.sqrtXdag(i))
    def add_sqrtY_gate(self, i): self.add_gate(_g.sqrtY(i))
    def add_sqrtYdag_gate(self, i): self.add_gate(_g.sqrtYdag(i))
    def add_P0_gate(self, i): self.add_gate(_g.P0(i))
    def add_P1_gate(self, i): self.add_gate(_g.P1(i))
    def add_CNOT_gate(self, c, t): self.add_gate(_g.CNOT(c, t))
    def add_CZ_gate(self, c, t): self.add_gate(_g.CZ(c, t))
    def add_SWAP_gate(self, a, b): self.add_gate(_g.SWAP(a, b))
    def add_RX_gate(self, i, angle): self.add_gate(_g.RX(i, angle))
    def add_RY_gate(self, i, angle): self.add_gate(_g.RY(i, angle))
    def add_RZ_gate(self, i, angle): self.add_gate(_g.RZ(i, angle))
    def add_RotX_gate(self, i, angle): self.add_gate(_g.RotX(i, angle))
    def add_RotY_gate(self, i, angle): self.add_gate(_g.RotY(i, angle))
    def add_RotZ_gate(self, i, angle): self.add_gate(_g.RotZ(i, angle))
    def add_RotInvX_gate(self, i, angle): self.add_gate(_g.RotInvX(i, angle))
    def add_RotInvY_gate(self, i, angle): self.add_gate(_g.RotInvY(i, angle))
    def add_RotInvZ_gate(self, i, angle): self.add_gate(_g.RotInvZ(i, angle))
    def add_U1_gate(self, i, lam): self.add_gate(_g.U1(i, lam))
    def add_U2_gate(self, i, phi, lam): self.add_gate(_g.U2(i, phi, lam))
    def add_U3_gate(self, i, theta, phi, lam): self.add_gate(_g.U3(i, theta, phi, lam))

    def add_dense_matrix_gate(self, targets, matrix):
        self.add_gate(_g.DenseMatrix(targets, matrix))

    def add_diagonal_observable_rotation_gate(self, *a):  # pragma: no cover - outside the hot path
        raise NotImplementedError

    def add_multi_Pauli_gate(self, index_list, pauli_ids):
        self.add_gate(_g.Pauli(index_list, pauli_ids))

    def add_multi_Pauli_rotation_gate(self, index_list, pauli_ids, angle):
        self.add_gate(_g.PauliRotation(index_list, pauli_ids, angle))

    def __repr__(self):
        return f"QuantumCircuit(qubits={self._n}, gates={len(self.gate_list)})"


class QuantumCircuitOptimizer:
    """Greedy gate fusion (src/cppsim/circuit_optimizer.cpp:113-168)."""

    def __init__(self, mpi_size=0, sv_block_size=0):
        self._mpi_size = mpi_size
        self._sv_block_size = sv_block_size

    def optimize(self, circuit, block_size, swap_level=0):
        if swap_level != 0:
            raise NotImplementedError("swap insertion is a multi-rank feature; one GPU holds every block")
        if self._sv_block_size and self._sv_block_size > circuit.get_qubit_count():
            raise RuntimeError("sv_block_size exceeds the local qubit count")
        gates = list(circuit.gate_list)
        if not gates:
            return
        rec, pool = _lib.pack_gates(gates)
        out = np.zeros(len(gates), dtype=_lib.GATE_DTYPE)
        cap = lib.qsim_optimize_pool_bound(len(gates), pool.size, int(block_size))
        pool_out = np.zeros(cap, dtype=np.float64)
        n_out = ctypes.c_uint64()
        plen = ctypes.c_uint64(cap)
        origin = np.zeros(len(gates), dtype=np.int64)
        call("qsim_optimize", ptr(rec), len(gates), ptr(pool), pool.size, int(block_size), ptr(out),
             ctypes.byref(n_out), ptr(pool_out), ctypes.byref(plen), ptr(origin))
        new = []
        mpool = pool_out.view(np.complex128)
        for i in range(n_out.value):
            if origin[i] >= 0:
                new.append(gates[origin[i]])
                continue
            r = out[i]
            k, nc = int(r["n_targets"]), int(r["n_controls"])
            d = 1 << k
            off = int(r["matrix_offset"])
            m = mpool[off:off + d * d].reshape(d, d).copy()
            new.append(_g.QuantumGate("DenseMatrix", r["targets"][:k].tolist(), m, r["controls"][:nc].tolist(),
                                      r["control_values"][:nc].tolist(), r["target_commute"][:k].tolist()))
        circuit.gate_list = new

    def merge_all(self, circuit):
        return _g.merge(circuit.gate_list)


class ParametricQuantumCircuit(QuantumCircuit):
    """QuantumCircuit with parametric rotations (src/vqcsim/parametric_circuit.cpp)."""

    def __init__(self, qubit_count):
        super().__init__(qubit_count)
        self._param_gates = []

    def add_parametric_gate(self, gate, position=None):
        if not gate.is_parametric():
            raise ValueError("not a parametric gate")
        gate._param_index = len(self._param_gates)
        self._param_gates.append(gate)
        self.add_gate(gate, position)

    def add_parametric_RX_gate(self, i, angle): self.add_parametric_gate(_g.ParametricRX(i, angle))
    def add_parametric_RY_gate(self, i, angle): self.add_parametric_gate(_g.ParametricRY(i, angle))
    def add_parametric_RZ_gate(self, i, angle): self.add_parametric_gate(_g.ParametricRZ(i, angle))

    def add_parametric_multi_Pauli_rotation_gate(self, index_list, pauli_ids, angle):
        self.add_parametric_gate(_g.ParametricPauliRotation(index_list, pauli_ids, angle))

    def get_parameter_count(self):
        return len(self._param_gates)

    def get_parameter(self, index):
        return self._param_gates[index].get_parameter_value()

    def set_parameter(self, index, value):
        self._param_gates[index].set_parameter_value(value)

    def get_parameters(self):
        return np.array([g.get_parameter_value() for g in self._param_gates])

    def set_parameters(self, values):
        for g, v in zip(self._param_gates, values):
            g.set_parameter_value(float(v))

    def get_parametric_gate_position(self, index):
        return self.gate_list.index(self._param_gates[index])

    def copy(self):
        c = ParametricQuantumCircuit(self._n)
        for g in self.gate_list:
            if g.is_parametric():
                c.add_parametric_gate(g.copy())
            else:
                c.add_gate(g.copy())
        return c

    def backprop(self, obs):
        """Adjoint gradient d<H>/dtheta from |0...0> (src/vqcsim/parametric_circuit.cpp:281-308)."""
        from .adjoint import adjoint_gradient

        return list(adjoint_gradient(self, obs)[1])
