"""Gate factory with Qulacs semantics (``qulacs.gate``).

Every gate is held in QuantumGateMatrix form (src/cppsim/gate_matrix.cpp:
19-250): targets (matrix index bit i <-> targets[i]), controls with values, a
2^k x 2^k complex matrix and per-target commutation flags that the circuit
optimizer reads (src/cppsim/gate_named_one.hpp:95-464,
src/cppsim/gate_named_two.hpp:83-224, src/cppsim/gate_named_pauli.hpp).
Rotations follow R_P(theta) = exp(+i theta P / 2)
(src/cppsim/gate_named_one.hpp:412-464); RotX/RotY/RotZ = R(-theta)
(src/cppsim/gate_factory.cpp:152-165).
"""
import numpy as np

from ._lib import COMMUTE_X, COMMUTE_Y, COMMUTE_Z

_I2 = np.eye(2, dtype=np.complex128)
_X = np.array([[0, 1], [1, 0]], dtype=np.complex128)
_Y = np.array([[0, -1j], [1j, 0]], dtype=np.complex128)
_Z = np.array([[1, 0], [0, -1]], dtype=np.complex128)
_PAULI = {0: _I2, 1: _X, 2: _Y, 3: _Z}
_PAULI_FLAG = {0: COMMUTE_X | COMMUTE_Y | COMMUTE_Z, 1: COMMUTE_X, 2: COMMUTE_Y, 3: COMMUTE_Z}


class QuantumGate:
    """A gate in QuantumGateMatrix form."""

    def __init__(self, name, targets, matrix, controls=(), control_values=None, commute=None):
        self._name = name
        self._targets = [int(t) for t in targets]
        self._controls = [int(c) for c in controls]
        self._cvals = [1] * len(self._controls) if control_values is None else [int(v) for v in control_values]
        self._matrix = np.array(matrix, dtype=np.complex128)
        self._commute = list(commute) if commute is not None else [0] * len(self._targets)
        qs = self._targets + self._controls
        if len(set(qs)) != len(qs):
            raise ValueError(f"{name}: duplicated qubit index")

    # --- Qulacs accessors
    def get_name(self):
        return self._name

    def set_name(self, name):
        self._name = name

    def get_target_index_list(self):
        return list(self._targets)

    def get_control_index_list(self):
        return list(self._controls)

    def get_control_value_list(self):
        return list(self._cvals)

    def get_matrix(self):
        return self._matrix.copy()

    def is_parametric(self):
        return False

    def is_diagonal(self):
        return all(c & COMMUTE_Z for c in self._commute)

    def add_control_qubit(self, index, control_value):
        """QuantumGateMatrix::add_control_qubit (gate_matrix.cpp)."""
        if index in self._targets or index in self._controls:
            raise ValueError("control qubit overlaps the gate")
        self._controls.append(int(index))
        self._cvals.append(int(control_value))

    def copy(self):
        g = QuantumGate(self._name, self._targets, self._matrix, self._controls, self._cvals, self._commute)
        return g

    def get_inverse(self):
        return QuantumGate(self._name, self._targets, self._matrix.conj().T, self._controls, self._cvals,
                           self._commute)

    def update_quantum_state(self, state):
        from .circuit import QuantumCircuit

        c = QuantumCircuit(state.get_qubit_count())
        c.add_gate(self)
        c.update_quantum_state(state)

    def __repr__(self):
        return f"{self._name}(targets={self._targets}, controls={self._controls})"


class ParametricGate(QuantumGate):
    """Parametric Pauli rotation exp(+i theta P / 2) (src/vqcsim/parametric_gate.hpp:102-200)."""

    def __init__(self, name, targets, pauli_ids, angle):
        self._pauli_ids = [int(p) for p in pauli_ids]
        self._angle = float(angle)
        self._param_index = -1
        super().__init__(name, targets, _rotation_matrix(self._pauli_ids, self._angle),
                         commute=[_PAULI_FLAG[p] for p in self._pauli_ids])

    def is_parametric(self):
        return True

    def get_parameter_value(self):
        return self._angle

    def set_parameter_value(self, angle):
        self._angle = float(angle)
        self._matrix = _rotation_matrix(self._pauli_ids, self._angle)

    def get_pauli_id_list(self):
        return list(self._pauli_ids)

    def copy(self):
        g = ParametricGate(self._name, self._targets, self._pauli_ids, self._angle)
        g._param_index = self._param_index
        return g

    def get_inverse(self):
        g = ParametricGate(self._name, self._targets, self._pauli_ids, -self._angle)
        return g


def _pauli_matrix(pauli_ids):
    m = np.ones((1, 1), dtype=np.complex128)
    for p in pauli_ids:  # targets[0] is the least significant matrix bit
        m = np.kron(_PAULI[p], m)
    return m


def _rotation_matrix(pauli_ids, angle):
    P = _pauli_matrix(pauli_ids)
    d = P.shape[0]
    return np.cos(angle / 2) * np.eye(d, dtype=np.complex128) + 1j * np.sin(angle / 2) * P


# ---------------------------------------------------------------- factory
def Identity(index):
    return QuantumGate("I", [index], _I2, commute=[COMMUTE_X | COMMUTE_Y | COMMUTE_Z])


def X(index):
    return QuantumGate("X", [index], _X, commute=[COMMUTE_X])


def Y(index):
    return QuantumGate("Y", [index], _Y, commute=[COMMUTE_Y])


def Z(index):
    return QuantumGate("Z", [index], _Z, commute=[COMMUTE_Z])


def H(index):
    return QuantumGate("H", [index], np.array([[1, 1], [1, -1]], dtype=np.complex128) / np.sqrt(2))


def S(index):
    return QuantumGate("S", [index], np.diag([1, 1j]), commute=[COMMUTE_Z])


def Sdag(index):
    return QuantumGate("Sdag", [index], np.diag([1, -1j]), commute=[COMMUTE_Z])


def T(index):
    return QuantumGate("T", [index], np.diag([1, np.exp(1j * np.pi / 4)]), commute=[COMMUTE_Z])


def Tdag(index):
    return QuantumGate("Tdag", [index], np.diag([1, np.exp(-1j * np.pi / 4)]), commute=[COMMUTE_Z])


def sqrtX(index):
    m = np.array([[0.5 + 0.5j, 0.5 - 0.5j], [0.5 - 0.5j, 0.5 + 0.5j]])
    return QuantumGate("sqrtX", [index], m, commute=[COMMUTE_X])


def sqrtXdag(index):
    m = np.array([[0.5 - 0.5j, 0.5 + 0.5j], [0.5 + 0.5j, 0.5 - 0.5j]])
    return QuantumGate("sqrtXdag", [index], m, commute=[COMMUTE_X])


def sqrtY(index):
    m = np.array([[0.5 + 0.5j, -0.5 - 0.5j], [0.5 + 0.5j, 0.5 + 0.5j]])
    return QuantumGate("sqrtY", [index], m, commute=[COMMUTE_Y])


def sqrtYdag(index):
    m = np.array([[0.5 - 0.5j, 0.5 - 0.5j], [-0.5 + 0.5j, 0.5 - 0.5j]])
    return QuantumGate("sqrtYdag", [index], m, commute=[COMMUTE_Y])


def P0(index):
    return QuantumGate("Projection-0", [index], np.diag([1, 0]))


def P1(index):
    return QuantumGate("Projection-1", [index], np.diag([0, 1]))


def _ibmq(theta, phi, lam):
    """get_IBMQ_matrix (src/cppsim/gate_factory.cpp:166-180)."""
    c, s = np.cos(theta / 2), np.sin(theta / 2)
    return np.array([[c, -np.exp(1j * lam) * s], [np.exp(1j * phi) * s, np.exp(1j * (phi + lam)) * c]])


def U1(index, lambda_):
    return QuantumGate("DenseMatrix", [index], _ibmq(0.0, 0.0, lambda_))


def U2(index, phi, lambda_):
    return QuantumGate("DenseMatrix", [index], _ibmq(np.pi / 2, phi, lambda_))


def U3(index, theta, phi, lambda_):
    return QuantumGate("DenseMatrix", [index], _ibmq(theta, phi, lambda_))


def RX(index, angle):
    return QuantumGate("X-rotation", [index], _rotation_matrix([1], angle), commute=[COMMUTE_X])


def RY(index, angle):
    return QuantumGate("Y-rotation", [index], _rotation_matrix([2], angle), commute=[COMMUTE_Y])


def RZ(index, angle):
    return QuantumGate("Z-rotation", [index], _rotation_matrix([3], angle), commute=[COMMUTE_Z])


def RotInvX(index, angle):
    return RX(index, angle)


def RotInvY(index, angle):
    return RY(index, angle)


def RotInvZ(index, angle):
    return RZ(index, angle)


def RotX(index, angle):
    return RX(index, -angle)


def RotY(index, angle):
    return RY(index, -angle)


def RotZ(index, angle):
    return RZ(index, -angle)


def CNOT(control, target):
    if control == target:
        raise ValueError("CNOT: control and target are the same qubit")
    return QuantumGate("CNOT", [target], _X, controls=[control], commute=[COMMUTE_X])


def CZ(control, target):
    if control == target:
        raise ValueError("CZ: control and target are the same qubit")
    return QuantumGate("CZ", [target], _Z, controls=[control], commute=[COMMUTE_Z])


def SWAP(index1, index2):
    if index1 == index2:
        raise ValueError("SWAP: two indices have the same value")
    m = np.eye(4, dtype=np.complex128)[[0, 2, 1, 3]]
    return QuantumGate("SWAP", [index1, index2], m)


def TOFFOLI(control1, control2, target):
    return QuantumGate("DenseMatrix", [target], _X, controls=[control1, control2], commute=[COMMUTE_X])


def FREDKIN(control, target1, target2):
    g = SWAP(target1, target2)
    g._name = "DenseMatrix"
    g.add_control_qubit(control, 1)
    return g


def DenseMatrix(targets, matrix):
    if np.isscalar(targets):
        targets = [targets]
    return QuantumGate("DenseMatrix", list(targets), matrix)


def DiagonalMatrix(targets, diagonal_element):
    if np.isscalar(targets):
        targets = [targets]
    d = np.asarray(diagonal_element, dtype=np.complex128)
    return QuantumGate("DiagonalMatrix", list(targets), np.diag(d), commute=[COMMUTE_Z] * len(targets))


def Pauli(index_list, pauli_ids):
    return QuantumGate("Pauli", list(index_list), _pauli_matrix(pauli_ids),
                       commute=[_PAULI_FLAG[p] for p in pauli_ids])


def PauliRotation(index_list, pauli_ids, angle):
    return QuantumGate("Pauli-rotation", list(index_list), _rotation_matrix(list(pauli_ids), angle),
                       commute=[_PAULI_FLAG[p] for p in pauli_ids])


def ParametricRX(index, angle=0.0):
    return ParametricGate("ParametricRX", [index], [1], angle)


def ParametricRY(index, angle=0.0):
    return ParametricGate("ParametricRY", [index], [2], angle)


def ParametricRZ(index, angle=0.0):
    return ParametricGate("ParametricRZ", [index], [3], angle)


def ParametricPauliRotation(index_list, pauli_ids, angle=0.0):
    return ParametricGate("ParametricPauliRotation", list(index_list), list(pauli_ids), angle)


def to_matrix_gate(gate):
    """gate::to_matrix_gate (src/cppsim/gate_merge.cpp:428-433)."""
    g = QuantumGate(gate.get_name(), gate.get_target_index_list(), gate.get_matrix(),
                    gate.get_control_index_list(), gate.get_control_value_list(), gate._commute)
    return g


def _extended(gate, new_targets):
    """Matrix of `gate` on the ordered target list `new_targets`; controls of
    the gate that appear there act as conditions."""
    K = len(new_targets)
    D = 1 << K
    pos = {q: i for i, q in enumerate(new_targets)}
    tpos = [pos[t] for t in gate._targets]
    gmask = sum(1 << p for p in tpos)
    cmask = cval = 0
    for q, v in zip(gate._controls, gate._cvals):
        if q in pos:
            cmask |= 1 << pos[q]
            cval |= v << pos[q]
    gd = 1 << len(tpos)
    M = np.zeros((D, D), dtype=np.complex128)
    for col in range(D):
        if (col & cmask) != cval:
            M[col, col] = 1
            continue
        cg = sum(((col >> p) & 1) << i for i, p in enumerate(tpos))
        for rg in range(gd):
            row = (col & ~gmask) | sum(((rg >> i) & 1) << p for i, p in enumerate(tpos))
            M[row, col] = gate._matrix[rg, cg]
    return M


def merge(gate_first, gate_second=None):
    """gate::merge (src/cppsim/gate_merge.cpp:316-356): product second*first on
    the sorted union of qubits; shared equal controls stay controls."""
    if gate_second is None:
        gates = list(gate_first)
        cur = to_matrix_gate(gates[0])
        for g in gates[1:]:
            cur = merge(cur, g)
        return cur
    g1, g2 = gate_first, gate_second
    targets = {}
    controls = {}
    for q, p in zip(g1._targets, g1._commute):
        if q in g2._targets:
            targets[q] = p & g2._commute[g2._targets.index(q)]
        elif q in g2._controls:
            targets[q] = p & COMMUTE_Z
        else:
            targets[q] = p
    for q, v in zip(g1._controls, g1._cvals):
        if q in g2._targets:
            targets[q] = g2._commute[g2._targets.index(q)] & COMMUTE_Z
        elif q in g2._controls:
            if g2._cvals[g2._controls.index(q)] == v:
                controls[q] = v
            else:
                targets[q] = COMMUTE_Z
        else:
            targets[q] = COMMUTE_Z
    for q, p in zip(g2._targets, g2._commute):
        if q not in g1._targets and q not in g1._controls:
            targets[q] = p
    for q in g2._controls:
        if q not in g1._targets and q not in g1._controls:
            targets[q] = COMMUTE_Z
    nt = sorted(targets)
    m = _extended(g2, nt) @ _extended(g1, nt)
    nc = sorted(controls)
    return QuantumGate("DenseMatrix", nt, m, nc, [controls[q] for q in nc], [targets[q] for q in nt])
