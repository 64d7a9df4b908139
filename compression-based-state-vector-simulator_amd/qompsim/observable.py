"""Pauli-sum observables (src/cppsim/observable.cpp, pauli_operator.cpp:132-183).

A term coef * P is held as an X-flip mask, a Z-phase mask (Y sets both) and
its Y count, which is what the HIP expectation and apply kernels take:
P|i> = i^nY (-1)^popcount(i & zmask) |i ^ xmask>.
"""
import ctypes

import numpy as np

from ._lib import call, ptr

_PAULI_CHARS = {"I": 0, "X": 1, "Y": 2, "Z": 3}


class PauliTerm:
    def __init__(self, coef, index_list, pauli_ids):
        self.coef = complex(coef)
        self.index_list = [int(i) for i in index_list]
        self.pauli_ids = [int(p) for p in pauli_ids]
        x = z = 0
        ny = 0
        for q, p in zip(self.index_list, self.pauli_ids):
            if p in (1, 2):
                x |= 1 << q
            if p in (2, 3):
                z |= 1 << q
            if p == 2:
                ny += 1
        self.xmask, self.zmask, self.ny = x, z, ny

    def get_coef(self):
        return self.coef

    def get_index_list(self):
        return list(self.index_list)

    def get_pauli_id_list(self):
        return list(self.pauli_ids)


def parse_pauli_string(s):
    toks = s.split()
    if len(toks) % 2:
        raise ValueError(f"bad Pauli string {s!r}")
    idx, ids = [], []
    for a, b in zip(toks[::2], toks[1::2]):
        p = _PAULI_CHARS[a.upper()]
        if p == 0:
            continue
        idx.append(int(b))
        ids.append(p)
    return idx, ids


class GeneralQuantumOperator:
    def __init__(self, qubit_count):
        self._n = int(qubit_count)
        self.terms = []

    def get_qubit_count(self):
        return self._n

    def get_term_count(self):
        return len(self.terms)

    def get_term(self, i):
        return self.terms[i]

    def add_operator(self, coef, pauli, *args):
        if isinstance(pauli, str):
            idx, ids = parse_pauli_string(pauli)
        else:
            idx, ids = list(pauli), list(args[0])
        if any(q >= self._n for q in idx):
            raise ValueError("Pauli index outside the register")
        self.terms.append(PauliTerm(coef, idx, ids))

    def _arrays(self):
        t = self.terms
        xm = np.array([p.xmask for p in t], dtype=np.uint64)
        zm = np.array([p.zmask for p in t], dtype=np.uint64)
        ny = np.array([p.ny for p in t], dtype=np.uint32)
        cr = np.array([p.coef.real for p in t], dtype=np.float64)
        ci = np.array([p.coef.imag for p in t], dtype=np.float64)
        return xm, zm, ny, cr, ci

    def get_expectation_value(self, state):
        if not self.terms:
            return 0.0
        xm, zm, ny, cr, ci = self._arrays()
        re, im = ctypes.c_double(), ctypes.c_double()
        call("qsim_sv_expectation_pauli", state._h, len(self.terms), ptr(xm), ptr(zm), ptr(ny), ptr(cr), ptr(ci),
             ctypes.byref(re), ctypes.byref(im))
        return complex(re.value, im.value)

    def apply_to_state(self, work_state, state_in, state_out):
        """state_out = H |state_in> (GeneralQuantumOperator::apply_to_state)."""
        xm, zm, ny, cr, ci = self._arrays()
        call("qsim_sv_apply_pauli_sum", state_in._h, state_out._h, len(self.terms), ptr(xm), ptr(zm), ptr(ny),
             ptr(cr), ptr(ci))


class Observable(GeneralQuantumOperator):
    """Hermitian Pauli sum; expectation values are real."""

    def add_operator(self, coef, pauli, *args):
        if np.iscomplexobj(coef) and np.imag(coef) != 0:
            raise ValueError("Observable coefficients must be real")
        super().add_operator(coef, pauli, *args)

    def get_expectation_value(self, state):
        return super().get_expectation_value(state).real
