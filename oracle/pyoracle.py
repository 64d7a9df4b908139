"""ORACLE — test infrastructure.  ctypes binding of oracle/liboracle.so (the CPU
restatement) and, when built, oracle/_ref/libcsim_ref.so (the reference's own
csim kernels compiled from its sources).  Only tests/, __graft_entry__.smoke()
and bench.py's cpu_baseline leg may import this module.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_so = os.path.join(HERE, "liboracle.so")
if not os.path.exists(_so):
    raise ImportError("oracle/liboracle.so missing: run python oracle/build.py")
lib = ctypes.CDLL(_so)

_vp, _u32, _u64, _dbl, _int = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_double, ctypes.c_int
_P = ctypes.POINTER
for name, res, args in [
    ("orc_num_threads", _int, []),
    ("orc_set_threads", None, [_int]),
    ("orc_apply_circuit", None, [_vp, _u32, _vp, _u64, _vp]),
    ("orc_norm", _dbl, [_vp, _u64]),
    ("orc_inner", None, [_vp, _vp, _u64, _P(_dbl), _P(_dbl)]),
    ("orc_expectation", None, [_vp, _u32, _u64, _vp, _vp, _vp, _vp, _vp, _P(_dbl), _P(_dbl)]),
    ("orc_apply_pauli_sum", None, [_vp, _vp, _u32, _u64, _vp, _vp, _vp, _vp, _vp]),
    ("orc_codec_roundtrip", None, [_vp, _u64, _dbl, _dbl, _int, _vp, _P(_u64), _vp]),
    ("orc_cs_create", _vp, [_u32, _u32, _dbl, _dbl, _int, _int, _dbl]),
    ("orc_cs_destroy", None, [_vp]),
    ("orc_cs_load", None, [_vp, _vp]),
    ("orc_cs_partition_and_register", _u32, [_vp, _vp, _u64, _vp, _u32]),
    ("orc_cs_stage", _u32, [_vp, _u32, _vp, _vp, _P(_u32)]),
    ("orc_cs_apply_circuit", None, [_vp]),
    ("orc_cs_run_stages", None, [_vp, _u32, _u32]),
    ("orc_cs_read_vector", None, [_vp, _vp, _int]),
    ("orc_cs_null_bitmap", None, [_vp, _vp]),
    ("orc_cs_bytes_history", _u64, [_vp, _vp, _u64]),
    ("orc_cs_total_bytes", _u64, [_vp]),
    ("orc_cs_sampling_blockwise", _int, [_vp, _u32, _u32, _vp]),
    ("orc_backprop", None, [_u32, _vp, _vp, _u64, _vp, _u64, _vp, _vp, _vp, _vp, _vp, _u32, _vp, _P(_dbl)]),
]:
    f = getattr(lib, name)
    f.restype = res
    f.argtypes = args


def _p(a):
    return None if a is None else ctypes.c_void_p(a.ctypes.data)


def _pack(gates):
    from qompsim._lib import pack_gates

    return pack_gates(gates)


def apply_circuit(psi, gates):
    """Apply gate objects (qompsim.gate) to a copy of psi on the CPU."""
    psi = np.ascontiguousarray(psi, dtype=np.complex128).copy()
    n = int(np.log2(psi.size))
    if gates:
        rec, pool = _pack(gates)
        lib.orc_apply_circuit(_p(psi), n, _p(rec), len(gates), _p(pool))
    return psi


def zero_state(n):
    v = np.zeros(1 << n, dtype=np.complex128)
    v[0] = 1
    return v


def norm(psi):
    psi = np.ascontiguousarray(psi, dtype=np.complex128)
    return lib.orc_norm(_p(psi), psi.size)


def inner(a, b):
    a = np.ascontiguousarray(a, dtype=np.complex128)
    b = np.ascontiguousarray(b, dtype=np.complex128)
    re, im = ctypes.c_double(), ctypes.c_double()
    lib.orc_inner(_p(a), _p(b), a.size, ctypes.byref(re), ctypes.byref(im))
    return complex(re.value, im.value)


def _terms(obs):
    return obs._arrays()


def expectation(psi, obs):
    psi = np.ascontiguousarray(psi, dtype=np.complex128)
    xm, zm, ny, cr, ci = _terms(obs)
    re, im = ctypes.c_double(), ctypes.c_double()
    lib.orc_expectation(_p(psi), int(np.log2(psi.size)), len(xm), _p(xm), _p(zm), _p(ny), _p(cr), _p(ci),
                        ctypes.byref(re), ctypes.byref(im))
    return complex(re.value, im.value)


def apply_pauli_sum(psi, obs):
    psi = np.ascontiguousarray(psi, dtype=np.complex128)
    out = np.empty_like(psi)
    xm, zm, ny, cr, ci = _terms(obs)
    lib.orc_apply_pauli_sum(_p(psi), _p(out), int(np.log2(psi.size)), len(xm), _p(xm), _p(zm), _p(ny), _p(cr),
                            _p(ci))
    return out


def codec_roundtrip(x, eps_rel, eps_abs, guard):
    x = np.ascontiguousarray(x, dtype=np.float64)
    assert x.size % 64 == 0
    out = np.empty_like(x)
    words = ctypes.c_uint64()
    widths = np.zeros(x.size // 64, dtype=np.uint8)
    lib.orc_codec_roundtrip(_p(x), x.size, eps_rel, eps_abs, guard, _p(out), ctypes.byref(words), _p(widths))
    return out, words.value, widths


def backprop(circuit, obs, init=None):
    n = circuit.get_qubit_count()
    init = zero_state(n) if init is None else np.ascontiguousarray(init, dtype=np.complex128)
    rec, pool = _pack(circuit.gate_list)
    xm, zm, ny, cr, ci = _terms(obs)
    np_ = circuit.get_parameter_count()
    grads = np.zeros(max(1, np_))
    e = ctypes.c_double()
    lib.orc_backprop(n, _p(init), _p(rec), len(circuit.gate_list), _p(pool), len(xm), _p(xm), _p(zm), _p(ny), _p(cr),
                     _p(ci), np_, _p(grads), ctypes.byref(e))
    return e.value, grads[:np_]


class OracleCompressedState:
    """CPU restatement of CompressedQuantumState's stage loop."""

    def __init__(self, n, B, eps_rel=1e-3, eps_abs=1e-7, guard=1, zero_mode="exact", zero_threshold=1e-6):
        self.n, self.B = n, B
        zm = 1 if zero_mode == "threshold" else 0
        self._h = lib.orc_cs_create(n, B, eps_rel, eps_abs, guard, zm, zero_threshold)
        if not self._h:
            raise ValueError("block_size must satisfy 5 <= block_size <= qubit_count")

    def __del__(self):
        if getattr(self, "_h", None):
            lib.orc_cs_destroy(self._h)
            self._h = None

    def load(self, vec):
        v = np.ascontiguousarray(vec, dtype=np.complex128)
        lib.orc_cs_load(self._h, _p(v))

    def partition_and_register_circuit(self, circuit, inner_size):
        rec, pool = _pack(circuit.gate_list)
        self._nst = lib.orc_cs_partition_and_register(self._h, _p(rec), len(circuit.gate_list), _p(pool), inner_size)
        return self._nst

    def stages(self):
        res = []
        for s in range(self._nst):
            cnt = lib.orc_cs_stage(self._h, s, None, None, None)
            ids = np.zeros(max(1, cnt), dtype=np.uint32)
            b2g = np.zeros(1 << (self.n - self.B), dtype=np.uint32)
            ng = ctypes.c_uint32()
            lib.orc_cs_stage(self._h, s, _p(ids), _p(b2g), ctypes.byref(ng))
            res.append((ids[:cnt].tolist(), b2g, ng.value))
        return res

    def apply_circuit(self):
        lib.orc_cs_apply_circuit(self._h)

    def run_stages(self, b, e):
        lib.orc_cs_run_stages(self._h, b, e)

    def get_vector(self, normalise=True):
        out = np.empty(1 << self.n, dtype=np.complex128)
        lib.orc_cs_read_vector(self._h, _p(out), 1 if normalise else 0)
        return out

    def null_bitmap(self):
        out = np.zeros(1 << (self.n - self.B), dtype=np.uint8)
        lib.orc_cs_null_bitmap(self._h, _p(out))
        return out.astype(bool)

    def sampling_blockwise(self, count, seed):
        out = np.zeros(max(1, count), dtype=np.uint64)
        if lib.orc_cs_sampling_blockwise(self._h, count, seed, _p(out)) != 0:
            raise ValueError("state has no finite positive probability mass")
        return out[:count].tolist()

    def bytes_history(self):
        n = lib.orc_cs_bytes_history(self._h, None, 0)
        out = np.zeros(max(1, n), dtype=np.uint64)
        lib.orc_cs_bytes_history(self._h, _p(out), n)
        return out[:n].tolist()


# --------------------------------------------- the reference's own csim kernels
REF_SO = os.path.join(HERE, "_ref", "libcsim_ref.so")


class RefCsim:
    """Reference csim kernels (src/csim), compiled from the reference tree."""

    def __init__(self):
        self.lib = ctypes.CDLL(REF_SO)
        L = self.lib
        self.mcmt = L["_Z49multi_qubit_control_multi_qubit_dense_matrix_gatePKjS0_jS0_jPKSt7complexIdEPS2_y"]
        self.mcmt.argtypes = [_vp, _vp, _u32, _vp, _u32, _vp, _vp, _u64]
        self.multi = L["_Z29multi_qubit_dense_matrix_gatePKjjPKSt7complexIdEPS2_y"]
        self.multi.argtypes = [_vp, _u32, _vp, _vp, _u64]
        self.single = L["_Z30single_qubit_dense_matrix_gatejPKSt7complexIdEPS0_y"]
        self.single.argtypes = [_u32, _vp, _vp, _u64]
        self.double = L["_Z32double_qubit_dense_matrix_gate_cjjPKSt7complexIdEPS0_y"]
        self.double.argtypes = [_u32, _u32, _vp, _vp, _u64]
        self.cnot = L["_Z9CNOT_gatejjPSt7complexIdEy"]
        self.cnot.argtypes = [_u32, _u32, _vp, _u64]
        self.cz = L["_Z7CZ_gatejjPSt7complexIdEy"]
        self.cz.argtypes = [_u32, _u32, _vp, _u64]
        self.h = L["_Z6H_gatejPSt7complexIdEy"]
        self.h.argtypes = [_u32, _vp, _u64]
        self.rx = L["_Z7RX_gatejdPSt7complexIdEy"]
        self.rx.argtypes = [_u32, _dbl, _vp, _u64]
        self.ry = L["_Z7RY_gatejdPSt7complexIdEy"]
        self.ry.argtypes = [_u32, _dbl, _vp, _u64]
        self.rz = L["_Z7RZ_gatejdPSt7complexIdEy"]
        self.rz.argtypes = [_u32, _dbl, _vp, _u64]
        self.norm = L["_Z18state_norm_squaredPKSt7complexIdEy"]
        self.norm.argtypes = [_vp, _u64]
        self.norm.restype = _dbl

    def apply_gate(self, psi, gate):
        """QuantumGateMatrix::update_quantum_state dispatch (gate_matrix.cpp:71-250)."""
        t = np.array(gate.get_target_index_list(), dtype=np.uint32)
        c = np.array(gate.get_control_index_list(), dtype=np.uint32)
        cv = np.array(gate.get_control_value_list(), dtype=np.uint32)
        m = np.ascontiguousarray(gate.get_matrix(), dtype=np.complex128)
        dim = psi.size
        if c.size == 0:
            if t.size == 1:
                self.single(int(t[0]), _p(m), _p(psi), dim)
            elif t.size == 2:
                self.double(int(t[0]), int(t[1]), _p(m), _p(psi), dim)
            else:
                self.multi(_p(t), t.size, _p(m), _p(psi), dim)
        else:
            self.mcmt(_p(c), _p(cv), c.size, _p(t), t.size, _p(m), _p(psi), dim)
