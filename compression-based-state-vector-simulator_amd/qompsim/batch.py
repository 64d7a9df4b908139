"""Batched energies and adjoint gradients of many small circuits
(qsim_batch_*): every instance is planned once, the whole batch is captured
into one hipGraph and replayed, and an autotuner picks the forward engine /
tile and the adjoint tile size per qubit count.
"""
import ctypes

import numpy as np

from ._lib import call, lib, pack_gates, ptr


class BatchedAdjoint:
    def __init__(self):
        h = ctypes.c_void_p()
        call("qsim_batch_create", ctypes.byref(h))
        self._h = h
        self.qubits = []

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib.qsim_batch_destroy(h)
            self._h = None

    def add(self, circuit, observable, n_params=None):
        """Add one circuit (gates with param indices) and its observable;
        returns the instance index."""
        rec, pool = pack_gates(circuit.gate_list)
        xm, zm, ny, cr, ci = observable._arrays()
        if n_params is None:
            n_params = circuit.get_parameter_count() if hasattr(circuit, "get_parameter_count") else 0
        idx = ctypes.c_uint64()
        call("qsim_batch_add", self._h, circuit.get_qubit_count(), ptr(rec), len(circuit.gate_list), ptr(pool),
             pool.size, len(observable.terms), ptr(xm), ptr(zm), ptr(ny), ptr(cr), ptr(ci), int(n_params),
             ctypes.byref(idx))
        self.qubits.append(circuit.get_qubit_count())
        return idx.value

    def tune(self, reps=2):
        call("qsim_batch_tune", self._h, int(reps))

    def config(self, n):
        e, f, ae, a = ctypes.c_int32(), ctypes.c_uint32(), ctypes.c_int32(), ctypes.c_uint32()
        call("qsim_batch_config", self._h, int(n), ctypes.byref(e), ctypes.byref(f), ctypes.byref(ae),
             ctypes.byref(a))
        names = {-1: "default", 0: "lds", 1: "wave"}
        return {"forward_engine": names.get(e.value, str(e.value)), "forward_tile_bits": f.value,
                "adjoint_engine": names.get(ae.value, str(ae.value)), "adjoint_tile_bits": a.value}

    def traffic(self):
        """(algorithmic HBM bytes of one run, instances in the persistent kernel)"""
        t, k = ctypes.c_uint64(), ctypes.c_uint64()
        call("qsim_batch_traffic", self._h, ctypes.byref(t), ctypes.byref(k))
        return t.value, k.value

    def run(self, use_graph=True):
        """(energies[n_instances], grads[n_instances, max_params])"""
        n = lib.qsim_batch_size(self._h)
        p = lib.qsim_batch_max_params(self._h)
        e = np.zeros(max(n, 1))
        g = np.zeros((max(n, 1), max(p, 1)))
        call("qsim_batch_run", self._h, 1 if use_graph else 0, ptr(e), ptr(g))
        return e[:n], g[:n, :p]
