"""QuantumState (uncompressed, resident in HBM) and CompressedQuantumState.

QuantumState mirrors the Qulacs state API (src/cppsim/state.hpp:445-1039);
CompressedQuantumState mirrors the pybind class of the reference
(python/cppsim_wrapper.cpp:447-519) over the compressed block store of
src/cppsim/compressed_state.hpp:87-1707.  Both are thin handles over
libqsim_hip.so; no amplitude arithmetic happens in Python.
"""
import ctypes
import os

import numpy as np

from . import _lib
from ._lib import call, lib, ptr


class QuantumState:
    """Uncompressed complex128 state vector of 2^n amplitudes in HBM."""

    def __init__(self, qubit_count, use_multi_cpu=False):
        h = ctypes.c_void_p()
        call("qsim_sv_create", int(qubit_count), ctypes.byref(h))
        self._h = h
        self._n = int(qubit_count)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            lib.qsim_sv_destroy(h)
            self._h = None

    @property
    def dim(self):
        return 1 << self._n

    def get_qubit_count(self):
        return self._n

    def set_zero_state(self):
        call("qsim_sv_set_basis", self._h, 0)

    def set_computational_basis(self, comp_basis):
        call("qsim_sv_set_basis", self._h, int(comp_basis))

    def load(self, state):
        if isinstance(state, QuantumState):
            call("qsim_sv_copy", self._h, state._h)
            return
        v = np.ascontiguousarray(np.asarray(state, dtype=np.complex128).reshape(-1))
        if v.size != self.dim:
            raise ValueError("vector length does not match the qubit count")
        call("qsim_sv_load", self._h, ptr(v))

    def get_vector(self):
        out = np.empty(self.dim, dtype=np.complex128)
        call("qsim_sv_read", self._h, ptr(out))
        return out

    def get_squared_norm(self):
        r = ctypes.c_double()
        call("qsim_sv_norm_squared", self._h, ctypes.byref(r))
        return r.value

    def normalize(self, squared_norm):
        call("qsim_sv_normalize", self._h, float(squared_norm))

    def multiply_coef(self, coef):
        if np.iscomplexobj(coef) and np.imag(coef) != 0:
            raise NotImplementedError("complex coefficient")
        call("qsim_sv_scale", self._h, float(np.real(coef)))

    def copy(self):
        s = QuantumState(self._n)
        s.load(self)
        return s

    def device_ptr(self):
        return lib.qsim_sv_device_ptr(self._h)

    def get_fused_swap_time(self):
        """Seconds spent in FusedSWAP exchanges between ranks
        (python/cppsim_wrapper.cpp:412-416).  One GPU holds the whole state,
        so there is no exchange: always 0.0."""
        return 0.0

    def __repr__(self):
        return f"QuantumState(qubit_count={self._n})"


def partition_commuting(circuit, block_size, inner_size):
    """Stages (lists of gate indices, each in circuit order) of the
    commutation-aware packing; host-side planning only."""
    gates = list(circuit.gate_list)
    rec, pool = _lib.pack_gates(gates)
    out = np.zeros(max(1, len(gates)), dtype=np.uint32)
    ns = ctypes.c_uint32()
    call("qsim_partition_commuting", int(circuit.get_qubit_count()), int(block_size), int(inner_size), ptr(rec),
         len(gates), ptr(pool), pool.size, ptr(out), ctypes.byref(ns))
    stages = [[] for _ in range(ns.value)]
    for i in range(len(gates)):
        stages[int(out[i])].append(i)
    return stages


def inner_product(state_bra, state_ket):
    """state::inner_product (src/cppsim/state.cpp:9-31): <bra|ket>."""
    re, im = ctypes.c_double(), ctypes.c_double()
    call("qsim_sv_inner_product", state_bra._h, state_ket._h, ctypes.byref(re), ctypes.byref(im))
    return complex(re.value, im.value)


def _format_size(size):
    """format_size of compressed_state.hpp:360-372."""
    if size > 1 << 30:
        return f"{size / (1 << 30):.3f} GiB"
    if size > 1 << 20:
        return f"{size / (1 << 20):.3f} MiB"
    if size > 1 << 10:
        return f"{size / (1 << 10):.3f} KiB"
    return f"{size} bytes"


def _env_float(name, default):
    v = os.environ.get(name)
    return float(v) if v is not None else default


def default_codec_config():
    """Codec settings from the environment (compressed_state.hpp:128-139, 184-189)."""
    cfg = _lib.CsConfig()
    call("qsim_cs_default_config", ctypes.byref(cfg))
    cfg.eps_rel = _env_float("BITCOMP_REL_EB", cfg.eps_rel)
    cfg.eps_abs = _env_float("BITCOMP_ABS_EB", cfg.eps_abs)
    cfg.guard_bits = int(os.environ.get("QOMPSIM_GUARD_BITS", cfg.guard_bits))
    mode = os.environ.get("QOMPSIM_ZERO_MODE", "exact")
    cfg.zero_mode = _lib.ZERO_THRESHOLD if mode == "threshold" else _lib.ZERO_EXACT
    cfg.zero_threshold = _env_float("ALL_ZERO_THRESHOLD", 10.0 * cfg.eps_abs)
    return cfg


class CompressedQuantumState:
    """Block-compressed state in HBM (CompressedQuantumState(qubit_count, block_size))."""

    def __init__(self, qubit_count, block_size, *, eps_rel=None, eps_abs=None, guard_bits=None,
                 zero_mode=None, zero_threshold=None, scratch_bytes=None, tile_bits=None, store="double",
                 hbm_budget=None, verbose=True):
        """store: "double" (two whole arenas) or "regions" (payload regions
        mapped on demand and freed as a stage consumes them: about one
        compressed state of peak memory instead of two; no checkpoint
        files).  hbm_budget: device bytes the store may hold (descriptors,
        staging buffer, arenas or mapped regions); growth past it raises
        RuntimeError."""
        cfg = default_codec_config()
        stores = {"double": _lib.STORE_DOUBLE, "regions": _lib.STORE_REGIONS}
        if store not in stores:
            raise ValueError(f"store must be one of {sorted(stores)}")
        cfg.store_mode = stores[store]
        if hbm_budget is not None:
            cfg.hbm_budget = int(hbm_budget)
        if eps_rel is not None:
            cfg.eps_rel = eps_rel
        if eps_abs is not None:
            cfg.eps_abs = eps_abs
            if zero_threshold is None and "ALL_ZERO_THRESHOLD" not in os.environ:
                cfg.zero_threshold = 10.0 * eps_abs
        if guard_bits is not None:
            cfg.guard_bits = guard_bits
        if zero_mode is not None:
            cfg.zero_mode = {"exact": _lib.ZERO_EXACT, "threshold": _lib.ZERO_THRESHOLD}.get(zero_mode, zero_mode)
        if zero_threshold is not None:
            cfg.zero_threshold = zero_threshold
        if scratch_bytes is not None:
            cfg.scratch_bytes = int(scratch_bytes)
        if tile_bits is not None:
            cfg.tile_bits = int(tile_bits)
        h = ctypes.c_void_p()
        call("qsim_cs_create", int(qubit_count), int(block_size), ctypes.byref(cfg), ctypes.byref(h))
        self._h = h
        self._n = int(qubit_count)
        self._B = int(block_size)
        self._cfg = cfg
        self._gates = []
        self._verbose = bool(verbose)
        if verbose:
            print("GPU (MI355X) block codec")
            print(f"  relative error bound: {cfg.eps_rel}")
            print(f"  absolute error bound: {cfg.eps_abs}")

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            lib.qsim_cs_destroy(h)
            self._h = None

    # ---- reference surface
    def print_info(self):
        print(f" block_size: {self._B}")
        print(f" namps_per_block: {1 << self._B}")
        print(f" nblocks: {1 << (self._n - self._B)}")

    def get_qubit_count(self):
        return self._n

    @property
    def nblocks(self):
        return 1 << (self._n - self._B)

    def partition_and_register_circuit(self, circuit, inner_size, verbose=0, reorder=False):
        """partition_circuit + make_block_groups + registration
        (compressed_state.hpp:673-726).  reorder=True packs stages with the
        commutation-aware rule instead (qsim_partition_commuting): fewer
        stages, the same circuit."""
        gates = list(circuit.gate_list)
        rec, pool = _lib.pack_gates(gates)
        fn = "qsim_cs_partition_commuting_and_register" if reorder else "qsim_cs_partition_and_register"
        call(fn, self._h, ptr(rec), len(gates), ptr(pool), pool.size, int(inner_size))
        self._gates = gates
        if verbose > 0:
            sizes = [len(s) for s in self.get_stages()]
            print("Circuit Stages: [" + ", ".join(str(s) for s in sizes) + "]")
        if verbose > 1:
            print("\n=== Block Groups ===")
            for si, groups in enumerate(self.get_block_groups()):
                print(f"Stage {si} block groups:")
                for gi, g in enumerate(groups):
                    print(f"  Group {gi}: [" + ", ".join(str(b) for b in g) + "]")

    def register_circuit_stages(self, circuit, stages, block_groups_list):
        gates = list(circuit.gate_list)
        rec, pool = _lib.pack_gates(gates)
        nb = self.nblocks
        offs = np.zeros(len(stages) + 1, dtype=np.uint32)
        ids = []
        for i, s in enumerate(stages):
            offs[i + 1] = offs[i] + len(s)
            ids.extend(int(x) for x in s)
        ids = np.asarray(ids if ids else [0], dtype=np.uint32)
        b2g = np.zeros((len(stages), nb), dtype=np.uint32)
        ng = np.zeros(len(stages), dtype=np.uint32)
        for si, groups in enumerate(block_groups_list):
            ng[si] = len(groups)
            seen = 0
            for gi, g in enumerate(groups):
                for b in g:
                    b2g[si, int(b)] = gi
                    seen += 1
            if seen != nb:
                raise ValueError("block groups must cover every block exactly once")
        call("qsim_cs_register_stages", self._h, ptr(rec), len(gates), ptr(pool), pool.size, len(stages), ptr(offs),
             ptr(ids), ptr(np.ascontiguousarray(b2g)), ptr(ng))
        self._gates = gates

    def apply_circuit(self):
        """The stage loop (compressed_state.hpp:840-937).  With verbose=True
        the total compressed size is printed before the first stage and
        after every stage, as append_total_SV_bytes prints it (:374-437,
        931-933), once the loop has finished."""
        call("qsim_cs_apply_circuit", self._h)
        if self._verbose:
            hist = self.get_bytes_history()
            ns = self.get_num_stages()
            hist = hist[-(ns + 1):]
            print(f"Total SV bytes: {_format_size(hist[0])}")
            for i, b in enumerate(hist[1:]):
                print(f"[Stage {i}/{ns - 1}] Total SV bytes: {_format_size(b)}")

    def run_stages(self, begin, end):
        call("qsim_cs_run_stages", self._h, int(begin), int(end))

    def get_vector(self):
        """Decompressed, renormalised copy (compressed_state.hpp:1683-1707)."""
        out = np.empty(1 << self._n, dtype=np.complex128)
        call("qsim_cs_read_vector", self._h, ptr(out), 1)
        return out

    def get_raw_vector(self):
        out = np.empty(1 << self._n, dtype=np.complex128)
        call("qsim_cs_read_vector", self._h, ptr(out), 0)
        return out

    def get_block(self, block_id):
        out = np.empty(1 << self._B, dtype=np.complex128)
        call("qsim_cs_read_block", self._h, int(block_id), ptr(out))
        return out

    def get_max_SV_bytes(self):
        r = ctypes.c_uint64()
        call("qsim_cs_max_sv_bytes", self._h, ctypes.byref(r))
        return r.value

    def get_total_times(self):
        t = np.zeros(len(_lib.TIMER_KEYS), dtype=np.float64)
        call("qsim_cs_times", self._h, ptr(t))
        return {k: float(v) for k, v in zip(_lib.TIMER_KEYS, t)}

    # ---- extensions (store inspection and measurement)
    def set_zero_state(self):
        call("qsim_cs_set_basis", self._h, 0)

    def set_computational_basis(self, comp_basis):
        call("qsim_cs_set_basis", self._h, int(comp_basis))

    def load(self, vector):
        v = np.ascontiguousarray(np.asarray(vector, dtype=np.complex128).reshape(-1))
        if v.size != 1 << self._n:
            raise ValueError("vector length does not match the qubit count")
        call("qsim_cs_load", self._h, ptr(v))

    def sampling_blockwise(self, sampling_count, random_seed):
        """Basis-state samples drawn block by block
        (CompressedQuantumState::sampling_blockwise,
        src/cppsim/compressed_state.hpp:1600-1681): a list of
        `sampling_count` indices; ValueError when the state has no finite
        positive probability mass."""
        n = int(sampling_count)
        if n < 0 or n >= 1 << 32:
            raise ValueError("sampling_count out of range")
        out = np.zeros(max(1, n), dtype=np.uint64)
        call("qsim_cs_sampling_blockwise", self._h, n, int(random_seed) & 0xFFFFFFFF, ptr(out))
        return [int(x) for x in out[:n]]

    def save_checkpoint(self, path, next_stage=0):
        """Write the compressed store (descriptors, null flags, arena) to
        `path`; `next_stage` is kept for resuming a staged run."""
        call("qsim_cs_save_checkpoint", self._h, os.fsencode(path), int(next_stage))

    def load_checkpoint(self, path):
        """Replace the store by a checkpoint written by save_checkpoint from a
        store of the same shape and codec settings; returns its next_stage."""
        ns = ctypes.c_uint64()
        call("qsim_cs_load_checkpoint", self._h, os.fsencode(path), ctypes.byref(ns))
        return int(ns.value)

    def get_null_bitmap(self):
        out = np.empty(self.nblocks, dtype=np.uint8)
        call("qsim_cs_null_bitmap", self._h, ptr(out))
        return out.astype(bool)

    def get_nonnull_blocks(self):
        return [int(b) for b in np.nonzero(~self.get_null_bitmap())[0]]

    def get_block_bytes(self):
        out = np.empty(self.nblocks, dtype=np.uint64)
        call("qsim_cs_block_bytes", self._h, ptr(out))
        return out

    def get_store_bytes(self):
        """Device bytes of the payload store: {"arena": held now, "arena_peak":
        most held at once, "descriptors": the descriptor tables (double store:
        input and output; region store: one table, encoded in place, and the
        copy of a stage batch's input descriptors)}."""
        out = np.zeros(3, dtype=np.uint64)
        call("qsim_cs_store_bytes", self._h, ptr(out))
        return {"arena": int(out[0]), "arena_peak": int(out[1]), "descriptors": int(out[2])}

    def get_bytes_history(self):
        n = ctypes.c_uint64()
        call("qsim_cs_bytes_history", self._h, None, 0, ctypes.byref(n))
        out = np.zeros(max(1, n.value), dtype=np.uint64)
        call("qsim_cs_bytes_history", self._h, ptr(out), n.value, ctypes.byref(n))
        return [int(x) for x in out[: n.value]]

    def get_stage_stats(self):
        """Per executed stage: gpu_seconds, bytes_in, bytes_out, decompressed amplitudes, passes."""
        n = ctypes.c_uint64()
        call("qsim_cs_stage_stats", self._h, None, 0, ctypes.byref(n))
        out = np.zeros((max(1, n.value), 5), dtype=np.float64)
        call("qsim_cs_stage_stats", self._h, ptr(out), n.value, ctypes.byref(n))
        keys = ("gpu_seconds", "bytes_in", "bytes_out", "amplitudes", "passes")
        return [dict(zip(keys, row)) for row in out[: n.value]]

    def get_num_stages(self):
        r = ctypes.c_uint32()
        call("qsim_cs_num_stages", self._h, ctypes.byref(r))
        return r.value

    def _stage_info(self, s, want_b2g=False):
        ng, ngr, npass = ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_uint32()
        call("qsim_cs_stage_info", self._h, s, ctypes.byref(ng), None, ctypes.byref(ngr), None, ctypes.byref(npass))
        ids = np.zeros(max(1, ng.value), dtype=np.uint32)
        b2g = np.zeros(self.nblocks, dtype=np.uint32) if want_b2g else None
        call("qsim_cs_stage_info", self._h, s, ctypes.byref(ng), ptr(ids), ctypes.byref(ngr), ptr(b2g),
             ctypes.byref(npass))
        return [int(x) for x in ids[: ng.value]], ngr.value, b2g, npass.value

    def get_stages(self):
        return [self._stage_info(s)[0] for s in range(self.get_num_stages())]

    def get_stage_passes(self):
        return [self._stage_info(s)[3] for s in range(self.get_num_stages())]

    def get_block_to_group_list(self):
        return [self._stage_info(s, True)[2] for s in range(self.get_num_stages())]

    def get_block_groups(self):
        res = []
        for s in range(self.get_num_stages()):
            _, ngr, b2g, _ = self._stage_info(s, True)
            groups = [[] for _ in range(ngr)]
            for b, g in enumerate(b2g):
                groups[int(g)].append(b)
            res.append(groups)
        return res

    def get_squared_norm(self):
        r = ctypes.c_double()
        call("qsim_cs_norm_squared", self._h, ctypes.byref(r))
        return r.value

    def inner_product_with(self, state):
        """<state|self> against an uncompressed QuantumState, on the GPU."""
        re, im = ctypes.c_double(), ctypes.c_double()
        call("qsim_cs_inner_product_sv", self._h, state._h, ctypes.byref(re), ctypes.byref(im))
        return complex(re.value, im.value)

    def get_expectation_value(self, observable):
        """<psi|H|psi> over the compressed store (decoded in batches on the GPU);
        real for an Observable."""
        if not observable.terms:
            return 0.0
        xm, zm, ny, cr, ci = observable._arrays()
        re, im = ctypes.c_double(), ctypes.c_double()
        call("qsim_cs_expectation_pauli", self._h, len(observable.terms), ptr(xm), ptr(zm), ptr(ny), ptr(cr), ptr(ci),
             ctypes.byref(re), ctypes.byref(im))
        from .observable import Observable

        return re.value if isinstance(observable, Observable) else complex(re.value, im.value)

    def fidelity_with(self, state):
        """|<state|self>| / ||self|| (run_sim.py:187-205 after the readout renormalisation)."""
        return abs(self.inner_product_with(state)) / np.sqrt(self.get_squared_norm())
