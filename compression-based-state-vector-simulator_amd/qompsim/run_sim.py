"""Harness with the modes of the reference's script/run_sim.py:60-231 —
``qulacs`` (uncompressed state), ``qompsim`` (compressed stage loop) and
``fidelity`` (both, then |<psi_qulacs|psi_qompsim>| after the readout
renormalisation) — on one GPU, without MPI ranks.

Result lines are JSON objects {times, sv_MiB, fidelity} as in the reference
(run_sim.py:19-23, 64-81); file names follow
``{sim}_{circuit}_{n}qubits_1ranks[_bs{B}_inner{I}].json``.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

from . import CompressedQuantumState, QuantumCircuitOptimizer, QuantumState
from . import workloads as W
from ._lib import call

CIRCUITS = {
    "ghz": W.ghz,
    "qft": lambda n: W.qft_on_product(n, seed=2),
    "hadamard": None,
    "ising": lambda n: W.tfim_trotter(n),
}


def get_circuit(name, n):
    from .circuit import QuantumCircuit

    if name == "hadamard":
        c = QuantumCircuit(n)
        for q in range(n):
            c.add_H_gate(q)
        return c
    if name not in CIRCUITS:
        raise ValueError(f"unknown circuit {name!r}; choose from {sorted(CIRCUITS)}")
    return CIRCUITS[name](n)


class RunResult:
    def __init__(self):
        self.times = {}
        self.sv_MiB = 0.0
        self.fidelity = None


def _write(result, outfile, append):
    print(f"Simulation time: {result.times['overall']:.3f} seconds")
    print(f"State vector size: {result.sv_MiB:.3f} MiB")
    if result.fidelity is not None:
        print(f"Fidelity: {result.fidelity:.6f}")
    if outfile is not None:
        with open(outfile, "a") as f:
            json.dump(vars(result), f)
            f.write("\n")
        print(outfile)
    else:
        print(vars(result))


def run_sim(simulator, circuit_name, nqubits, outdir=None, append=False, block_size=None, inner_size=None,
            on_gpu_fidelity=True):
    if simulator in ("qompsim", "fidelity"):
        inner_size = 4 if inner_size is None else inner_size
        if block_size is None:
            block_size = min(nqubits - 1 - inner_size - 1, 26)  # run_sim.py:82-86 with one rank
    circuit = get_circuit(circuit_name, nqubits)
    QuantumCircuitOptimizer(sv_block_size=block_size or 0).optimize(circuit, 2)
    results = {}
    qs = None
    if simulator in ("qulacs", "fidelity"):
        out = f"{outdir}/qulacs_{circuit_name}_{nqubits}qubits_1ranks.json" if outdir else None
        if out and os.path.exists(out) and not append:
            print(f"{out} already exists. Use --append to append results.")
            return None
        r = RunResult()
        t0 = time.time()
        qs = QuantumState(nqubits)
        t1 = time.time()
        circuit.update_quantum_state(qs)
        call("qsim_synchronize")
        t2 = time.time()
        r.times = {"init": t1 - t0, "overall": t2 - t1, "fused_swap": qs.get_fused_swap_time()}
        r.sv_MiB = 2 ** (nqubits + 4) / 2 ** 20
        _write(r, out, append)
        results["qulacs"] = r
    if simulator in ("qompsim", "fidelity"):
        assert block_size + inner_size < nqubits, f"{block_size=} and {inner_size=} are too large"
        out = (f"{outdir}/qompsim_{circuit_name}_{nqubits}qubits_1ranks_bs{block_size}_inner{inner_size}.json"
               if outdir else None)
        if out and os.path.exists(out) and not append:
            print(f"{out} already exists. Use --append to append results.")
            return None
        r = RunResult()
        t0 = time.time()
        cqs = CompressedQuantumState(nqubits, block_size)
        cqs.partition_and_register_circuit(circuit, inner_size, verbose=1)
        t1 = time.time()
        cqs.apply_circuit()
        call("qsim_synchronize")
        t2 = time.time()
        r.times = {"init": t1 - t0, "overall": t2 - t1, **cqs.get_total_times()}
        r.sv_MiB = cqs.get_max_SV_bytes() / 2 ** 20
        if simulator == "fidelity":
            if on_gpu_fidelity:
                r.fidelity = float(cqs.fidelity_with(qs))
            else:
                r.fidelity = float(abs(np.vdot(qs.get_vector(), cqs.get_vector())))
        _write(r, out, append)
        results["qompsim"] = r
    return results


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("-s", "--simulator", choices=["qulacs", "qompsim", "fidelity"], required=True)
    ap.add_argument("-c", "--circuit_name", choices=sorted(CIRCUITS), required=True)
    ap.add_argument("-n", "--nqubits", type=int, required=True)
    ap.add_argument("-o", "--outdir", default=None)
    ap.add_argument("-a", "--append", action="store_true")
    ap.add_argument("-b", "--block_size", type=int, default=None)
    ap.add_argument("-i", "--inner_size", type=int, default=None)
    a = ap.parse_args(argv)
    run_sim(a.simulator, a.circuit_name, a.nqubits, a.outdir, a.append, a.block_size, a.inner_size)


if __name__ == "__main__":
    main(sys.argv[1:])
