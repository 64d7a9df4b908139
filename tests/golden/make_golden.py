"""Writes the golden fixtures of tests/golden/ (data only, numpy + closed forms).

ghz20_b12_i4.json — GHZ-20 (H(0), CNOT(i,i+1) i=0..18) after
  QuantumCircuitOptimizer.optimize(c, 2), block_size 12, inner_size 4:
  derived by reading src/cppsim/circuit_optimizer.cpp:113-168 and
  src/cppsim/compressed_state.hpp:509-671, 840-1029 — 19 gates
  (Dense{0,1} + CNOT(1..18)), stages [15, 3, 1], outer qubits per stage, and
  the final non-null block set {0, 15, 127, 255} (blocks 15 and 127 hold
  zeros but stay non-null: a block is only zero-checked while null).
qft_product_n10.npz — a seeded RY*RZ product state and its unitary DFT
  (numpy.fft.ifft with norm="ortho"), the closed form of QFT-with-swaps in
  the index = sum_q 2^q b_q convention.
"""
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def ghz20():
    return {
        "n": 20, "block_size": 12, "inner_size": 4,
        "optimized_gate_count": 19,
        "optimized_first_gate_targets": [0, 1],
        "stage_sizes": [15, 3, 1],
        "stage_outer_qubits": [[16, 17, 18, 19], [12, 13, 14, 19], [12, 13, 14, 15, 16, 17]],
        "stage_num_groups": [16, 16, 64],
        "final_nonnull_blocks": [0, 15, 127, 255],
        "final_amplitudes": {"0": [0.7071067811865476, 0.0], "1048575": [0.7071067811865476, 0.0]},
    }


def product_state(n, seed):
    rng = np.random.default_rng(seed)
    th = rng.uniform(0, 2 * np.pi, n)
    ph = rng.uniform(0, 2 * np.pi, n)
    # RZ(phi) RY(theta)|0> with R_P(a) = exp(+i a P/2)
    v = np.ones(1, dtype=np.complex128)
    for q in range(n):
        ry = np.array([np.cos(th[q] / 2), -np.sin(th[q] / 2)])
        rz = np.array([np.exp(1j * ph[q] / 2), np.exp(-1j * ph[q] / 2)])
        v = np.kron(rz * ry, v)
    return th, ph, v


if __name__ == "__main__":
    with open(os.path.join(HERE, "ghz20_b12_i4.json"), "w") as f:
        json.dump(ghz20(), f, indent=1)
    th, ph, v = product_state(10, 2)
    np.savez(os.path.join(HERE, "qft_product_n10.npz"), theta=th, phi=ph, input=v,
             output=np.fft.ifft(v, norm="ortho"))
    print("written")
