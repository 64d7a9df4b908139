"""Batched energies and adjoint gradients (qsim_batch_*) against the
oracle's restated backprop, eager and graph-replayed, before and after the
autotuner."""
import numpy as np
import pytest

import pyoracle as O
from qompsim import QuantumState
from qompsim import workloads as W
from qompsim.batch import BatchedAdjoint

pytestmark = pytest.mark.gpu


def _instances(rng, sizes=(8, 10, 12, 11), p=2):
    out = []
    for n in sizes:
        edges = W.random_3regular(n, rng) if n % 2 == 0 else [(i, (i + 1) % n) for i in range(n)]
        out.append((n, edges, rng.uniform(0, np.pi, p), rng.uniform(0, np.pi, p)))
    return out


def _reference(insts):
    ref = []
    for n, edges, gm, bt in insts:
        c = W.qaoa_maxcut_shared(n, edges, gm, bt)
        e, g = O.backprop(c, W.maxcut_observable(n, edges))
        ref.append((e, g[: 2 * len(gm)]))
    return ref


def _build(insts):
    b = BatchedAdjoint()
    for n, edges, gm, bt in insts:
        b.add(W.qaoa_maxcut_shared(n, edges, gm, bt), W.maxcut_observable(n, edges), n_params=2 * len(gm))
    return b


def _check(e, g, ref):
    for i, (e0, g0) in enumerate(ref):
        assert abs(e[i] - e0) < 1e-10
        assert np.max(np.abs(g[i, : len(g0)] - g0)) < 1e-10


def test_batch_eager_and_graph_replays_match_oracle(rng):
    insts = _instances(rng)
    ref = _reference(insts)
    b = _build(insts)
    e_eager, g_eager = b.run(use_graph=False)
    _check(e_eager, g_eager, ref)
    for _ in range(3):
        # the program's own eager work between replays
        s = QuantumState(10)
        s.get_vector()
        e, g = b.run(use_graph=True)
        assert np.array_equal(e, e_eager) and np.array_equal(g, g_eager)


def test_batch_autotuner_keeps_results(rng):
    insts = _instances(rng, sizes=(10, 12, 14), p=2)
    ref = _reference(insts)
    b = _build(insts)
    b.tune(reps=1)
    for n in (10, 12, 14):
        cfg = b.config(n)
        assert cfg["forward_engine"] in ("lds", "wave") and cfg["adjoint_tile_bits"] >= 4
    e, g = b.run(use_graph=True)
    _check(e, g, ref)
    e2, g2 = b.run(use_graph=True)
    assert np.array_equal(e, e2) and np.array_equal(g, g2)


def test_batch_persistent_kernel_matches_launches_and_oracle(rng, monkeypatch, capfd):
    """Instances of 10..16 qubits run in the persistent batch kernel (one
    work-group per instance, taken from a counter), an 18-qubit instance
    through its own launches beside it: energies and gradients against the
    oracle's backprop (1e-10) and the per-instance launches
    (QOMPSIM_BATCH_PERSISTENT=0, 1e-12); graph replays equal the eager run."""
    monkeypatch.setenv("QOMPSIM_PLAN_LOG", "1")
    insts = _instances(rng, sizes=(10, 12, 16, 11, 14, 13, 18, 12, 10), p=3)
    ref = _reference(insts)
    b = _build(insts)
    e, g = b.run(use_graph=False)
    assert "[batch] persistent kernel instances: 8 of 9" in capfd.readouterr().err
    _check(e, g, ref)
    for _ in range(3):
        s = QuantumState(10)
        s.get_vector()
        e2, g2 = b.run(use_graph=True)
        assert np.array_equal(e, e2) and np.array_equal(g, g2)
    monkeypatch.setenv("QOMPSIM_BATCH_PERSISTENT", "0")
    b0 = _build(insts)
    e0, g0 = b0.run(use_graph=False)
    assert "persistent kernel" not in capfd.readouterr().err
    assert np.max(np.abs(e - e0)) < 1e-12 and np.max(np.abs(g - g0)) < 1e-12


def test_batch_persistent_many_instances(rng):
    """96 instances of 10..13 qubits, more than one per work-group of the
    persistent grid on a small device share: every result against the
    oracle."""
    sizes = [10 + (i % 4) for i in range(96)]
    insts = _instances(rng, sizes=sizes, p=2)
    ref = _reference(insts)
    b = _build(insts)
    e, g = b.run(use_graph=True)
    _check(e, g, ref)
