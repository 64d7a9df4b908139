"""The allocation-free Pauli expectation (qsim_sv_expectation_pauli_noalloc)
against the allocating entry point (qsim_sv_expectation_pauli), called through
the C ABI: bit-identical results on two calls in a row, the second on a
smaller state with fewer terms (the shared scratch keeps the first call's
contents), the term limit enforced, and the oracle's expectation as the
numerical pin."""
import ctypes

import numpy as np
import pytest

import pyoracle as O
from gpu_util import rand_state
from qompsim import QuantumState, _lib
from qompsim.observable import GeneralQuantumOperator, Observable

pytestmark = pytest.mark.gpu


def _random_observable(n, count, rng):
    obs = GeneralQuantumOperator(n)  # complex coefficients
    for _ in range(count):
        k = int(rng.integers(1, min(4, n) + 1))
        qs = rng.choice(n, size=k, replace=False)
        ps = rng.choice(["X", "Y", "Z"], size=k)
        obs.add_operator(complex(rng.normal(), rng.normal()), " ".join(f"{p} {q}" for p, q in zip(ps, qs)))
    return obs


def _call(name, state, obs):
    xm, zm, ny, cr, ci = obs._arrays()
    re, im = ctypes.c_double(), ctypes.c_double()
    _lib.call(name, state._h, len(obs.terms), _lib.ptr(xm), _lib.ptr(zm), _lib.ptr(ny), _lib.ptr(cr), _lib.ptr(ci),
              ctypes.byref(re), ctypes.byref(im))
    return re.value, im.value


@pytest.mark.parametrize("n1,t1,n2,t2", [(18, 300, 12, 40), (20, 1024, 9, 1)])
def test_noalloc_expectation_matches_allocating_entry_point(n1, t1, n2, t2):
    rng = np.random.default_rng(n1 * 1000 + t1)
    for n, t in ((n1, t1), (n2, t2)):
        psi = rand_state(n, rng)
        s = QuantumState(n)
        s.load(psi)
        obs = _random_observable(n, t, rng)
        got = _call("qsim_sv_expectation_pauli_noalloc", s, obs)
        ref = _call("qsim_sv_expectation_pauli", s, obs)
        assert got == ref, (n, t, got, ref)  # same tables, same reduction order
        want = O.expectation(psi, obs)
        assert abs(complex(*got) - want) < 1e-10 * max(1.0, abs(want))


def test_noalloc_expectation_term_limit():
    s = QuantumState(8)
    obs = Observable(8)
    for i in range(1025):
        obs.add_operator(1.0, f"Z {i % 8}")
    with pytest.raises(ValueError, match="at most 1024 terms"):
        _call("qsim_sv_expectation_pauli_noalloc", s, obs)
