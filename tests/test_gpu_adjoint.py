"""GPU adjoint gradients against the oracle's restated backprop and central
finite differences (h=1e-5) on the uncompressed FP64 path."""
import numpy as np
import pytest

import pyoracle as O
from qompsim import QuantumState
from qompsim import workloads as W
from qompsim.adjoint import adjoint_gradient

pytestmark = pytest.mark.gpu


def test_adjoint_matches_oracle_and_fd():
    n = 8
    c = W.hea(n, layers=2, seed=4)
    obs = W.tfim_observable(n)
    e, g = adjoint_gradient(c, obs)
    e0, g0 = O.backprop(c, obs)
    assert abs(e - e0) < 1e-12
    assert np.max(np.abs(g - g0)) < 1e-10
    th = c.get_parameters()
    h = 1e-5
    for i in range(0, len(th), 5):
        cp = c.copy()
        cp.set_parameter(i, th[i] + h)
        s = QuantumState(n)
        cp.update_quantum_state(s)
        ep = obs.get_expectation_value(s)
        cp.set_parameter(i, th[i] - h)
        s = QuantumState(n)
        cp.update_quantum_state(s)
        em = obs.get_expectation_value(s)
        assert abs((ep - em) / (2 * h) - g[i]) < 1e-6
