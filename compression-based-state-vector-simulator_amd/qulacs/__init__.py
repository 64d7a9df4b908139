"""`qulacs` import surface over qompsim, so that code written against the
reference's Python package (pysrc/qulacs/__init__.py re-exports the pybind
module qulacs_core, python/cppsim_wrapper.cpp) runs unchanged:

    import qulacs
    from qulacs import QuantumState, CompressedQuantumState
    from qulacs.circuit import QuantumCircuitOptimizer

as script/run_sim.py:4-6 does.  Every class is qompsim's (libqsim_hip.so on
the GPU); this package only maps the reference's module layout onto it:
qulacs.circuit, qulacs.gate, qulacs.state and qulacs.observable are the
reference's submodules (cppsim_wrapper.cpp:1693, 913, 718, 325).
"""
from qompsim import (  # noqa: F401
    CompressedQuantumState,
    GeneralQuantumOperator,
    Observable,
    ParametricQuantumCircuit,
    PauliTerm,
    QuantumCircuit,
    QuantumCircuitOptimizer,
    QuantumState,
    check_build_for_mpi,
    gate,
    inner_product,
)

from . import circuit, observable, state  # noqa: F401,E402

__version__ = "0.0.0+qompsim.mi355x"
__version_tuple__ = (0, 0, 0, "qompsim", "mi355x")
