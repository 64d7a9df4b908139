"""qompsim — MI355X-native compressed state-vector stage loop with the Qulacs /
QompSim Python surface (``qulacs`` package of the reference,
pysrc/qulacs/__init__.py and python/cppsim_wrapper.cpp).

All amplitude work runs in libqsim_hip.so (HIP kernels for gfx950).
"""
from . import gate
from ._lib import LIB_PATH
from .circuit import ParametricQuantumCircuit, QuantumCircuit, QuantumCircuitOptimizer
from .observable import GeneralQuantumOperator, Observable, PauliTerm
from .state import CompressedQuantumState, QuantumState, inner_product

__all__ = [
    "gate", "QuantumCircuit", "QuantumCircuitOptimizer", "ParametricQuantumCircuit", "QuantumState",
    "CompressedQuantumState", "Observable", "GeneralQuantumOperator", "PauliTerm", "inner_product",
    "check_build_for_mpi", "LIB_PATH",
]


def check_build_for_mpi():
    """One GPU holds every block group: the build has no MPI ranks."""
    return False
