"""qulacs.circuit (cppsim_wrapper.cpp:1693-1715): QuantumCircuitOptimizer."""
from qompsim.circuit import QuantumCircuitOptimizer  # noqa: F401
