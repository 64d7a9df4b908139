"""qulacs.observable (cppsim_wrapper.cpp:325-): the Observable class."""
from qompsim.observable import Observable  # noqa: F401
