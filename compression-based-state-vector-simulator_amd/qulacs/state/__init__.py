"""qulacs.state (cppsim_wrapper.cpp:718-): inner_product."""
from qompsim.state import inner_product  # noqa: F401
