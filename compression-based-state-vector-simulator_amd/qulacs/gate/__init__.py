"""qulacs.gate (cppsim_wrapper.cpp:913-1220): the gate factories."""
from qompsim.gate import *  # noqa: F401,F403
