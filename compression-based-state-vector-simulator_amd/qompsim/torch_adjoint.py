"""torch.autograd.Function over the GPU adjoint: <H>(theta) for a
ParametricQuantumCircuit, differentiable in theta.  Parameters live in a
float64 CPU or GPU tensor of length get_parameter_count().

Without checkpoints the forward call runs qsim_sv_backprop once (energy and
the full gradient together) and backward scales the stored gradient.  With
`segment_gates` the forward call runs forward_with_checkpoints and returns
<H>; backward runs the checkpointed reverse sweep (qsim_adj_backward).
"""
import numpy as np
import torch

from .adjoint import CheckpointedAdjoint, adjoint_gradient


def _set_theta(theta, circuit):
    th = theta.detach().to("cpu", torch.float64).numpy()
    if th.shape != (circuit.get_parameter_count(),):
        raise ValueError("theta must hold one value per circuit parameter")
    circuit.set_parameters(th)


class AdjointExpectation(torch.autograd.Function):
    @staticmethod
    def forward(ctx, theta, circuit, obs, register=None):
        _set_theta(theta, circuit)
        energy, grads = adjoint_gradient(circuit, obs, register=register)
        ctx.save_for_backward(torch.as_tensor(np.asarray(grads), dtype=theta.dtype, device=theta.device))
        return torch.as_tensor(energy, dtype=theta.dtype, device=theta.device)

    @staticmethod
    def backward(ctx, grad_output):
        (grads,) = ctx.saved_tensors
        return grad_output * grads, None, None, None


class CheckpointedExpectation(torch.autograd.Function):
    @staticmethod
    def forward(ctx, theta, circuit, obs, segment_gates, block_size, codec):
        _set_theta(theta, circuit)
        adj = CheckpointedAdjoint(circuit, segment_gates, block_size, codec)
        psi = adj.forward_with_checkpoints()
        energy = obs.get_expectation_value(psi).real
        ctx.adj, ctx.obs = adj, obs
        ctx.meta = (theta.dtype, theta.device)
        return torch.as_tensor(energy, dtype=theta.dtype, device=theta.device)

    @staticmethod
    def backward(ctx, grad_output):
        _, grads = ctx.adj.adjoint(ctx.obs)
        dtype, device = ctx.meta
        g = torch.as_tensor(np.asarray(grads), dtype=dtype, device=device)
        ctx.adj = None
        return grad_output * g, None, None, None, None, None


def expectation(theta, circuit, obs, segment_gates=None, block_size=20, codec=None, register=None):
    """<H> as a differentiable function of theta; with `segment_gates` the
    backward pass uses compressed forward checkpoints every that many gates.
    Without checkpoints, `register` (a QuantumState of the circuit's qubit
    count) is the working state of every call and holds U|0...0> afterwards;
    the library keeps its lambda and psi buffers with it, so repeated calls
    allocate nothing.  Without it each call allocates a fresh state."""
    if segment_gates is None:
        return AdjointExpectation.apply(theta, circuit, obs, register)
    return CheckpointedExpectation.apply(theta, circuit, obs, int(segment_gates), int(block_size), codec)
