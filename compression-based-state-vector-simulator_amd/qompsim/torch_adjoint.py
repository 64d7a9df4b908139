"""torch.autograd.Function over the GPU adjoint: <H>(theta) for a
ParametricQuantumCircuit, differentiable in theta.  The forward call runs
qsim_sv_backprop once (energy and the full gradient together); backward scales
the stored gradient by the incoming grad_output.  Parameters live in a float64
CPU or GPU tensor of length get_parameter_count().
"""
import numpy as np
import torch

from .adjoint import adjoint_gradient


class AdjointExpectation(torch.autograd.Function):
    @staticmethod
    def forward(ctx, theta, circuit, obs):
        th = theta.detach().to("cpu", torch.float64).numpy()
        if th.shape != (circuit.get_parameter_count(),):
            raise ValueError("theta must hold one value per circuit parameter")
        circuit.set_parameters(th)
        energy, grads = adjoint_gradient(circuit, obs)
        ctx.save_for_backward(torch.as_tensor(np.asarray(grads), dtype=theta.dtype, device=theta.device))
        return torch.as_tensor(energy, dtype=theta.dtype, device=theta.device)

    @staticmethod
    def backward(ctx, grad_output):
        (grads,) = ctx.saved_tensors
        return grad_output * grads, None, None


def expectation(theta, circuit, obs):
    """<H> as a differentiable function of theta."""
    return AdjointExpectation.apply(theta, circuit, obs)
