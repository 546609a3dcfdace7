"""torch.autograd access to the free energy of a grain.

`free_energy(sim, delta)` returns F(Delta) as a 0-d float64 tensor that
torch can differentiate: its backward is the discrete adjoint of one
iteration on the GPU (`Simulation.free_energy_gradient`, C ABI
`sc_free_energy_gradient`).  `delta` is a real float64 tensor of shape
(Nc, 2, N0, N0), the real and imaginary planes of every order-parameter
component (the C ABI's layout); for a complex tensor pass
`torch.view_as_real(delta).movedim(-1, 1)`.  The vector potential and the
boundary in-values of `sim` are held fixed, and the call sets `sim`'s order
parameter to `delta`.

The frontend of the reference runs the solver as a subprocess
(frontend/common/cpp.py:16-39); here the grain lives in-process.
"""
import numpy as np
import torch


class FreeEnergy(torch.autograd.Function):
    """F(Delta) of one iteration; backward by the discrete adjoint."""

    @staticmethod
    def forward(ctx, delta, sim):
        d = delta.detach().to("cpu", torch.float64).numpy()
        if d.shape != (sim.Nc, 2, sim.N0, sim.N0):
            raise ValueError("delta must have shape (Nc, 2, N0, N0) = %s"
                             % ((sim.Nc, 2, sim.N0, sim.N0),))
        for c in range(sim.Nc):
            sim.set_order_parameter(c, d[c, 0] + 1j * d[c, 1])
        f, g = sim.free_energy_gradient()
        grad = np.stack([g.real, g.imag], axis=1)
        ctx.save_for_backward(torch.from_numpy(grad).to(delta.device))
        return torch.tensor(f, dtype=torch.float64, device=delta.device)

    @staticmethod
    def backward(ctx, grad_output):
        (grad,) = ctx.saved_tensors
        return grad_output * grad, None


def free_energy(sim, delta):
    """F(Delta) as a differentiable tensor (see the module docstring)."""
    return FreeEnergy.apply(delta, sim)
