"""torch.autograd access to the free energy of a grain.

`free_energy(sim, delta)` returns F(Delta) as a 0-d float64 tensor that
torch can differentiate: its backward is the discrete adjoint of one
iteration on the GPU (`Simulation.free_energy_gradient`, C ABI
`sc_free_energy_gradient_ex`).  `delta` is a real float64 tensor of shape
(Nc, 2, N0, N0), the real and imaginary planes of every order-parameter
component (the C ABI's layout); for a complex tensor pass
`torch.view_as_real(delta).movedim(-1, 1)`.  `free_energy(sim, delta, a)`
also takes the induced vector potential, a (2, N0, N0) float64 tensor (x
and y components, in the units `set_vector_potential` takes), and
differentiates F(Delta, A) in both.  The boundary in-values of `sim` (and
its vector potential, when `a` is not given) are held fixed, and the call
sets `sim`'s order parameter (and vector potential) to the arguments.

The frontend of the reference runs the solver as a subprocess
(frontend/common/cpp.py:16-39); here the grain lives in-process.
"""
import numpy as np
import torch


class FreeEnergy(torch.autograd.Function):
    """F(Delta[, A]) of one iteration; backward by the discrete adjoint."""

    @staticmethod
    def forward(ctx, delta, a, sim):
        d = delta.detach().to("cpu", torch.float64).numpy()
        if d.shape != (sim.Nc, 2, sim.N0, sim.N0):
            raise ValueError("delta must have shape (Nc, 2, N0, N0) = %s"
                             % ((sim.Nc, 2, sim.N0, sim.N0),))
        for c in range(sim.Nc):
            sim.set_order_parameter(c, d[c, 0] + 1j * d[c, 1])
        if a is not None:
            av = a.detach().to("cpu", torch.float64).numpy()
            if av.shape != (2, sim.N0, sim.N0):
                raise ValueError("a must have shape (2, N0, N0) = %s" % ((2, sim.N0, sim.N0),))
            sim.set_vector_potential(av[0], av[1])
            f, g, ga = sim.free_energy_gradient(vector_potential=True)
            grad_a = torch.from_numpy(ga).to(delta.device)
        else:
            f, g = sim.free_energy_gradient()
            grad_a = torch.zeros(0, dtype=torch.float64, device=delta.device)
        grad = np.stack([g.real, g.imag], axis=1)
        ctx.has_a = a is not None
        ctx.save_for_backward(torch.from_numpy(grad).to(delta.device), grad_a)
        return torch.tensor(f, dtype=torch.float64, device=delta.device)

    @staticmethod
    def backward(ctx, grad_output):
        grad, grad_a = ctx.saved_tensors
        return grad_output * grad, (grad_output * grad_a if ctx.has_a else None), None


def free_energy(sim, delta, a=None):
    """F(Delta) or F(Delta, A) as a differentiable tensor (see the module
    docstring)."""
    return FreeEnergy.apply(delta, a, sim)
