"""
The HF energy as a differentiable function of the nuclear positions.

``HFEnergy.apply(positions, solver, guess, sink)`` runs the SCF in forward();
backward() returns dE/dR from the analytic gradient (one-electron derivative
kernel with D and W, derivative-ERI kernel with Gamma), so any torch code --
an integrator, an optimiser -- obtains forces from ``energy.backward()``.
The SCF result of forward() is handed back through ``sink`` (a dict) so that
callers can read the density, dipole and convergence data of the same run.
"""

from __future__ import annotations

import numpy as np
import torch

from aimd_amd.qc import engine as qengine
from aimd_amd.qc.gradients import scf_gradient


class HFEnergy(torch.autograd.Function):

    @staticmethod
    def forward(ctx, positions: torch.Tensor, solver, guess=None, sink: dict | None = None,
                gradient_screening: float = qengine.DEFAULT_SCHWARZ_GRAD):
        x = positions.detach().to("cpu", torch.float64).numpy()
        res = solver.run(x, guess=guess)
        ctx.result = res
        ctx.screening = float(gradient_screening)
        ctx.pos_device = positions.device
        if sink is not None:
            sink["result"] = res
        return torch.tensor(res.energy, dtype=torch.float64, device=positions.device)

    @staticmethod
    def backward(ctx, grad_energy: torch.Tensor):
        g = scf_gradient(ctx.result, schwarz_threshold=ctx.screening)
        grad = torch.as_tensor(g, dtype=torch.float64, device=ctx.pos_device) * grad_energy
        return grad, None, None, None, None


def energy_and_forces(solver, positions, guess=None, gradient_screening=qengine.DEFAULT_SCHWARZ_GRAD):
    """(energy, gradient (natm,3) numpy, SCFResult) through HFEnergy's forward and backward."""
    pos = torch.tensor(np.asarray(positions, dtype=float), dtype=torch.float64, requires_grad=True)
    sink: dict = {}
    E = HFEnergy.apply(pos, solver, guess, sink, gradient_screening)
    E.backward()
    return float(E.detach()), pos.grad.numpy().copy(), sink["result"]
