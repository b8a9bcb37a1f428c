"""torch.autograd binding of the hand-written discrete adjoint.

`BrinkmanObjective.apply(alpha, inlet_scale, lattice_factory, config)` returns the
time-averaged objective J of a fresh lattice built by `lattice_factory()` with
porosity `alpha` (FP64 tensor [nx, ny, nz]) and the first velocity plane scaled by
`inlet_scale` (FP64 scalar tensor); backward hands out the adjoint gradients that
the native library computed in the same forward call (no tape: the adjoint
recomputes the forward from checkpoints). It exists so that the gradient can be
checked with torch tooling (`torch.autograd.gradcheck` on a few entries, or
chained into a larger torch computation); the lattice itself runs only through
the native library.
"""
import dataclasses

import numpy as np
import torch


@dataclasses.dataclass
class AdjointConfig:
    nsteps: int
    checkpoint_every: int
    inlet_x: int
    outlet_x: int
    weight_pressure_drop: float = 1.0
    weight_drag: float = 1.0
    weight_flux: float = 0.0
    first_step: int = 0


class BrinkmanObjective(torch.autograd.Function):
    """J(alpha, s) of the BGK + Brinkman channel; gradients from the native adjoint."""

    @staticmethod
    def forward(ctx, alpha, inlet_scale, lattice_factory, config):
        a = alpha.detach().to("cpu", torch.float64).contiguous().numpy()
        s = float(inlet_scale.detach().cpu())
        lat, init_populations, inlet_plane, inlet_values = lattice_factory()
        try:
            lat.setBoundaryVelocity(inlet_plane, np.asarray(inlet_values) * s)
            lat.setPorosity(a)
            lat.setPopulations(init_populations)
            J, ga, gs, _ = lat.adjointGradient(config.nsteps, config.checkpoint_every, config.inlet_x,
                                               config.outlet_x, config.weight_pressure_drop,
                                               config.weight_drag, config.weight_flux,
                                               config.first_step)
        finally:
            lat.close()
        # dJ/ds_total = dJ/d(scale of the plane) / s, because the plane values were
        # multiplied by s before the adjoint measured the scale derivative at 1.
        g_s = gs[inlet_plane] / s if s != 0.0 else float("nan")
        ctx.save_for_backward(torch.from_numpy(ga).to(alpha.device),
                              torch.tensor(g_s, dtype=torch.float64, device=inlet_scale.device))
        return torch.tensor(J, dtype=torch.float64, device=alpha.device)

    @staticmethod
    def backward(ctx, grad_output):
        ga, gs = ctx.saved_tensors
        return grad_output * ga, grad_output * gs, None, None
