"""torch.autograd front end of the radius-trajectory adjoint of libbemhip.

`radius_loss(sim, x0, phi0, target, dt, every)` returns the scalar
L = sum_n (R_n - target_n)^2 of `len(target)` fixed RK4 steps of `sim` started
from (x0, phi0) as a torch tensor whose backward yields dL/dx0 and dL/dphi0
from the library's hand-written adjoint (bemhip_sim_adjoint_radius).  The
tensors are FP64 on the host; the device work happens inside the library.
V_0, the forcing and the connectivity of `sim` are held fixed.
"""

import numpy as np
import torch


class RadiusTrajectoryLoss(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x0, phi0, sim, tris, target, dt, every):
        sim.set_mesh(x0.detach().cpu().numpy(), tris)
        sim.set_phi(phi0.detach().cpu().numpy())
        tgt = np.asarray(target, dtype=np.float64)
        loss, R, gx, gp = sim.adjoint_radius(len(tgt), dt, tgt, every)
        ctx.save_for_backward(torch.from_numpy(gx), torch.from_numpy(gp))
        ctx.radius = R
        return torch.tensor(loss, dtype=torch.float64)

    @staticmethod
    def backward(ctx, g):
        gx, gp = ctx.saved_tensors
        return g * gx, g * gp, None, None, None, None, None


def radius_loss(sim, tris, x0, phi0, target, dt, every=10):
    return RadiusTrajectoryLoss.apply(x0, phi0, sim, tris, target, dt, every)
