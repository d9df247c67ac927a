"""torch.autograd.Function over the hand-written adjoint of the coupled step.

    phi_T = CoupledSteps.apply(phi0, sigma, mobility, config, nsteps, checkpoint_every, precision)

phi0 (interior shape, index [i, j(, k)], passed to the library in that
C order), sigma and mobility are tensors; the forward runs nsteps coupled
steps on the GPU from the case's initial state built on phi0
(clip_adjoint_forward) and the backward runs the reverse sweep
(clip_adjoint_backward): gradients with respect to phi0, sigma and
mobility.  Periodic, wall, velocity-inlet, free-convect, Neumann and
do-nothing sides; "f64" or "f32".  normal_floor
(default 0.1) bounds the derivative of the interface normal on flat
plateaus (clip_adjoint_set_normal_floor); 0 gives the exact derivative.
"""
import torch

from . import Solver


class CoupledSteps(torch.autograd.Function):
    @staticmethod
    def forward(ctx, phi0, sigma, mobility, config, nsteps, checkpoint_every=50, precision="f64", normal_floor=0.1):
        cfg = config.with_params(sigma=float(sigma), mobility=float(mobility))
        solver = Solver(cfg, precision)
        out = solver.adjoint_forward(phi0.detach().cpu().numpy(), int(nsteps), int(checkpoint_every))
        ctx.solver = solver
        ctx.normal_floor = float(normal_floor)
        ctx.meta = (phi0.dtype, phi0.device, sigma.dtype, sigma.device, mobility.dtype, mobility.device)
        return torch.as_tensor(out, dtype=phi0.dtype, device=phi0.device)

    @staticmethod
    def backward(ctx, grad_out):
        pd, pdev, sd, sdev, md, mdev = ctx.meta
        g0, gs, gm = ctx.solver.adjoint_backward(grad_out.detach().cpu().numpy(), ctx.normal_floor)
        ctx.solver.close()
        return (torch.as_tensor(g0, dtype=pd, device=pdev), torch.tensor(gs, dtype=sd, device=sdev),
                torch.tensor(gm, dtype=md, device=mdev), None, None, None, None, None)


def coupled_steps(phi0, sigma, mobility, config, nsteps, checkpoint_every=50, precision="f64", normal_floor=0.1):
    sigma = sigma if torch.is_tensor(sigma) else torch.tensor(float(sigma), dtype=torch.float64)
    mobility = mobility if torch.is_tensor(mobility) else torch.tensor(float(mobility), dtype=torch.float64)
    return CoupledSteps.apply(phi0, sigma, mobility, config, nsteps, checkpoint_every, precision, normal_floor)


__all__ = ["CoupledSteps", "coupled_steps"]
