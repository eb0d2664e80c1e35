"""PyTorch autograd wrapper of the drag adjoint (cfdmx_adjoint_drag), for checks of the
hand-written adjoint against finite differences and torch.autograd.gradcheck.

The device work happens in the HIP library; the tensors here are small CPU float64
parameter vectors:
    inflow -- the inflow profile u_in(j) of face 0 at the interior rows (length nyc)
    mask   -- (cx, cy, R) of the smoothed immersed-boundary mask (width eps fixed)
J = DragObjective.apply(inflow, mask, solver, nsteps, checkpoint_every) is the mean
drag coefficient over nsteps steps from the solver's current state (restored after
every call), with U^0 the start boundary conditions of that state.
"""
import numpy as np
import torch


def _set_params(solver, inflow, mask, eps):
    prof = np.ones(solver.ny)
    prof[2:-2] = inflow.detach().cpu().double().numpy()
    solver.set_inflow(prof)
    cx, cy, r = (float(v) for v in mask.detach().cpu().double())
    solver.set_eta_smoothed(cx, cy, r, eps)


class DragObjective(torch.autograd.Function):
    @staticmethod
    def forward(ctx, inflow, mask, solver, nsteps, checkpoint_every, eps):
        _set_params(solver, inflow, mask, eps)
        if not (inflow.requires_grad or mask.requires_grad):
            return torch.tensor(solver.drag_objective(nsteps), dtype=torch.float64)
        J, gi, gm, _ = solver.adjoint_drag(nsteps, checkpoint_every)
        ctx.save_for_backward(torch.from_numpy(gi[2:-2].copy()), torch.from_numpy(gm.copy()))
        return torch.tensor(J, dtype=torch.float64)

    @staticmethod
    def backward(ctx, grad_out):
        gi, gm = ctx.saved_tensors
        return grad_out * gi, grad_out * gm, None, None, None, None


def drag_objective(inflow, mask, solver, nsteps, checkpoint_every=10, eps=None):
    """J(inflow, mask); eps defaults to 1.5 cells of the x spacing."""
    if eps is None:
        eps = 1.5 * solver.cfg.lx / solver.cfg.nx
    return DragObjective.apply(inflow, mask, solver, nsteps, checkpoint_every, eps)
