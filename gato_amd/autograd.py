"""Differentiable batched solve: torch.autograd.Function over the HIP solve and its adjoint.

    XU = solve(solver, xs, ref, Qd, Rd, xu0)

runs the batched SQP solve from the initial trajectory xu0 with initial states xs [B, nx], end-effector
references ref [B, 6N] and per-problem diagonal joint weights Qd [B, nx], Rd [B, nu] (the cost's
joint-space terms 1/2 Qd (x - 0)^2 + 1/2 Rd u^2 next to the end-effector tracking term), and returns the
solution XU [B, traj]. Backward gives dL/dxs, dL/dref, dL/dQd, dL/dRd through the adjoint of the SQP step
at the solution (gato_backward in the C-ABI). No reference counterpart: the reference has no backward.
"""
import torch


class _SQPSolve(torch.autograd.Function):
    @staticmethod
    def forward(ctx, solver, xs, ref, Qd, Rd, xu0):
        xs_c = xs.detach().contiguous().float()
        ref_c = ref.detach().contiguous().float()
        solver.set_joint_cost(Qd.detach().contiguous().float(), Rd.detach().contiguous().float())
        xu = xu0.detach().clone().contiguous().float()
        solver.solve_device(xu, xs_c, ref_c)
        ctx.solver = solver
        ctx.save_for_backward(xu, xs_c, ref_c)
        return xu.clone()

    @staticmethod
    def backward(ctx, grad_xu):
        xu, xs, ref = ctx.saved_tensors
        g = ctx.solver.backward_device(xu, xs, ref, grad_xu.contiguous().float())
        return None, g["xs"], g["ref"], g["Qd"], g["Rd"], None


def solve(solver, xs, ref, Qd, Rd, xu0):
    """Differentiable solve (see the module docstring)."""
    return _SQPSolve.apply(solver, xs, ref, Qd, Rd, xu0)
